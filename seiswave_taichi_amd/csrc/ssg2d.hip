// 2-D P-SV sweeps: standard staggered grid in its two staggerings and the
// rotated staggered grid, with CPML memory variables.
//
// Reference: src/Elastic2D_CPML.py:127-200 (update_SSG_VS), :202-274
// (update_SSG_SV), :276-353 (update_RSG).  Device layout [nz][nx], x fastest.
// The 2-D grids are small, so memory variables are full-grid arrays and the
// CPML branch follows the reference's coordinate test exactly (per-axis
// in-PML flags computed on the host with the same f32 comparison).
#include "sw_internal.h"

namespace sw {

namespace {

struct Rot {
    // Rotated stencils (src/Differential2D.py:32-56) in the [nz][nx] layout:
    // (i+a, j+b) is p + a + b*SZ.
    template <int N>
    static __device__ __forceinline__ float rx_fm(const float *f, long p, long SZ, const float *c)
    {
        float v = 0.f;
#pragma unroll
        for (int id = 0; id < N; ++id) v += c[id] * (f[p + (id + 1) * (1 + SZ)] - f[p - id * (1 + SZ)]);
        return v;
    }
    template <int N>
    static __device__ __forceinline__ float rz_fm(const float *f, long p, long SZ, const float *c)
    {
        float v = 0.f;
#pragma unroll
        for (int id = 0; id < N; ++id) v += c[id] * (f[p - id + (id + 1) * SZ] - f[p + (id + 1) - id * SZ]);
        return v;
    }
    template <int N>
    static __device__ __forceinline__ float rx_bm(const float *f, long p, long SZ, const float *c)
    {
        float v = 0.f;
#pragma unroll
        for (int id = 0; id < N; ++id) v += c[id] * (f[p + id * (1 + SZ)] - f[p - (id + 1) * (1 + SZ)]);
        return v;
    }
    template <int N>
    static __device__ __forceinline__ float rz_bm(const float *f, long p, long SZ, const float *c)
    {
        float v = 0.f;
#pragma unroll
        for (int id = 0; id < N; ++id) v += c[id] * (f[p - (id + 1) + id * SZ] - f[p + id - (id + 1) * SZ]);
        return v;
    }
};

// With SAVE (VS scheme), each sweep also writes the CPML-corrected derivative
// sums its updates multiply, which are all the model gradient needs of the
// forward state: planes 0-1 of sv (velocity sweep: vx cells dsxx/dx +
// dsxz/dz, vz cells dszz/dz + dsxz/dx) and 2-4 (stress sweep: sxx cells
// dvx/dx and dvz/dz, sxz cells dvx/dz + dvz/dx), at the cells each sweep
// updates; shot b's planes start b*ss floats after sv.
template <int N, int SCHEME, bool SAVE>
__global__ void __launch_bounds__(256) k_velocity2d(sw_ssg2d g, long bs, float *sv, long ss, Fold2D fo)
{
    const Blk bk = xcd_block3();
    shift_state(g, (long)bk.z * bs);
    if (SAVE) sv += (long)bk.z * ss;
    const int i = bk.x * blockDim.x + threadIdx.x + N;
    const int j = bk.y * blockDim.y + threadIdx.y + N;
    const int nx = g.nx, nz = g.nz;
    if (i >= nx - N || j >= nz - N) return;
    const long SZ = nx, NC = (long)nx * nz;
    const long p = j * SZ + i;
    const float dx = g.dx, dz = g.dz, dt = g.dt;
    const float *c = g.c;
    float *m = g.mem;
    const sw_axis &X = g.px, &Z = g.pz;

    if (SCHEME == SW_2D_VS) {
        // vx at (i, j): src/Elastic2D_CPML.py:144-154
        if (i >= N + 1 && j >= N + 1) {
            float dsxxdx = d_bm<N>(g.sxx, p, 1, c) / dx;
            float dsxzdz = d_bm<N>(g.sxz, p, SZ, c) / dz;
            if (X.in[i] | Z.in[j]) {
                dsxxdx = cpml_apply(dsxxdx, m + 0 * NC + p, X.a[i], X.b[i], X.k[i]);
                dsxzdz = cpml_apply(dsxzdz, m + 3 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
            }
            const float sum = dsxxdx + dsxzdz;
            g.vx[p] += sum * dt / g.rho[p];
            if (SAVE) sv[p] = sum;
        }
        // vz at (i+1/2, j+1/2): :156-167
        if (i < nx - N - 1 && j < nz - N - 1) {
            const float rho_avg = 0.25f * (g.rho[p] + g.rho[p + SZ] + g.rho[p + 1 + SZ] + g.rho[p + 1]);
            float dsxzdx = d_fm<N>(g.sxz, p, 1, c) / dx;
            float dszzdz = d_fm<N>(g.szz, p, SZ, c) / dz;
            if (X.in_h[i] | Z.in_h[j]) {
                dsxzdx = cpml_apply(dsxzdx, m + 2 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
                dszzdz = cpml_apply(dszzdz, m + 5 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
            }
            const float sum = dszzdz + dsxzdx;
            g.vz[p] += sum * dt / rho_avg;
            if (SAVE) sv[NC + p] = sum;
        }
    } else if (SCHEME == SW_2D_SV) {
        // vx at (i+1/2, j): :218-229
        if (i < nx - N - 1 && j >= N + 1) {
            float dsxxdx = d_fm<N>(g.sxx, p, 1, c) / dx;
            float dsxzdz = d_bm<N>(g.sxz, p, SZ, c) / dz;
            const float rho_half_x = (g.rho[p] + g.rho[p + 1]) / 2.f;
            if (X.in_h[i] | Z.in[j]) {
                dsxxdx = cpml_apply(dsxxdx, m + 0 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
                dsxzdz = cpml_apply(dsxzdz, m + 3 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
            }
            g.vx[p] += (dsxxdx + dsxzdz) * dt / rho_half_x;
        }
        // vz at (i, j+1/2): :231-242
        if (i >= N + 1 && j < nz - N - 1) {
            const float rho_half_z = 0.5f * (g.rho[p] + g.rho[p + SZ]);
            float dsxzdx = d_bm<N>(g.sxz, p, 1, c) / dx;
            float dszzdz = d_fm<N>(g.szz, p, SZ, c) / dz;
            if (X.in[i] | Z.in_h[j]) {
                dsxzdx = cpml_apply(dsxzdx, m + 2 * NC + p, X.a[i], X.b[i], X.k[i]);
                dszzdz = cpml_apply(dszzdz, m + 5 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
            }
            g.vz[p] += (dszzdz + dsxzdx) * dt / rho_half_z;
        }
    } else {
        // rotated staggered grid, vx and vz at (i, j): :292-318
        if (i >= N + 1 && j >= N + 1) {
            const float sxxrx = Rot::rx_bm<N>(g.sxx, p, SZ, c), sxxrz = Rot::rz_bm<N>(g.sxx, p, SZ, c);
            const float szzrx = Rot::rx_bm<N>(g.szz, p, SZ, c), szzrz = Rot::rz_bm<N>(g.szz, p, SZ, c);
            const float sxzrx = Rot::rx_bm<N>(g.sxz, p, SZ, c), sxzrz = Rot::rz_bm<N>(g.sxz, p, SZ, c);
            const float dx2 = 2.f * dx, dz2 = 2.f * dz;
            float dsxxdx = (sxxrz - sxxrx) / dx2;
            float dszzdz = (szzrz + szzrx) / dz2;
            float dsxzdx = (sxzrz - sxzrx) / dx2;
            float dsxzdz = (sxzrz + sxzrx) / dz2;
            const float rho = g.rho[p];
            if (X.in[i] | Z.in[j]) {
                dsxxdx = cpml_apply(dsxxdx, m + 0 * NC + p, X.a[i], X.b[i], X.k[i]);
                dszzdz = cpml_apply(dszzdz, m + 5 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
                dsxzdx = cpml_apply(dsxzdx, m + 2 * NC + p, X.a[i], X.b[i], X.k[i]);
                dsxzdz = cpml_apply(dsxzdz, m + 3 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
            }
            g.vx[p] += (dsxxdx + dsxzdz) * dt / rho;
            g.vz[p] += (dszzdz + dsxzdx) * dt / rho;
        }
    }
    if (SCHEME != SW_2D_RSG && fo.force && i == fo.ix[bk.z] && j == fo.iz[bk.z])
        force_src2d(g, fo.force, fo.amp[fo.it_force], p);
}

template <int N, int SCHEME, bool SAVE>
__global__ void __launch_bounds__(256) k_stress2d(sw_ssg2d g, long bs, float *sv, long ss, Fold2D fo)
{
    const Blk bk = xcd_block3();
    shift_state(g, (long)bk.z * bs);
    if (SAVE) sv += (long)bk.z * ss;
    if (fo.record && bk.x == 0 && bk.y == 0) {
        // velocity receivers of this step: data[(it*ncomp + comp)*nrec + r]
        const int n = fo.rec.nrec * fo.rec.ncomp;
        float *out = fo.rec.data + (long)bk.z * fo.ds + (long)fo.it_rec * n;
        for (int t = threadIdx.y * blockDim.x + threadIdx.x; t < n; t += blockDim.x * blockDim.y) {
            const int comp = t / fo.rec.nrec, r = t - comp * fo.rec.nrec;
            out[t] = (fo.rec.comp[comp] == 0 ? g.vx : g.vz)[fo.rec.idx[r]];
        }
    }
    const int i = bk.x * blockDim.x + threadIdx.x + N;
    const int j = bk.y * blockDim.y + threadIdx.y + N;
    const int nx = g.nx, nz = g.nz;
    if (i >= nx - N || j >= nz - N) return;
    const long SZ = nx, NC = (long)nx * nz;
    const long p = j * SZ + i;
    const float dx = g.dx, dz = g.dz, dt = g.dt;
    const float *c = g.c;
    float *m = g.mem;
    const sw_axis &X = g.px, &Z = g.pz;

    if (SCHEME == SW_2D_VS) {
        // sxx, szz at (i+1/2, j): src/Elastic2D_CPML.py:172-186
        if (i < nx - N - 1 && j >= N + 1) {
            const float lam_h = 0.5f * (g.lam[p] + g.lam[p + 1]);
            const float mu_h = 0.5f * (g.mu[p] + g.mu[p + 1]);
            const float lpm = lam_h + 2.f * mu_h;
            float dvxdx = d_fm<N>(g.vx, p, 1, c) / dx;
            float dvzdz = d_bm<N>(g.vz, p, SZ, c) / dz;
            if (X.in_h[i] | Z.in[j]) {
                dvxdx = cpml_apply(dvxdx, m + 6 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
                dvzdz = cpml_apply(dvzdz, m + 9 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
            }
            g.sxx[p] += (lpm * dvxdx + lam_h * dvzdz) * dt;
            g.szz[p] += (lpm * dvzdz + lam_h * dvxdx) * dt;
            if (SAVE) {
                sv[2 * NC + p] = dvxdx;
                sv[3 * NC + p] = dvzdz;
            }
        }
        // sxz at (i, j+1/2): :188-199
        if (i >= N + 1 && j < nz - N - 1) {
            const float mu_half_z = 0.5f * (g.mu[p + SZ] + g.mu[p]);
            float dvzdx = d_bm<N>(g.vz, p, 1, c) / dx;
            float dvxdz = d_fm<N>(g.vx, p, SZ, c) / dz;
            if (X.in[i] | Z.in_h[j]) {
                dvzdx = cpml_apply(dvzdx, m + 8 * NC + p, X.a[i], X.b[i], X.k[i]);
                dvxdz = cpml_apply(dvxdz, m + 7 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
            }
            const float sum = dvxdz + dvzdx;
            g.sxz[p] += mu_half_z * sum * dt;
            if (SAVE) sv[4 * NC + p] = sum;
        }
    } else if (SCHEME == SW_2D_SV) {
        // sxx, szz at (i, j): :247-260
        if (i >= N + 1 && j >= N + 1) {
            const float lam = g.lam[p];
            const float lpm = g.lam[p] + 2.f * g.mu[p];
            float dvxdx = d_bm<N>(g.vx, p, 1, c) / dx;
            float dvzdz = d_bm<N>(g.vz, p, SZ, c) / dz;
            if (X.in[i] | Z.in[j]) {
                dvxdx = cpml_apply(dvxdx, m + 6 * NC + p, X.a[i], X.b[i], X.k[i]);
                dvzdz = cpml_apply(dvzdz, m + 9 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
            }
            g.sxx[p] += (lpm * dvxdx + lam * dvzdz) * dt;
            g.szz[p] += (lpm * dvzdz + lam * dvxdx) * dt;
        }
        // sxz at (i+1/2, j+1/2): :262-273
        if (i >= N + 1 && j >= N + 1) {
            const float mu_a = 0.25f * (g.mu[p + SZ] + g.mu[p] + g.mu[p + 1 + SZ] + g.mu[p + 1]);
            float dvzdx = d_fm<N>(g.vz, p, 1, c) / dx;
            float dvxdz = d_fm<N>(g.vx, p, SZ, c) / dz;
            if (X.in_h[i] | Z.in_h[j]) {
                dvzdx = cpml_apply(dvzdx, m + 8 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
                dvxdz = cpml_apply(dvxdz, m + 7 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
            }
            g.sxz[p] += mu_a * (dvxdz + dvzdx) * dt;
        }
    } else {
        // rotated staggered grid stresses at (i+1/2, j+1/2): :326-353
        if (i < nx - N - 1 && j < nz - N - 1) {
            const float lam = g.lam[p], mu = g.mu[p];
            const float lpm = lam + 2.f * mu;
            const float vxrx = Rot::rx_fm<N>(g.vx, p, SZ, c), vxrz = Rot::rz_fm<N>(g.vx, p, SZ, c);
            const float vzrx = Rot::rx_fm<N>(g.vz, p, SZ, c), vzrz = Rot::rz_fm<N>(g.vz, p, SZ, c);
            const float dx2 = 2.f * dx, dz2 = 2.f * dz;
            float dvxdx = (vxrz - vxrx) / dx2;
            float dvxdz = (vxrz + vxrx) / dz2;
            float dvzdx = (vzrz - vzrx) / dx2;
            float dvzdz = (vzrz + vzrx) / dz2;
            if (X.in_h[i] | Z.in_h[j]) {
                dvxdx = cpml_apply(dvxdx, m + 6 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
                dvxdz = cpml_apply(dvxdz, m + 7 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
                dvzdx = cpml_apply(dvzdx, m + 8 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
                dvzdz = cpml_apply(dvzdz, m + 9 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
            }
            g.sxx[p] += (lpm * dvxdx + lam * dvzdz) * dt;
            g.szz[p] += (lpm * dvzdz + lam * dvxdx) * dt;
            g.sxz[p] += (mu * (dvxdz + dvzdx)) * dt;
        }
    }
    if (SCHEME != SW_2D_RSG && fo.stress && i == fo.ix[bk.z] && j == fo.iz[bk.z])
        stress_src2d(g, fo.stress, fo.m0, fo.m2, fo.m4, fo.amp[fo.it_next], p);
}

inline dim3 box_grid2(const sw_ssg2d &g, int N, dim3 blk, int nbatch)
{
    return dim3((g.nx - 2 * N + blk.x - 1) / blk.x, (g.nz - 2 * N + blk.y - 1) / blk.y, nbatch);
}

}  // namespace

int launch_ssg2d_velocity_batch(const sw_ssg2d *g, int nbatch, long bs, const Fold2D &fo, hipStream_t st)
{
    const dim3 blk(kTile2DX, kTile2DY, 1);
    SW_DISPATCH_ORDER(g->order, {
        const dim3 grd = box_grid2(*g, N, blk, nbatch);
        if (g->scheme == SW_2D_VS)
            hipLaunchKernelGGL((k_velocity2d<N, SW_2D_VS, false>), grd, blk, 0, st, *g, bs, nullptr, 0L, fo);
        else if (g->scheme == SW_2D_SV)
            hipLaunchKernelGGL((k_velocity2d<N, SW_2D_SV, false>), grd, blk, 0, st, *g, bs, nullptr, 0L, fo);
        else
            hipLaunchKernelGGL((k_velocity2d<N, SW_2D_RSG, false>), grd, blk, 0, st, *g, bs, nullptr, 0L, fo);
    });
    return launch_status("ssg2d velocity");
}

int launch_ssg2d_stress_batch(const sw_ssg2d *g, int nbatch, long bs, const Fold2D &fo, hipStream_t st)
{
    const dim3 blk(kTile2DX, kTile2DY, 1);
    SW_DISPATCH_ORDER(g->order, {
        const dim3 grd = box_grid2(*g, N, blk, nbatch);
        if (g->scheme == SW_2D_VS)
            hipLaunchKernelGGL((k_stress2d<N, SW_2D_VS, false>), grd, blk, 0, st, *g, bs, nullptr, 0L, fo);
        else if (g->scheme == SW_2D_SV)
            hipLaunchKernelGGL((k_stress2d<N, SW_2D_SV, false>), grd, blk, 0, st, *g, bs, nullptr, 0L, fo);
        else
            hipLaunchKernelGGL((k_stress2d<N, SW_2D_RSG, false>), grd, blk, 0, st, *g, bs, nullptr, 0L, fo);
    });
    return launch_status("ssg2d stress");
}

int launch_ssg2d_sweeps_saved(const sw_ssg2d *g, int nbatch, long bs, bool stress, float *sv, long ss,
                              const Fold2D &fo, hipStream_t st)
{
    const dim3 blk(kTile2DX, kTile2DY, 1);
    SW_REQUIRE(g->scheme == SW_2D_VS, "saved sweeps exist for the VS scheme only");
    SW_DISPATCH_ORDER(g->order, {
        const dim3 grd = box_grid2(*g, N, blk, nbatch);
        if (stress)
            hipLaunchKernelGGL((k_stress2d<N, SW_2D_VS, true>), grd, blk, 0, st, *g, bs, sv, ss, fo);
        else
            hipLaunchKernelGGL((k_velocity2d<N, SW_2D_VS, true>), grd, blk, 0, st, *g, bs, sv, ss, fo);
    });
    return launch_status(stress ? "ssg2d stress (saved)" : "ssg2d velocity (saved)");
}

int launch_ssg2d_velocity(const sw_ssg2d *g, hipStream_t st)
{
    return launch_ssg2d_velocity_batch(g, 1, 0, Fold2D{}, st);
}

int launch_ssg2d_stress(const sw_ssg2d *g, hipStream_t st)
{
    return launch_ssg2d_stress_batch(g, 1, 0, Fold2D{}, st);
}

}  // namespace sw
