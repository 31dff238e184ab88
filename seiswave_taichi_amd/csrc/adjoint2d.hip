// Exact discrete adjoint of the 2-D VS step (src/Elastic2D_CPML.py:127-200)
// with CPML memory variables, and the model gradient it yields.
//
// The forward step, as a map of the state u = (vx, vz, sxx, szz, sxz, m0..m9):
//   V (velocity stage):  vx += (Dx' sxx + Dz' sxz) * dt/rho          (vx cells)
//                        vz += (Dz' szz + Dx' sxz) * dt/rho4         (vz cells)
//   S (stress stage):    sxx += (lpm Dx' vx + lamh Dz' vz) dt,
//                        szz += (lpm Dz' vz + lamh Dx' vx) dt         (sxx cells)
//                        sxz += muz (Dz' vx + Dx' vz) dt              (sxz cells)
// where D' = D/kappa + m_new, m_new = b m + a D inside the PML test and D' = D
// outside.  The adjoint runs S^T then V^T per step, backwards in time.  Each
// transpose is two kernels: pass A evaluates, per output cell, the adjoint of
// its pre-CPML derivatives (and updates the memory-variable adjoints and the
// model-coefficient gradients, which need the forward derivatives recomputed
// from the stored forward state); pass B gathers those into the adjoint of the
// differentiated fields, the transpose of the staggered stencils.
#include "sw_internal.h"

namespace sw {

namespace {

// Adjoint of one CPML-corrected derivative.  Forward: m1 = b m + a D,
// D' = D/k + m1.  Given Dbar' (adjoint of D') and mbar (adjoint of m1 from
// later steps), returns Dbar and replaces mbar by the adjoint of m.
__device__ __forceinline__ float cpml_adjoint(float dbar_p, float *mbar, float a, float b, float k)
{
    const float mt = *mbar + dbar_p;
    *mbar = b * mt;
    return dbar_p / k + a * mt;
}

__device__ __forceinline__ float fwd_cpml(float D, float m, float a, float b, float k)
{
    return D / k + (b * m + a * D);
}

// ------------------------------------------------------------- stress stage
// SAVED: the forward derivatives come from the planes the saved stress sweep
// wrote (sv + z * ss for shot z, planes 2-4) instead of being recomputed from
// the forward state; they are the same values, so the result is the same.
// The four stage passes of one cell (i, j) of one shot (f, q, sv already at
// the shot's offsets), writing / reading the pass-A outputs at tmp (4 planes).
template <int N, bool SAVED>
__device__ __forceinline__ void stress_a_cell(const sw_ssg2d &f, const sw_adj2d &q, const float *sv, float *tmp, int i,
                                              int j)
{
    const int nx = f.nx, nz = f.nz;
    const long SZ = nx, NC = (long)nx * nz;
    const long p = (long)j * SZ + i;
    const float dx = f.dx, dz = f.dz, dt = f.dt;
    const float *c = f.c;
    const sw_axis &X = f.px, &Z = f.pz;
    float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
    // sxx, szz at (i+1/2, j): forward src/Elastic2D_CPML.py:172-186
    if (i >= N && i < nx - N - 1 && j >= N + 1 && j < nz - N) {
        const float gxx = q.sxx[p] * dt, gzz = q.szz[p] * dt;
        const float lam_h = 0.5f * (f.lam[p] + f.lam[p + 1]);
        const float mu_h = 0.5f * (f.mu[p] + f.mu[p + 1]);
        const float lpm = lam_h + 2.f * mu_h;
        const bool in = X.in_h[i] | Z.in[j];
        float Dxp, Dzp;
        if (SAVED) {
            Dxp = sv[2 * NC + p];
            Dzp = sv[3 * NC + p];
        } else {
            Dxp = d_fm<N>(f.vx, p, 1, c) / dx;
            Dzp = d_bm<N>(f.vz, p, SZ, c) / dz;
            if (in) {
                Dxp = fwd_cpml(Dxp, f.mem[6 * NC + p], X.a_h[i], X.b_h[i], X.k_h[i]);
                Dzp = fwd_cpml(Dzp, f.mem[9 * NC + p], Z.a[j], Z.b[j], Z.k[j]);
            }
        }
        // d/d lam_h = (gxx+gzz)(Dx'+Dz'); d/d mu_h = 2(gxx Dx' + gzz Dz').  The vs and
        // rho gradients need mu_bar - 2 lam_bar, whose two terms nearly cancel for
        // P-dominated data; accumulate it in its cancellation-free form instead.
        q.g_lamh[p] += (gxx + gzz) * (Dxp + Dzp);
        q.g_muh[p] += -2.f * (gxx * Dzp + gzz * Dxp);
        float bx = lpm * gxx + lam_h * gzz;
        float bz = lam_h * gxx + lpm * gzz;
        if (in) {
            bx = cpml_adjoint(bx, q.mem + 6 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
            bz = cpml_adjoint(bz, q.mem + 9 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
        }
        t0 = bx / dx;
        t1 = bz / dz;
    }
    // sxz at (i, j+1/2): :188-199
    if (i >= N + 1 && i < nx - N && j >= N && j < nz - N - 1) {
        const float g = q.sxz[p] * dt;
        const float mu_z = 0.5f * (f.mu[p + SZ] + f.mu[p]);
        const bool in = X.in[i] | Z.in_h[j];
        float sum;
        if (SAVED) {
            sum = sv[4 * NC + p];
        } else {
            float Dxp = d_bm<N>(f.vz, p, 1, c) / dx;
            float Dzp = d_fm<N>(f.vx, p, SZ, c) / dz;
            if (in) {
                Dxp = fwd_cpml(Dxp, f.mem[8 * NC + p], X.a[i], X.b[i], X.k[i]);
                Dzp = fwd_cpml(Dzp, f.mem[7 * NC + p], Z.a_h[j], Z.b_h[j], Z.k_h[j]);
            }
            sum = Dzp + Dxp;
        }
        q.g_muz[p] += g * sum;
        float bx = mu_z * g, bz = mu_z * g;
        if (in) {
            bx = cpml_adjoint(bx, q.mem + 8 * NC + p, X.a[i], X.b[i], X.k[i]);
            bz = cpml_adjoint(bz, q.mem + 7 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
        }
        t2 = bx / dx;
        t3 = bz / dz;
    }
    tmp[0 * NC + p] = t0;
    tmp[1 * NC + p] = t1;
    tmp[2 * NC + p] = t2;
    tmp[3 * NC + p] = t3;
}

__device__ __forceinline__ float tap(const float *t, long idx, long n)
{
    return (idx >= 0 && idx < n) ? t[idx] : 0.f;
}

template <int N>
__device__ __forceinline__ void stress_b_cell(const sw_ssg2d &f, const sw_adj2d &q, const float *tmp, int i, int j)
{
    const int nx = f.nx, nz = f.nz;
    const long SZ = nx, NC = (long)nx * nz;
    const long p = (long)j * SZ + i;
    const float *c = f.c;
    const float *T0 = tmp, *T1 = tmp + NC, *T2 = tmp + 2 * NC, *T3 = tmp + 3 * NC;
    float ax = 0.f, az = 0.f;
#pragma unroll
    for (int id = 0; id < N; ++id) {
        // sxx cells read vx[p+id+1] (+) and vx[p-id] (-) along x (d_fm)
        ax += c[id] * (((i - id - 1) >= 0 ? T0[p - id - 1] : 0.f) - ((i + id) < nx ? T0[p + id] : 0.f));
        // sxz cells read vx[p+(id+1)SZ] (+) and vx[p-id SZ] (-) along z (d_fm)
        ax += c[id] * (tap(T3, p - (id + 1) * SZ, NC) - tap(T3, p + id * SZ, NC));
        // sxx cells read vz[p+id SZ] (+) and vz[p-(id+1)SZ] (-) along z (d_bm)
        az += c[id] * (tap(T1, p - id * SZ, NC) - tap(T1, p + (id + 1) * SZ, NC));
        // sxz cells read vz[p+id] (+) and vz[p-id-1] (-) along x (d_bm)
        az += c[id] * (((i - id) >= 0 ? T2[p - id] : 0.f) - ((i + id + 1) < nx ? T2[p + id + 1] : 0.f));
    }
    q.vx[p] += ax;
    q.vz[p] += az;
}

// ----------------------------------------------------------- velocity stage
// SAVED: planes 0-1 of the saved velocity sweep, as in k_adj_stress_a.
template <int N, bool SAVED>
__device__ __forceinline__ void velocity_a_cell(const sw_ssg2d &f, const sw_adj2d &q, const float *sv, float *tmp,
                                                int i, int j)
{
    const int nx = f.nx, nz = f.nz;
    const long SZ = nx, NC = (long)nx * nz;
    const long p = (long)j * SZ + i;
    const float dx = f.dx, dz = f.dz, dt = f.dt;
    const float *c = f.c;
    const sw_axis &X = f.px, &Z = f.pz;
    float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
    // vx at (i, j): src/Elastic2D_CPML.py:144-154
    if (i >= N + 1 && i < nx - N && j >= N + 1 && j < nz - N) {
        const float vb = q.vx[p];
        const float g = vb * dt / f.rho[p];
        const bool in = X.in[i] | Z.in[j];
        float sum;
        if (SAVED) {
            sum = sv[p];
        } else {
            float Dxp = d_bm<N>(f.sxx, p, 1, c) / dx;
            float Dzp = d_bm<N>(f.sxz, p, SZ, c) / dz;
            if (in) {
                Dxp = fwd_cpml(Dxp, f.mem[0 * NC + p], X.a[i], X.b[i], X.k[i]);
                Dzp = fwd_cpml(Dzp, f.mem[3 * NC + p], Z.a[j], Z.b[j], Z.k[j]);
            }
            sum = Dxp + Dzp;
        }
        q.g_b1[p] += vb * sum;
        float bx = g, bz = g;
        if (in) {
            bx = cpml_adjoint(bx, q.mem + 0 * NC + p, X.a[i], X.b[i], X.k[i]);
            bz = cpml_adjoint(bz, q.mem + 3 * NC + p, Z.a[j], Z.b[j], Z.k[j]);
        }
        t0 = bx / dx;
        t1 = bz / dz;
    }
    // vz at (i+1/2, j+1/2): :156-167
    if (i >= N && i < nx - N - 1 && j >= N && j < nz - N - 1) {
        const float vb = q.vz[p];
        const float rho4 = 0.25f * (f.rho[p] + f.rho[p + SZ] + f.rho[p + 1 + SZ] + f.rho[p + 1]);
        const float g = vb * dt / rho4;
        const bool in = X.in_h[i] | Z.in_h[j];
        float sum;
        if (SAVED) {
            sum = sv[NC + p];
        } else {
            float Dxp = d_fm<N>(f.sxz, p, 1, c) / dx;
            float Dzp = d_fm<N>(f.szz, p, SZ, c) / dz;
            if (in) {
                Dxp = fwd_cpml(Dxp, f.mem[2 * NC + p], X.a_h[i], X.b_h[i], X.k_h[i]);
                Dzp = fwd_cpml(Dzp, f.mem[5 * NC + p], Z.a_h[j], Z.b_h[j], Z.k_h[j]);
            }
            sum = Dzp + Dxp;
        }
        q.g_b4[p] += vb * sum;
        float bx = g, bz = g;
        if (in) {
            bx = cpml_adjoint(bx, q.mem + 2 * NC + p, X.a_h[i], X.b_h[i], X.k_h[i]);
            bz = cpml_adjoint(bz, q.mem + 5 * NC + p, Z.a_h[j], Z.b_h[j], Z.k_h[j]);
        }
        t2 = bx / dx;
        t3 = bz / dz;
    }
    tmp[0 * NC + p] = t0;
    tmp[1 * NC + p] = t1;
    tmp[2 * NC + p] = t2;
    tmp[3 * NC + p] = t3;
}

template <int N>
__device__ __forceinline__ void velocity_b_cell(const sw_ssg2d &f, const sw_adj2d &q, const float *tmp, int i, int j)
{
    const int nx = f.nx, nz = f.nz;
    const long SZ = nx, NC = (long)nx * nz;
    const long p = (long)j * SZ + i;
    const float *c = f.c;
    const float *T0 = tmp, *T1 = tmp + NC, *T2 = tmp + 2 * NC, *T3 = tmp + 3 * NC;
    float axx = 0.f, axz = 0.f, azz = 0.f;
#pragma unroll
    for (int id = 0; id < N; ++id) {
        // vx cells read sxx[p+id] (+), sxx[p-id-1] (-) along x (d_bm)
        axx += c[id] * (((i - id) >= 0 ? T0[p - id] : 0.f) - ((i + id + 1) < nx ? T0[p + id + 1] : 0.f));
        // vx cells read sxz[p+id SZ] (+), sxz[p-(id+1)SZ] (-) along z (d_bm)
        axz += c[id] * (tap(T1, p - id * SZ, NC) - tap(T1, p + (id + 1) * SZ, NC));
        // vz cells read sxz[p+id+1] (+), sxz[p-id] (-) along x (d_fm)
        axz += c[id] * (((i - id - 1) >= 0 ? T2[p - id - 1] : 0.f) - ((i + id) < nx ? T2[p + id] : 0.f));
        // vz cells read szz[p+(id+1)SZ] (+), szz[p-id SZ] (-) along z (d_fm)
        azz += c[id] * (tap(T3, p - (id + 1) * SZ, NC) - tap(T3, p + id * SZ, NC));
    }
    q.sxx[p] += axx;
    q.sxz[p] += axz;
    q.szz[p] += azz;
}

// One thread per cell of shot bk.z: the forward state at + z * fs (SAVED:
// derivative planes at sv + z * ss), the adjoint block at + z * as.  The
// stage kernels keep the dispatch order of their blocks: with the XCD
// renumbering the sweeps use, the two fused kernels of a backward step ran
// 70.1 and 84.0 us on MI355X (16 shots of 800x300) against 65.9 and 78.8.
#define SW_ADJ_CELL_PROLOGUE                                            \
    const Blk bk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};    \
    shift_state(f, (long)bk.z * fs);                                    \
    shift_adj(q, (long)bk.z * as);                                      \
    const int i = bk.x * blockDim.x + threadIdx.x;                      \
    const int j = bk.y * blockDim.y + threadIdx.y;                      \
    if (i >= f.nx || j >= f.nz) return;

template <int N, bool SAVED>
__global__ void __launch_bounds__(256) k_adj_stress_a(sw_ssg2d f, sw_adj2d q, long fs, long as, const float *sv,
                                                      long ss)
{
    SW_ADJ_CELL_PROLOGUE
    stress_a_cell<N, SAVED>(f, q, SAVED ? sv + (long)bk.z * ss : nullptr, q.tmp, i, j);
}

template <int N>
__global__ void __launch_bounds__(256) k_adj_stress_b(sw_ssg2d f, sw_adj2d q, long fs, long as)
{
    SW_ADJ_CELL_PROLOGUE
    stress_b_cell<N>(f, q, q.tmp, i, j);
}

template <int N, bool SAVED>
__global__ void __launch_bounds__(256) k_adj_velocity_a(sw_ssg2d f, sw_adj2d q, long fs, long as, const float *sv,
                                                        long ss)
{
    SW_ADJ_CELL_PROLOGUE
    velocity_a_cell<N, SAVED>(f, q, SAVED ? sv + (long)bk.z * ss : nullptr, q.tmp, i, j);
}

template <int N>
__global__ void __launch_bounds__(256) k_adj_velocity_b(sw_ssg2d f, sw_adj2d q, long fs, long as)
{
    SW_ADJ_CELL_PROLOGUE
    velocity_b_cell<N>(f, q, q.tmp, i, j);
}

// Fused passes of the segment: a gather (pass B) and the next stage's pass A
// of the same cell, which reads only that cell of what the gather wrote.  The
// stress stage's pass-A outputs live in q.tmp, the velocity stage's in tv
// (tv + z * ts for shot z), so neither pass overwrites what the other reads.
template <int N>
__global__ void __launch_bounds__(256) k_adj_sb_va(sw_ssg2d f, sw_adj2d q, long as, const float *sv, long ss,
                                                   float *tv, long ts)
{
    const long fs = 0;
    SW_ADJ_CELL_PROLOGUE
    stress_b_cell<N>(f, q, q.tmp, i, j);
    velocity_a_cell<N, true>(f, q, sv + (long)bk.z * ss, tv + (long)bk.z * ts, i, j);
}

template <int N, bool WITH_SA>
__global__ void __launch_bounds__(256) k_adj_vb_sa(sw_ssg2d f, sw_adj2d q, long as, const float *sv, long ss,
                                                   float *tv, long ts)
{
    const long fs = 0;
    SW_ADJ_CELL_PROLOGUE
    velocity_b_cell<N>(f, q, tv + (long)bk.z * ts, i, j);
    if (WITH_SA) stress_a_cell<N, true>(f, q, sv + (long)bk.z * ss, q.tmp, i, j);
}
#undef SW_ADJ_CELL_PROLOGUE

// Residual injection: adjoint of the receiver gather data[it] = field[idx].
// One thread per (component, receiver).  Receivers at the same cell add up;
// with such duplicates their sum order is not fixed (hardware f32 atomics).
// Shot blockIdx.y of a batch: adjoint at + y * as, residuals at + y * ds.
__global__ void k_adj_inject(sw_adj2d q, sw_receivers r, int it, long as, long ds)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= r.ncomp * r.nrec) return;
    shift_adj(q, (long)blockIdx.y * as);
    r.data += (long)blockIdx.y * ds;
    const int comp = t / r.nrec, k = t % r.nrec;
    float *fld = r.comp[comp] == 0 ? q.vx : (r.comp[comp] == 2 ? q.vz : (r.comp[comp] == 3 ? q.sxx :
                 (r.comp[comp] == 5 ? q.szz : q.sxz)));
    const float res = r.data[((long)it * r.ncomp + comp) * r.nrec + k];
    atomicAdd(fld + r.idx[k], res);
}

// Sum over the nb shots of a batch of one accumulator at cell p.
__device__ __forceinline__ float acc_sum(const float *a, long p, int nb, long as)
{
    float v = a[p];
    for (int b = 1; b < nb; ++b) v += a[p + (long)b * as];
    return v;
}

// Model gradient from the accumulated coefficient gradients (summed over the
// nb shots of a batch first):
//   lam_h = (lam[p]+lam[p+1])/2, mu_h likewise (sxx cells); mu_z = (mu[p+SZ]+mu[p])/2
//   (sxz cells); b1 = dt/rho (vx cells); b4 = dt/rho4 (vz cells); then
//   mu = rho vs^2, lam = rho (vp^2 - 2 vs^2).
__global__ void k_model_gradient(sw_ssg2d f, sw_adj2d q, const float *vp, const float *vs, float *gvp,
                                 float *gvs, float *grho, int nb, long as)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int nx = f.nx, nz = f.nz;
    if (i >= nx || j >= nz) return;
    const long SZ = nx;
    const long p = (long)j * SZ + i;
    const float dt = f.dt;
    // gl = d/d lam; gb = d/d mu - 2 d/d lam (g_muh holds the sxx-cell part of the
    // latter, g_muz the sxz-cell mu gradient)
    float gl = 0.5f * acc_sum(q.g_lamh, p, nb, as) + (i >= 1 ? 0.5f * acc_sum(q.g_lamh, p - 1, nb, as) : 0.f);
    float gb = 0.5f * acc_sum(q.g_muh, p, nb, as) + (i >= 1 ? 0.5f * acc_sum(q.g_muh, p - 1, nb, as) : 0.f);
    gb += 0.5f * acc_sum(q.g_muz, p, nb, as) + (j >= 1 ? 0.5f * acc_sum(q.g_muz, p - SZ, nb, as) : 0.f);
    const float r = f.rho[p];
    float gr = acc_sum(q.g_b1, p, nb, as) * (-dt / (r * r));
    // vz cell p' uses rho4 over (p', p'+SZ, p'+1+SZ, p'+1): cell p is in the stencils
    // of p, p-SZ, p-1-SZ, p-1
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int ii = i - a, jj = j - b;
            if (ii < 0 || jj < 0 || ii >= nx - 1 || jj >= nz - 1) continue;
            const long pp = (long)jj * SZ + ii;
            const float r4 = 0.25f * (f.rho[pp] + f.rho[pp + SZ] + f.rho[pp + 1 + SZ] + f.rho[pp + 1]);
            gr += 0.25f * acc_sum(q.g_b4, pp, nb, as) * (-dt / (r4 * r4));
        }
    // lam = rho (vp^2 - 2 vs^2), mu = rho vs^2:
    //   d/dvp = 2 rho vp lam_bar;  d/dvs = 2 rho vs (mu_bar - 2 lam_bar);
    //   d/drho (moduli part) = lam_bar vp^2 + vs^2 (mu_bar - 2 lam_bar)
    const float a = vp[p], s = vs[p];
    gvp[p] = gl * 2.f * r * a;
    gvs[p] = gb * 2.f * r * s;
    grho[p] = gr + gl * (a * a) + gb * (s * s);
}

inline dim3 grid2(const sw_ssg2d &g, dim3 blk, int nbatch = 1)
{
    return dim3((g.nx + blk.x - 1) / blk.x, (g.nz + blk.y - 1) / blk.y, nbatch);
}

}  // namespace

// sv == nullptr: recompute the forward derivatives from the state f holds
// (shot b's at + b * fs); otherwise read them from the saved planes.
int launch_adj2d_stress_batch(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long fs, long as, const float *sv,
                              long ss, hipStream_t st)
{
    const dim3 blk(kTile2DX, kTile2DY, 1);
    SW_DISPATCH_ORDER(f->order, {
        const dim3 grd = grid2(*f, blk, nbatch);
        if (sv)
            hipLaunchKernelGGL((k_adj_stress_a<N, true>), grd, blk, 0, st, *f, *q, fs, as, sv, ss);
        else
            hipLaunchKernelGGL((k_adj_stress_a<N, false>), grd, blk, 0, st, *f, *q, fs, as, nullptr, 0L);
        hipLaunchKernelGGL((k_adj_stress_b<N>), grd, blk, 0, st, *f, *q, fs, as);
    });
    return launch_status("2-D adjoint stress stage");
}

int launch_adj2d_velocity_batch(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long fs, long as,
                                const float *sv, long ss, hipStream_t st)
{
    const dim3 blk(kTile2DX, kTile2DY, 1);
    SW_DISPATCH_ORDER(f->order, {
        const dim3 grd = grid2(*f, blk, nbatch);
        if (sv)
            hipLaunchKernelGGL((k_adj_velocity_a<N, true>), grd, blk, 0, st, *f, *q, fs, as, sv, ss);
        else
            hipLaunchKernelGGL((k_adj_velocity_a<N, false>), grd, blk, 0, st, *f, *q, fs, as, nullptr, 0L);
        hipLaunchKernelGGL((k_adj_velocity_b<N>), grd, blk, 0, st, *f, *q, fs, as);
    });
    return launch_status("2-D adjoint velocity stage");
}

// Passes of the fused segment schedule (SAVED derivative planes sv + z * ss,
// velocity-stage scratch tv + z * ts): 0 stress pass A alone, 1 stress
// gather + velocity pass A, 2 velocity gather + stress pass A, 3 velocity
// gather alone.
int launch_adj2d_fused(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long as, int pass, const float *sv, long ss,
                       float *tv, long ts, hipStream_t st)
{
    const dim3 blk(kTile2DX, kTile2DY, 1);
    SW_DISPATCH_ORDER(f->order, {
        const dim3 grd = grid2(*f, blk, nbatch);
        if (pass == 0)
            hipLaunchKernelGGL((k_adj_stress_a<N, true>), grd, blk, 0, st, *f, *q, 0L, as, sv, ss);
        else if (pass == 1)
            hipLaunchKernelGGL((k_adj_sb_va<N>), grd, blk, 0, st, *f, *q, as, sv, ss, tv, ts);
        else if (pass == 2)
            hipLaunchKernelGGL((k_adj_vb_sa<N, true>), grd, blk, 0, st, *f, *q, as, sv, ss, tv, ts);
        else
            hipLaunchKernelGGL((k_adj_vb_sa<N, false>), grd, blk, 0, st, *f, *q, as, sv, ss, tv, ts);
    });
    return launch_status("2-D adjoint fused pass");
}

int launch_adj2d_stress(const sw_ssg2d *f, const sw_adj2d *q, hipStream_t st)
{
    return launch_adj2d_stress_batch(f, q, 1, 0, 0, nullptr, 0, st);
}

int launch_adj2d_velocity(const sw_ssg2d *f, const sw_adj2d *q, hipStream_t st)
{
    return launch_adj2d_velocity_batch(f, q, 1, 0, 0, nullptr, 0, st);
}

int launch_adj2d_inject_batch(const sw_adj2d *q, const sw_receivers *r, int it, int nbatch, long as, long ds,
                              hipStream_t st)
{
    const int n = r->ncomp * r->nrec;
    hipLaunchKernelGGL(k_adj_inject, dim3((n + 255) / 256, nbatch), dim3(256), 0, st, *q, *r, it, as, ds);
    return launch_status("2-D adjoint residual injection");
}

int launch_adj2d_inject(const sw_adj2d *q, const sw_receivers *r, int it, hipStream_t st)
{
    return launch_adj2d_inject_batch(q, r, it, 1, 0, 0, st);
}

int launch_adj2d_model_gradient_batch(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long as, const float *vp,
                                      const float *vs, float *gvp, float *gvs, float *grho, hipStream_t st)
{
    const dim3 blk(kTile2DX, kTile2DY, 1);
    hipLaunchKernelGGL(k_model_gradient, grid2(*f, blk), blk, 0, st, *f, *q, vp, vs, gvp, gvs, grho, nbatch, as);
    return launch_status("2-D model gradient");
}

int launch_adj2d_model_gradient(const sw_ssg2d *f, const sw_adj2d *q, const float *vp, const float *vs, float *gvp,
                                float *gvs, float *grho, hipStream_t st)
{
    return launch_adj2d_model_gradient_batch(f, q, 1, 0, vp, vs, gvp, gvs, grho, st);
}

}  // namespace sw
