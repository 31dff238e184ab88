// 3-D split-field PML sweeps.
//
// Reference: src/Elastic3D_SSG_SPML.py:160-393 (update_SSG_SPML).  Inside the
// PML region (the kernel's f32 coordinate test on any axis) every component is
// split by derivative direction and advanced with a Crank-Nicolson step on the
// damping d only; the component is the sum of its splits.  Elsewhere the plain
// update runs.  Split arrays are full-grid; only PML-region cells touch them.
#include <algorithm>

#include "sw_internal.h"

namespace sw {

namespace {

// A box of cells [i0,i1) x [j0,j1) x [k0,k1).
struct SBox {
    int i0, i1, j0, j1, k0, k1;
};

__device__ __forceinline__ float cn_split(float *q, float n, float d, float incr)
{
    const float v = (n * (*q) + incr) / d;
    *q = v;
    return v;
}

template <int N>
__global__ void __launch_bounds__(256) k_velocity_spml(sw_ssg3d g, SBox bx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + bx.i0;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + bx.j0;
    const int k = blockIdx.z + bx.k0;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= bx.i1 || j >= bx.j1) return;
    const long SY = nx, SZ = (long)nx * ny, NC = SZ * nz;
    const long p = k * SZ + j * SY + i;
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt, hdt = 0.5f * g.dt;
    const float *c = g.c;
    float *q = g.split;

    // vx: src/Elastic3D_SSG_SPML.py:182-205
    if (i < nx - N - 1 && j >= N + 1 && k >= N + 1) {
        const float dsxxdx = d_fm<N>(g.sxx, p, 1, c) / dx;
        const float dsxydy = d_bm<N>(g.sxy, p, SY, c) / dy;
        const float dsxzdz = d_bm<N>(g.sxz, p, SZ, c) / dz;
        const float rho_over = 1.f / g.rho[p];
        if (g.px.in_h[i] | g.py.in[j] | g.pz.in[k]) {
            const float dxp = g.px.d_h[i], dyp = g.py.d[j], dzp = g.pz.d[k];
            const float a = cn_split(q + 0 * NC + p, 1.f - hdt * dxp, 1.f + hdt * dxp, dsxxdx * dt * rho_over);
            const float b = cn_split(q + 1 * NC + p, 1.f - hdt * dyp, 1.f + hdt * dyp, dsxydy * dt * rho_over);
            const float e = cn_split(q + 2 * NC + p, 1.f - hdt * dzp, 1.f + hdt * dzp, dsxzdz * dt * rho_over);
            g.vx[p] = a + b + e;
        } else {
            g.vx[p] += (dsxxdx + dsxydy + dsxzdz) * dt * rho_over;
        }
    }
    // vy: :209-233
    if (j < ny - N - 1 && i >= N + 1 && k >= N + 1) {
        const float rho_half_y = 0.5f * (g.rho[p] + g.rho[p + SY]);
        const float rho_over = 1.f / rho_half_y;
        const float dsxydx = d_bm<N>(g.sxy, p, 1, c) / dx;
        const float dsyydy = d_fm<N>(g.syy, p, SY, c) / dy;
        const float dsyzdz = d_bm<N>(g.syz, p, SZ, c) / dz;
        if (g.px.in[i] | g.py.in_h[j] | g.pz.in[k]) {
            const float dxp = g.px.d[i], dyp = g.py.d_h[j], dzp = g.pz.d[k];
            const float a = cn_split(q + 3 * NC + p, 1.f - hdt * dxp, 1.f + hdt * dxp, dt * dsxydx * rho_over);
            const float b = cn_split(q + 4 * NC + p, 1.f - hdt * dyp, 1.f + hdt * dyp, dt * dsyydy * rho_over);
            const float e = cn_split(q + 5 * NC + p, 1.f - hdt * dzp, 1.f + hdt * dzp, dt * dsyzdz * rho_over);
            g.vy[p] = a + b + e;
        } else {
            g.vy[p] += (dsxydx + dsyydy + dsyzdz) * dt * rho_over;
        }
    }
    // vz: :235-259
    if (k < nz - N - 1 && i >= N + 1 && j >= N + 1) {
        const float rho_avg = 0.25f * (g.rho[p] + g.rho[p + SY] + g.rho[p + 1 + SY] + g.rho[p + 1]);
        const float rho_over = 1.f / rho_avg;
        const float dsxzdx = d_bm<N>(g.sxz, p, 1, c) / dx;
        const float dsyzdy = d_bm<N>(g.syz, p, SY, c) / dy;
        const float dszzdz = d_fm<N>(g.szz, p, SZ, c) / dz;
        if (g.px.in[i] | g.py.in[j] | g.pz.in_h[k]) {
            const float dxp = g.px.d[i], dyp = g.py.d[j], dzp = g.pz.d_h[k];
            const float a = cn_split(q + 6 * NC + p, 1.f - hdt * dxp, 1.f + hdt * dxp, dt * dsxzdx * rho_over);
            const float b = cn_split(q + 7 * NC + p, 1.f - hdt * dyp, 1.f + hdt * dyp, dt * dsyzdy * rho_over);
            const float e = cn_split(q + 8 * NC + p, 1.f - hdt * dzp, 1.f + hdt * dzp, dt * dszzdz * rho_over);
            g.vz[p] = a + b + e;
        } else {
            g.vz[p] += (dsxzdx + dsyzdy + dszzdz) * dt * rho_over;
        }
    }
}

template <int N>
__global__ void __launch_bounds__(256) k_stress_spml(sw_ssg3d g, SBox bx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + bx.i0;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + bx.j0;
    const int k = blockIdx.z + bx.k0;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= bx.i1 || j >= bx.j1) return;
    const long SY = nx, SZ = (long)nx * ny, NC = SZ * nz;
    const long p = k * SZ + j * SY + i;
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt, hdt = 0.5f * g.dt;
    const float *c = g.c;
    float *q = g.split;

    // normal stresses, point lambda and mu: src/Elastic3D_SSG_SPML.py:296-333
    if (i >= N + 1 && j >= N + 1 && k >= N + 1) {
        const float lam = g.lam[p], mu = g.mu[p];
        const float l2m = lam + 2.f * mu;
        const float dvxdx = d_bm<N>(g.vx, p, 1, c) / dx;
        const float dvydy = d_bm<N>(g.vy, p, SY, c) / dy;
        const float dvzdz = d_bm<N>(g.vz, p, SZ, c) / dz;
        if (g.px.in[i] | g.py.in[j] | g.pz.in[k]) {
            const float dxp = g.px.d[i], dyp = g.py.d[j], dzp = g.pz.d[k];
            const float xn = 1.f - hdt * dxp, yn = 1.f - hdt * dyp, zn = 1.f - hdt * dzp;
            const float xd = 1.f + hdt * dxp, yd = 1.f + hdt * dyp, zd = 1.f + hdt * dzp;
            const float a0 = cn_split(q + 9 * NC + p, xn, xd, dt * l2m * dvxdx);
            const float a1 = cn_split(q + 10 * NC + p, yn, yd, dt * lam * dvydy);
            const float a2 = cn_split(q + 11 * NC + p, zn, zd, dt * lam * dvzdz);
            const float b0 = cn_split(q + 12 * NC + p, xn, xd, dt * lam * dvxdx);
            const float b1 = cn_split(q + 13 * NC + p, yn, yd, dt * l2m * dvydy);
            const float b2 = cn_split(q + 14 * NC + p, zn, zd, dt * lam * dvzdz);
            const float c0 = cn_split(q + 21 * NC + p, xn, xd, dt * lam * dvxdx);
            const float c1 = cn_split(q + 22 * NC + p, yn, yd, dt * lam * dvydy);
            const float c2 = cn_split(q + 23 * NC + p, zn, zd, dt * l2m * dvzdz);
            g.sxx[p] = a0 + a1 + a2;
            g.syy[p] = b0 + b1 + b2;
            g.szz[p] = c0 + c1 + c2;
        } else {
            g.sxx[p] += (l2m * dvxdx + lam * (dvzdz + dvydy)) * dt;
            g.syy[p] += (l2m * dvydy + lam * (dvxdx + dvzdz)) * dt;
            g.szz[p] += (l2m * dvzdz + lam * (dvxdx + dvydy)) * dt;
        }
    }
    // sxy: :334-350
    if (i >= N + 1 && j >= N + 1) {
        const float mu_a = 0.25f * (g.mu[p + SY] + g.mu[p] + g.mu[p] + g.mu[p + 1]);
        const float dvydx = d_fm<N>(g.vy, p, 1, c) / dx;
        const float dvxdy = d_fm<N>(g.vx, p, SY, c) / dy;
        if (g.px.in_h[i] | g.py.in_h[j] | g.pz.in[k]) {
            const float dxp = g.px.d_h[i], dyp = g.py.d_h[j];
            const float a = cn_split(q + 15 * NC + p, 1.f - hdt * dxp, 1.f + hdt * dxp, dt * mu_a * dvydx);
            const float b = cn_split(q + 16 * NC + p, 1.f - hdt * dyp, 1.f + hdt * dyp, dt * mu_a * dvxdy);
            g.sxy[p] = a + b;
        } else {
            g.sxy[p] += mu_a * (dvxdy + dvydx) * dt;
        }
    }
    // sxz: :352-371
    if (i < nx - N - 1 && k < nz - N - 1) {
        const float mu_a = 0.25f * (g.mu[p] + g.mu[p + 1] + g.mu[p] + g.mu[p + SY]);
        const float dvzdx = d_fm<N>(g.vz, p, 1, c) / dx;
        const float dvxdz = d_fm<N>(g.vx, p, SZ, c) / dz;
        if (g.px.in_h[i] | g.py.in[j] | g.pz.in_h[k]) {
            const float dxp = g.px.d_h[i], dzp = g.pz.d_h[k];
            const float a = cn_split(q + 17 * NC + p, 1.f - hdt * dxp, 1.f + hdt * dxp, dt * mu_a * dvzdx);
            const float b = cn_split(q + 18 * NC + p, 1.f - hdt * dzp, 1.f + hdt * dzp, dt * mu_a * dvxdz);
            g.sxz[p] = a + b;
        } else {
            g.sxz[p] += mu_a * (dvxdz + dvzdx) * dt;
        }
    }
    // syz: :373-391
    if (j < ny - N - 1 && k < nz - N - 1) {
        const float mu_a = 0.25f * (g.mu[p + SY] + g.mu[p] + g.mu[p + SZ] + g.mu[p]);
        const float dvzdy = d_fm<N>(g.vz, p, SY, c) / dy;
        const float dvydz = d_fm<N>(g.vy, p, SZ, c) / dz;
        if (g.px.in[i] | g.py.in_h[j] | g.pz.in_h[k]) {
            const float dyp = g.py.d_h[j], dzp = g.pz.d_h[k];
            const float a = cn_split(q + 19 * NC + p, 1.f - hdt * dyp, 1.f + hdt * dyp, dt * mu_a * dvzdy);
            const float b = cn_split(q + 20 * NC + p, 1.f - hdt * dzp, 1.f + hdt * dzp, dt * mu_a * dvydz);
            g.syz[p] = a + b;
        } else {
            g.syz[p] += mu_a * (dvydz + dvzdy) * dt;
        }
    }
}

constexpr int kSX = 64, kSY = 4;

// Split-free interior box (plain updates, z-marched in ssg3d.hip) and the
// shell of up to six boxes around it (thread-per-cell kernels above).
struct SPlan {
    int n = 0;
    SBox shell[6];
    SBox inner{0, 0, 0, 0, 0, 0};
};

void add(SPlan &p, int i0, int i1, int j0, int j1, int k0, int k1)
{
    if (i1 > i0 && j1 > j0 && k1 > k0) p.shell[p.n++] = SBox{i0, i1, j0, j1, k0, k1};
}

SPlan make_splan(const sw_ssg3d &g, int N)
{
    SPlan p;
    const int x0 = N, x1 = g.nx - N, y0 = N, y1 = g.ny - N, z0 = N, z1 = g.nz - N;
    const int xl = std::max(x0, g.px.lo), xh = std::min(x1, g.px.hi);
    const int yl = std::max(y0, g.py.lo), yh = std::min(y1, g.py.hi);
    const int zl = std::max(z0, g.pz.lo), zh = std::min(z1, g.pz.hi);
    if (xl < xh && yl < yh && zl < zh) {
        p.inner = SBox{xl, xh, yl, yh, zl, zh};
        add(p, x0, x1, y0, y1, z0, zl);
        add(p, x0, x1, y0, y1, zh, z1);
        add(p, x0, x1, y0, yl, zl, zh);
        add(p, x0, x1, yh, y1, zl, zh);
        add(p, x0, xl, yl, yh, zl, zh);
        add(p, xh, x1, yl, yh, zl, zh);
    } else {
        add(p, x0, x1, y0, y1, z0, z1);
    }
    return p;
}

inline dim3 sgrid(const SBox &b)
{
    return dim3((b.i1 - b.i0 + kSX - 1) / kSX, (b.j1 - b.j0 + kSY - 1) / kSY, b.k1 - b.k0);
}

}  // namespace

int launch_ssg3d_spml_interior(const sw_ssg3d *g, int i0, int i1, int j0, int j1, int k0, int k1, bool stress,
                               hipStream_t st);

int launch_ssg3d_velocity_spml(const sw_ssg3d *g, hipStream_t st)
{
    const dim3 blk(kSX, kSY, 1);
    SW_DISPATCH_ORDER(g->order, {
        const SPlan p = make_splan(*g, N);
        for (int t = 0; t < p.n; ++t)
            hipLaunchKernelGGL((k_velocity_spml<N>), sgrid(p.shell[t]), blk, 0, st, *g, p.shell[t]);
        const SBox &b = p.inner;
        int rc = launch_ssg3d_spml_interior(g, b.i0, b.i1, b.j0, b.j1, b.k0, b.k1, false, st);
        if (rc) return rc;
    });
    return launch_status("ssg3d velocity (split-field PML)");
}

int launch_ssg3d_stress_spml(const sw_ssg3d *g, hipStream_t st)
{
    const dim3 blk(kSX, kSY, 1);
    SW_DISPATCH_ORDER(g->order, {
        const SPlan p = make_splan(*g, N);
        for (int t = 0; t < p.n; ++t)
            hipLaunchKernelGGL((k_stress_spml<N>), sgrid(p.shell[t]), blk, 0, st, *g, p.shell[t]);
        const SBox &b = p.inner;
        int rc = launch_ssg3d_spml_interior(g, b.i0, b.i1, b.j0, b.j1, b.k0, b.k1, true, st);
        if (rc) return rc;
    });
    return launch_status("ssg3d stress (split-field PML)");
}

}  // namespace sw
