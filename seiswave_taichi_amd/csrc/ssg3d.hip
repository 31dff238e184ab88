// 3-D staggered-grid velocity/stress sweeps with slab-confined CPML.
//
// Reference: src/Elastic3D_SSG_CPML.py:144-309 (update_SSG) and
// src/Elastic3D_SSG_Tensor.py:166-345 (same maths, other vz term order).
// One thread per cell of the [N, n-N) box; each thread evaluates every
// component whose reference loop range contains the cell.
#include "sw_internal.h"

namespace sw {

namespace {

struct Slabs {
    int xs, ys, zs;
    long ix, iy, iz;     // index inside the slab arrays
    long nxs, nys, nzs;  // size of one slab array
};

__device__ __forceinline__ Slabs slabs_of(const sw_ssg3d &g, int i, int j, int k)
{
    Slabs s;
    const int wx = g.px.lo + (g.nx - g.px.hi);
    const int wy = g.py.lo + (g.ny - g.py.hi);
    const int wz = g.pz.lo + (g.nz - g.pz.hi);
    s.xs = slab_index(i, g.px.lo, g.px.hi);
    s.ys = slab_index(j, g.py.lo, g.py.hi);
    s.zs = slab_index(k, g.pz.lo, g.pz.hi);
    s.ix = ((long)k * g.ny + j) * wx + s.xs;
    s.iy = ((long)k * wy + s.ys) * g.nx + i;
    s.iz = ((long)s.zs * g.ny + j) * g.nx + i;
    s.nxs = (long)g.nz * g.ny * wx;
    s.nys = (long)g.nz * wy * g.nx;
    s.nzs = (long)wz * g.ny * g.nx;
    return s;
}

template <int N, int CLS>
__global__ void __launch_bounds__(256) k_velocity(sw_ssg3d g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + N;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + N;
    const int k = blockIdx.z + N;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= nx - N || j >= ny - N || k >= nz - N) return;
    const long SY = nx, SZ = (long)nx * ny;
    const long p = k * SZ + j * SY + i;
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    const float *c = g.c;
    const Slabs s = slabs_of(g, i, j, k);

    // vx at (i+1/2, j, k): src/Elastic3D_SSG_CPML.py:164-178
    if (i < nx - N - 1 && j >= N + 1 && k >= N + 1) {
        float dsxx = d_fm<N>(g.sxx, p, 1, c) / dx;
        float dsxy = d_bm<N>(g.sxy, p, SY, c) / dy;
        float dsxz = d_bm<N>(g.sxz, p, SZ, c) / dz;
        if (s.xs >= 0) dsxx = cpml_apply(dsxx, g.mem_x + 0 * s.nxs + s.ix, g.px.a_h[i], g.px.b_h[i], g.px.k_h[i]);
        if (s.ys >= 0) dsxy = cpml_apply(dsxy, g.mem_y + 0 * s.nys + s.iy, g.py.a[j], g.py.b[j], g.py.k[j]);
        if (s.zs >= 0) dsxz = cpml_apply(dsxz, g.mem_z + 0 * s.nzs + s.iz, g.pz.a[k], g.pz.b[k], g.pz.k[k]);
        g.vx[p] += (dsxx + dsxz + dsxy) * dt / g.rho[p];
    }
    // vy at (i, j+1/2, k): :180-195
    if (j < ny - N - 1 && i >= N + 1 && k >= N + 1) {
        const float rho_half_y = 0.5f * (g.rho[p] + g.rho[p + SY]);
        float dsxy = d_bm<N>(g.sxy, p, 1, c) / dx;
        float dsyy = d_fm<N>(g.syy, p, SY, c) / dy;
        float dsyz = d_bm<N>(g.syz, p, SZ, c) / dz;
        if (s.xs >= 0) dsxy = cpml_apply(dsxy, g.mem_x + 1 * s.nxs + s.ix, g.px.a[i], g.px.b[i], g.px.k[i]);
        if (s.ys >= 0) dsyy = cpml_apply(dsyy, g.mem_y + 1 * s.nys + s.iy, g.py.a_h[j], g.py.b_h[j], g.py.k_h[j]);
        if (s.zs >= 0) dsyz = cpml_apply(dsyz, g.mem_z + 1 * s.nzs + s.iz, g.pz.a[k], g.pz.b[k], g.pz.k[k]);
        g.vy[p] += (dsxy + dsyy + dsyz) * dt / rho_half_y;
    }
    // vz at (i, j, k+1/2): :197-212 (Tensor :228-243)
    if (k < nz - N - 1 && i >= N + 1 && j >= N + 1) {
        const float rho_avg = 0.25f * (g.rho[p] + g.rho[p + SY] + g.rho[p + 1 + SY] + g.rho[p + 1]);
        float dsxz = d_bm<N>(g.sxz, p, 1, c) / dx;
        float dsyz = d_bm<N>(g.syz, p, SY, c) / dy;
        float dszz = d_fm<N>(g.szz, p, SZ, c) / dz;
        if (s.xs >= 0) dsxz = cpml_apply(dsxz, g.mem_x + 2 * s.nxs + s.ix, g.px.a[i], g.px.b[i], g.px.k[i]);
        if (s.ys >= 0) dsyz = cpml_apply(dsyz, g.mem_y + 2 * s.nys + s.iy, g.py.a[j], g.py.b[j], g.py.k[j]);
        if (s.zs >= 0) dszz = cpml_apply(dszz, g.mem_z + 2 * s.nzs + s.iz, g.pz.a_h[k], g.pz.b_h[k], g.pz.k_h[k]);
        if (CLS == SW_CLASS_TENSOR)
            g.vz[p] += (dsxz + dsyz + dszz) * dt / rho_avg;
        else
            g.vz[p] += (dszz + dsxz + dsyz) * dt / rho_avg;
    }
}

template <int N>
__global__ void __launch_bounds__(256) k_stress(sw_ssg3d g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + N;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + N;
    const int k = blockIdx.z + N;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= nx - N || j >= ny - N || k >= nz - N) return;
    const long SY = nx, SZ = (long)nx * ny;
    const long p = k * SZ + j * SY + i;
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    const float *c = g.c;
    const Slabs s = slabs_of(g, i, j, k);

    // normal stresses at (i, j, k): src/Elastic3D_SSG_CPML.py:247-266
    if (i >= N + 1 && j >= N + 1 && k >= N + 1) {
        const float lam_h = 0.5f * (g.lam[p] + g.lam[p + 1]);
        const float mu_h = 0.5f * (g.mu[p] + g.mu[p + 1]);
        const float lpm = lam_h + 2.f * mu_h;
        float dvxx = d_bm<N>(g.vx, p, 1, c) / dx;
        float dvyy = d_bm<N>(g.vy, p, SY, c) / dy;
        float dvzz = d_bm<N>(g.vz, p, SZ, c) / dz;
        if (s.xs >= 0) dvxx = cpml_apply(dvxx, g.mem_x + 3 * s.nxs + s.ix, g.px.a[i], g.px.b[i], g.px.k[i]);
        if (s.ys >= 0) dvyy = cpml_apply(dvyy, g.mem_y + 3 * s.nys + s.iy, g.py.a[j], g.py.b[j], g.py.k[j]);
        if (s.zs >= 0) dvzz = cpml_apply(dvzz, g.mem_z + 3 * s.nzs + s.iz, g.pz.a[k], g.pz.b[k], g.pz.k[k]);
        g.sxx[p] += (lpm * dvxx + lam_h * (dvzz + dvyy)) * dt;
        g.syy[p] += (lpm * dvyy + lam_h * (dvxx + dvzz)) * dt;
        g.szz[p] += (lpm * dvzz + lam_h * (dvxx + dvyy)) * dt;
    }
    // sxy at (i+1/2, j+1/2, k): :267-279
    if (i >= N + 1 && j >= N + 1) {
        const float mu_a = 0.25f * (g.mu[p + SY] + g.mu[p] + g.mu[p] + g.mu[p + 1]);
        float dvyx = d_fm<N>(g.vy, p, 1, c) / dx;
        float dvxy = d_fm<N>(g.vx, p, SY, c) / dy;
        if (s.xs >= 0) dvyx = cpml_apply(dvyx, g.mem_x + 4 * s.nxs + s.ix, g.px.a_h[i], g.px.b_h[i], g.px.k_h[i]);
        if (s.ys >= 0) dvxy = cpml_apply(dvxy, g.mem_y + 4 * s.nys + s.iy, g.py.a_h[j], g.py.b_h[j], g.py.k_h[j]);
        g.sxy[p] += mu_a * (dvxy + dvyx) * dt;
    }
    // sxz at (i+1/2, j, k+1/2): :281-293
    if (i < nx - N - 1 && k < nz - N - 1) {
        const float mu_a = 0.25f * (g.mu[p] + g.mu[p + 1] + g.mu[p] + g.mu[p + SY]);
        float dvzx = d_fm<N>(g.vz, p, 1, c) / dx;
        float dvxz = d_fm<N>(g.vx, p, SZ, c) / dz;
        if (s.xs >= 0) dvzx = cpml_apply(dvzx, g.mem_x + 5 * s.nxs + s.ix, g.px.a_h[i], g.px.b_h[i], g.px.k_h[i]);
        if (s.zs >= 0) dvxz = cpml_apply(dvxz, g.mem_z + 4 * s.nzs + s.iz, g.pz.a_h[k], g.pz.b_h[k], g.pz.k_h[k]);
        g.sxz[p] += mu_a * (dvxz + dvzx) * dt;
    }
    // syz at (i, j+1/2, k+1/2): :295-307
    if (j < ny - N - 1 && k < nz - N - 1) {
        const float mu_a = 0.25f * (g.mu[p + SY] + g.mu[p] + g.mu[p + SZ] + g.mu[p]);
        float dvyz = d_fm<N>(g.vy, p, SZ, c) / dz;
        float dvzy = d_fm<N>(g.vz, p, SY, c) / dy;
        if (s.ys >= 0) dvzy = cpml_apply(dvzy, g.mem_y + 5 * s.nys + s.iy, g.py.a_h[j], g.py.b_h[j], g.py.k_h[j]);
        if (s.zs >= 0) dvyz = cpml_apply(dvyz, g.mem_z + 5 * s.nzs + s.iz, g.pz.a_h[k], g.pz.b_h[k], g.pz.k_h[k]);
        g.syz[p] += mu_a * (dvyz + dvzy) * dt;
    }
}

inline dim3 box_grid(const sw_ssg3d &g, int N, dim3 blk)
{
    const int ex = g.nx - 2 * N, ey = g.ny - 2 * N, ez = g.nz - 2 * N;
    return dim3((ex + blk.x - 1) / blk.x, (ey + blk.y - 1) / blk.y, ez);
}

}  // namespace

int launch_ssg3d_velocity(const sw_ssg3d *g, hipStream_t st)
{
    const dim3 blk(64, 4, 1);
    SW_DISPATCH_ORDER(g->order, {
        const dim3 grd = box_grid(*g, N, blk);
        if (g->cls == SW_CLASS_TENSOR)
            hipLaunchKernelGGL((k_velocity<N, SW_CLASS_TENSOR>), grd, blk, 0, st, *g);
        else
            hipLaunchKernelGGL((k_velocity<N, SW_CLASS_CPML>), grd, blk, 0, st, *g);
    });
    return launch_status("ssg3d velocity");
}

int launch_ssg3d_stress(const sw_ssg3d *g, hipStream_t st)
{
    const dim3 blk(64, 4, 1);
    SW_DISPATCH_ORDER(g->order, {
        const dim3 grd = box_grid(*g, N, blk);
        hipLaunchKernelGGL((k_stress<N>), grd, blk, 0, st, *g);
    });
    return launch_status("ssg3d stress");
}

}  // namespace sw
