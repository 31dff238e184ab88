// 3-D staggered-grid velocity/stress sweeps with slab-confined CPML.
//
// Reference: src/Elastic3D_SSG_CPML.py:144-309 (update_SSG) and
// src/Elastic3D_SSG_Tensor.py:166-345 (same maths, other vz term order).
//
// Execution: each thread owns one (i, j) column of an x-y tile (x fastest,
// 64 lanes = one wavefront row) and marches along z over a chunk of planes.
// The z-differentiated fields live in per-thread register windows of 2N
// values that slide with k, so each z plane of those fields is read from
// memory once per column; x and y taps are plain loads that hit the L1/L2
// lines the neighbouring lanes and waves of the tile already fetched.  Blocks
// are renumbered so that the tiles of one x-y neighbourhood run on the same
// XCD (MI355X dispatches consecutive workgroups round-robin over 8 XCDs, each
// with its own L2).  The z-slab of the CPML is a wave-uniform branch; the x
// and y slabs are per-lane.
#include "sw_internal.h"

namespace sw {

namespace {

constexpr int kXCD = 8;  // XCDs per MI355X; any value keeps the renumbering a bijection

// Physical workgroup b runs on XCD b % kXCD.  Give each XCD a contiguous range
// of logical tiles so halo lines are shared within one L2.
__device__ __forceinline__ int xcd_logical_block(int b, int nb)
{
    const int q = nb / kXCD, r = nb % kXCD;
    const int x = b % kXCD, s = b / kXCD;
    return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

struct Geo {
    int nbx, nby, nbz, cz;
};

struct Cell {
    int i, j, k0, k1;
};

template <int TX, int TY>
__device__ __forceinline__ Cell tile_cell(const sw_ssg3d &g, const Geo geo, int N)
{
    const int b = xcd_logical_block(blockIdx.x, gridDim.x);
    const int bx = b % geo.nbx;
    const int r = b / geo.nbx;
    const int by = r % geo.nby;
    const int bz = r / geo.nby;
    Cell c;
    c.i = bx * TX + threadIdx.x;
    c.j = by * TY + threadIdx.y;
    c.k0 = N + bz * geo.cz;
    c.k1 = min(c.k0 + geo.cz, g.nz - N);
    return c;
}

// z taps from a window w[t] = f[k - N + t] (t = 0..2N-1): backward difference
template <int N>
__device__ __forceinline__ float win_bm(const float *w, const float *c)
{
    float v = 0.f;
#pragma unroll
    for (int id = 0; id < N; ++id) v += c[id] * (w[N + id] - w[N - id - 1]);
    return v;
}

// forward difference from a window w[t] = f[k - N + 1 + t]
template <int N>
__device__ __forceinline__ float win_fm(const float *w, const float *c)
{
    float v = 0.f;
#pragma unroll
    for (int id = 0; id < N; ++id) v += c[id] * (w[N + id] - w[N - 1 - id]);
    return v;
}

template <int N>
__device__ __forceinline__ void win_shift(float *w)
{
#pragma unroll
    for (int t = 0; t < 2 * N - 1; ++t) w[t] = w[t + 1];
}

struct SlabIdx {
    int xs, ys;        // slab index along x / y (-1 outside)
    long rx, ry;       // (k*ny + j)*wx + xs at k = 0, and (k*wy + ys)*nx + i at k = 0
    long sx, sy, sz;   // per-k stride of the x / y slab arrays, z-slab row pitch
    long nxs, nys, nzs;
};

__device__ __forceinline__ SlabIdx slab_setup(const sw_ssg3d &g, int i, int j)
{
    SlabIdx s;
    const int wx = g.px.lo + (g.nx - g.px.hi);
    const int wy = g.py.lo + (g.ny - g.py.hi);
    const int wz = g.pz.lo + (g.nz - g.pz.hi);
    s.xs = slab_index(i, g.px.lo, g.px.hi);
    s.ys = slab_index(j, g.py.lo, g.py.hi);
    s.rx = (long)j * wx + s.xs;
    s.ry = (long)s.ys * g.nx + i;
    s.sx = (long)g.ny * wx;
    s.sy = (long)wy * g.nx;
    s.sz = (long)g.ny * g.nx;
    s.nxs = (long)g.nz * g.ny * wx;
    s.nys = (long)g.nz * wy * g.nx;
    s.nzs = (long)wz * g.ny * g.nx;
    return s;
}

template <int N, int CLS, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) k_velocity(sw_ssg3d g, Geo geo)
{
    const Cell cl = tile_cell<TX, TY>(g, geo, N);
    const int i = cl.i, j = cl.j;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i < N || i >= nx - N || j < N || j >= ny - N) return;
    const long SY = nx, SZ = (long)nx * ny;
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    float c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = g.c[t];
    const bool vx_col = i < nx - N - 1 && j >= N + 1;
    const bool vy_col = j < ny - N - 1 && i >= N + 1;
    const bool vz_col = i >= N + 1 && j >= N + 1;
    const SlabIdx sl = slab_setup(g, i, j);
    // per-column CPML coefficients along x and y
    float ax_h = 0.f, bx_h = 1.f, kx_h = 1.f, ax = 0.f, bx = 1.f, kx = 1.f;
    float ay_h = 0.f, by_h = 1.f, ky_h = 1.f, ay = 0.f, by = 1.f, ky = 1.f;
    if (sl.xs >= 0) {
        ax_h = g.px.a_h[i]; bx_h = g.px.b_h[i]; kx_h = g.px.k_h[i];
        ax = g.px.a[i]; bx = g.px.b[i]; kx = g.px.k[i];
    }
    if (sl.ys >= 0) {
        ay_h = g.py.a_h[j]; by_h = g.py.b_h[j]; ky_h = g.py.k_h[j];
        ay = g.py.a[j]; by = g.py.b[j]; ky = g.py.k[j];
    }
    const long pc = (long)j * SY + i;  // column base
    // windows: wxz[t] = sxz[k-N+t], wyz[t] = syz[k-N+t], wzz[t] = szz[k-N+1+t]
    float wxz[2 * N], wyz[2 * N], wzz[2 * N];
    const int k0 = cl.k0;
#pragma unroll
    for (int t = 0; t < 2 * N - 1; ++t) {
        wxz[t] = g.sxz[pc + (long)(k0 - N + t) * SZ];
        wyz[t] = g.syz[pc + (long)(k0 - N + t) * SZ];
        wzz[t] = g.szz[pc + (long)(k0 - N + 1 + t) * SZ];
    }
    for (int k = k0; k < cl.k1; ++k) {
        const long p = pc + (long)k * SZ;
        wxz[2 * N - 1] = g.sxz[p + (long)(N - 1) * SZ];
        wyz[2 * N - 1] = g.syz[p + (long)(N - 1) * SZ];
        wzz[2 * N - 1] = g.szz[p + (long)N * SZ];
        const int zs = slab_index(k, g.pz.lo, g.pz.hi);  // wave-uniform
        // vx at (i+1/2, j, k): src/Elastic3D_SSG_CPML.py:164-178
        if (vx_col && k >= N + 1) {
            float dsxx = d_fm<N>(g.sxx, p, 1, c) / dx;
            float dsxy = d_bm<N>(g.sxy, p, SY, c) / dy;
            float dsxz = win_bm<N>(wxz, c) / dz;
            if (sl.xs >= 0) dsxx = cpml_apply(dsxx, g.mem_x + 0 * sl.nxs + sl.rx + k * sl.sx, ax_h, bx_h, kx_h);
            if (sl.ys >= 0) dsxy = cpml_apply(dsxy, g.mem_y + 0 * sl.nys + sl.ry + k * sl.sy, ay, by, ky);
            if (zs >= 0)
                dsxz = cpml_apply(dsxz, g.mem_z + 0 * sl.nzs + zs * sl.sz + pc, g.pz.a[k], g.pz.b[k], g.pz.k[k]);
            g.vx[p] += (dsxx + dsxz + dsxy) * dt / g.rho[p];
        }
        // vy at (i, j+1/2, k): :180-195
        if (vy_col && k >= N + 1) {
            const float rho_half_y = 0.5f * (g.rho[p] + g.rho[p + SY]);
            float dsxy = d_bm<N>(g.sxy, p, 1, c) / dx;
            float dsyy = d_fm<N>(g.syy, p, SY, c) / dy;
            float dsyz = win_bm<N>(wyz, c) / dz;
            if (sl.xs >= 0) dsxy = cpml_apply(dsxy, g.mem_x + 1 * sl.nxs + sl.rx + k * sl.sx, ax, bx, kx);
            if (sl.ys >= 0) dsyy = cpml_apply(dsyy, g.mem_y + 1 * sl.nys + sl.ry + k * sl.sy, ay_h, by_h, ky_h);
            if (zs >= 0)
                dsyz = cpml_apply(dsyz, g.mem_z + 1 * sl.nzs + zs * sl.sz + pc, g.pz.a[k], g.pz.b[k], g.pz.k[k]);
            g.vy[p] += (dsxy + dsyy + dsyz) * dt / rho_half_y;
        }
        // vz at (i, j, k+1/2): :197-212 (Tensor :228-243)
        if (vz_col && k < nz - N - 1) {
            const float rho_avg = 0.25f * (g.rho[p] + g.rho[p + SY] + g.rho[p + 1 + SY] + g.rho[p + 1]);
            float dsxz = d_bm<N>(g.sxz, p, 1, c) / dx;
            float dsyz = d_bm<N>(g.syz, p, SY, c) / dy;
            float dszz = win_fm<N>(wzz, c) / dz;
            if (sl.xs >= 0) dsxz = cpml_apply(dsxz, g.mem_x + 2 * sl.nxs + sl.rx + k * sl.sx, ax, bx, kx);
            if (sl.ys >= 0) dsyz = cpml_apply(dsyz, g.mem_y + 2 * sl.nys + sl.ry + k * sl.sy, ay, by, ky);
            if (zs >= 0)
                dszz = cpml_apply(dszz, g.mem_z + 2 * sl.nzs + zs * sl.sz + pc, g.pz.a_h[k], g.pz.b_h[k],
                                  g.pz.k_h[k]);
            if (CLS == SW_CLASS_TENSOR)
                g.vz[p] += (dsxz + dsyz + dszz) * dt / rho_avg;
            else
                g.vz[p] += (dszz + dsxz + dsyz) * dt / rho_avg;
        }
        win_shift<N>(wxz);
        win_shift<N>(wyz);
        win_shift<N>(wzz);
    }
}

template <int N, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) k_stress(sw_ssg3d g, Geo geo)
{
    const Cell cl = tile_cell<TX, TY>(g, geo, N);
    const int i = cl.i, j = cl.j;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i < N || i >= nx - N || j < N || j >= ny - N) return;
    const long SY = nx, SZ = (long)nx * ny;
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    float c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = g.c[t];
    const bool nrm_col = i >= N + 1 && j >= N + 1;
    const bool sxy_col = i >= N + 1 && j >= N + 1;
    const bool sxz_col = i < nx - N - 1;
    const bool syz_col = j < ny - N - 1;
    const SlabIdx sl = slab_setup(g, i, j);
    float ax_h = 0.f, bx_h = 1.f, kx_h = 1.f, ax = 0.f, bx = 1.f, kx = 1.f;
    float ay_h = 0.f, by_h = 1.f, ky_h = 1.f, ay = 0.f, by = 1.f, ky = 1.f;
    if (sl.xs >= 0) {
        ax_h = g.px.a_h[i]; bx_h = g.px.b_h[i]; kx_h = g.px.k_h[i];
        ax = g.px.a[i]; bx = g.px.b[i]; kx = g.px.k[i];
    }
    if (sl.ys >= 0) {
        ay_h = g.py.a_h[j]; by_h = g.py.b_h[j]; ky_h = g.py.k_h[j];
        ay = g.py.a[j]; by = g.py.b[j]; ky = g.py.k[j];
    }
    const long pc = (long)j * SY + i;
    // windows: wvz[t] = vz[k-N+t] (backward), wvx[t] = vx[k-N+1+t], wvy[t] = vy[k-N+1+t] (forward)
    float wvz[2 * N], wvx[2 * N], wvy[2 * N];
    const int k0 = cl.k0;
#pragma unroll
    for (int t = 0; t < 2 * N - 1; ++t) {
        wvz[t] = g.vz[pc + (long)(k0 - N + t) * SZ];
        wvx[t] = g.vx[pc + (long)(k0 - N + 1 + t) * SZ];
        wvy[t] = g.vy[pc + (long)(k0 - N + 1 + t) * SZ];
    }
    for (int k = k0; k < cl.k1; ++k) {
        const long p = pc + (long)k * SZ;
        wvz[2 * N - 1] = g.vz[p + (long)(N - 1) * SZ];
        wvx[2 * N - 1] = g.vx[p + (long)N * SZ];
        wvy[2 * N - 1] = g.vy[p + (long)N * SZ];
        const int zs = slab_index(k, g.pz.lo, g.pz.hi);
        // normal stresses at (i, j, k): src/Elastic3D_SSG_CPML.py:247-266
        if (nrm_col && k >= N + 1) {
            const float lam_h = 0.5f * (g.lam[p] + g.lam[p + 1]);
            const float mu_h = 0.5f * (g.mu[p] + g.mu[p + 1]);
            const float lpm = lam_h + 2.f * mu_h;
            float dvxx = d_bm<N>(g.vx, p, 1, c) / dx;
            float dvyy = d_bm<N>(g.vy, p, SY, c) / dy;
            float dvzz = win_bm<N>(wvz, c) / dz;
            if (sl.xs >= 0) dvxx = cpml_apply(dvxx, g.mem_x + 3 * sl.nxs + sl.rx + k * sl.sx, ax, bx, kx);
            if (sl.ys >= 0) dvyy = cpml_apply(dvyy, g.mem_y + 3 * sl.nys + sl.ry + k * sl.sy, ay, by, ky);
            if (zs >= 0)
                dvzz = cpml_apply(dvzz, g.mem_z + 3 * sl.nzs + zs * sl.sz + pc, g.pz.a[k], g.pz.b[k], g.pz.k[k]);
            g.sxx[p] += (lpm * dvxx + lam_h * (dvzz + dvyy)) * dt;
            g.syy[p] += (lpm * dvyy + lam_h * (dvxx + dvzz)) * dt;
            g.szz[p] += (lpm * dvzz + lam_h * (dvxx + dvyy)) * dt;
        }
        // sxy at (i+1/2, j+1/2, k): :267-279
        if (sxy_col) {
            const float mu_a = 0.25f * (g.mu[p + SY] + g.mu[p] + g.mu[p] + g.mu[p + 1]);
            float dvyx = d_fm<N>(g.vy, p, 1, c) / dx;
            float dvxy = d_fm<N>(g.vx, p, SY, c) / dy;
            if (sl.xs >= 0) dvyx = cpml_apply(dvyx, g.mem_x + 4 * sl.nxs + sl.rx + k * sl.sx, ax_h, bx_h, kx_h);
            if (sl.ys >= 0) dvxy = cpml_apply(dvxy, g.mem_y + 4 * sl.nys + sl.ry + k * sl.sy, ay_h, by_h, ky_h);
            g.sxy[p] += mu_a * (dvxy + dvyx) * dt;
        }
        // sxz at (i+1/2, j, k+1/2): :281-293
        if (sxz_col && k < nz - N - 1) {
            const float mu_a = 0.25f * (g.mu[p] + g.mu[p + 1] + g.mu[p] + g.mu[p + SY]);
            float dvzx = d_fm<N>(g.vz, p, 1, c) / dx;
            float dvxz = win_fm<N>(wvx, c) / dz;
            if (sl.xs >= 0) dvzx = cpml_apply(dvzx, g.mem_x + 5 * sl.nxs + sl.rx + k * sl.sx, ax_h, bx_h, kx_h);
            if (zs >= 0)
                dvxz = cpml_apply(dvxz, g.mem_z + 4 * sl.nzs + zs * sl.sz + pc, g.pz.a_h[k], g.pz.b_h[k],
                                  g.pz.k_h[k]);
            g.sxz[p] += mu_a * (dvxz + dvzx) * dt;
        }
        // syz at (i, j+1/2, k+1/2): :295-307
        if (syz_col && k < nz - N - 1) {
            const float mu_a = 0.25f * (g.mu[p + SY] + g.mu[p] + g.mu[p + SZ] + g.mu[p]);
            float dvyz = win_fm<N>(wvy, c) / dz;
            float dvzy = d_fm<N>(g.vz, p, SY, c) / dy;
            if (sl.ys >= 0) dvzy = cpml_apply(dvzy, g.mem_y + 5 * sl.nys + sl.ry + k * sl.sy, ay_h, by_h, ky_h);
            if (zs >= 0)
                dvyz = cpml_apply(dvyz, g.mem_z + 5 * sl.nzs + zs * sl.sz + pc, g.pz.a_h[k], g.pz.b_h[k],
                                  g.pz.k_h[k]);
            g.syz[p] += mu_a * (dvyz + dvzy) * dt;
        }
        win_shift<N>(wvz);
        win_shift<N>(wvx);
        win_shift<N>(wvy);
    }
}

constexpr int kTX = 64, kTY = 4;

Geo make_geo(const sw_ssg3d &g, int N, int cz)
{
    Geo geo;
    geo.cz = cz;
    geo.nbx = (g.nx + kTX - 1) / kTX;
    geo.nby = (g.ny + kTY - 1) / kTY;
    geo.nbz = (g.nz - 2 * N + cz - 1) / cz;
    return geo;
}

// z planes per block: enough blocks to fill 256 CUs several times over while
// keeping the window warm-up (2N-1 planes) a small fraction of the chunk.
int choose_cz(const sw_ssg3d &g, int N)
{
    const int cols = ((g.nx + kTX - 1) / kTX) * ((g.ny + kTY - 1) / kTY);
    const int nzi = g.nz - 2 * N;
    int cz = 16;
    while (cz < 64 && (long)cols * ((nzi + 2 * cz - 1) / (2 * cz)) >= 4096) cz *= 2;
    return cz > nzi ? nzi : cz;
}

}  // namespace

int launch_ssg3d_velocity(const sw_ssg3d *g, hipStream_t st)
{
    const dim3 blk(kTX, kTY, 1);
    SW_DISPATCH_ORDER(g->order, {
        const Geo geo = make_geo(*g, N, choose_cz(*g, N));
        const dim3 grd(geo.nbx * geo.nby * geo.nbz);
        if (g->cls == SW_CLASS_TENSOR)
            hipLaunchKernelGGL((k_velocity<N, SW_CLASS_TENSOR, kTX, kTY>), grd, blk, 0, st, *g, geo);
        else
            hipLaunchKernelGGL((k_velocity<N, SW_CLASS_CPML, kTX, kTY>), grd, blk, 0, st, *g, geo);
    });
    return launch_status("ssg3d velocity");
}

int launch_ssg3d_stress(const sw_ssg3d *g, hipStream_t st)
{
    const dim3 blk(kTX, kTY, 1);
    SW_DISPATCH_ORDER(g->order, {
        const Geo geo = make_geo(*g, N, choose_cz(*g, N));
        const dim3 grd(geo.nbx * geo.nby * geo.nbz);
        hipLaunchKernelGGL((k_stress<N, kTX, kTY>), grd, blk, 0, st, *g, geo);
    });
    return launch_status("ssg3d stress");
}

}  // namespace sw
