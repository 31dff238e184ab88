// 3-D staggered-grid velocity/stress sweeps with slab-confined CPML.
//
// Reference: src/Elastic3D_SSG_CPML.py:144-309 (update_SSG),
// src/Elastic3D_SSG_Tensor.py:166-345 (same maths, other vz term order) and
// src/Elastic3D_SSG_SPML.py:160-393 (split-field PML: plain updates with
// 1/rho and point lambda, mu in its split-free interior; the PML shell runs
// the kernels of ssg3d_split.hip).
//
// Execution: each thread owns one (i, j) column and marches along z over a
// chunk of planes; a block takes 256 consecutive columns of a box in x-fastest
// order, its waves starting on 128-byte lines.  On large grids a sweep is
// five launches over boxes of full-width rows, instantiated with the CPML of
// only the axes whose slabs they meet (x; x and y; x, y and z), so the bulk
// of the grid runs with fewer registers and more waves in flight; on grids
// where the CPML shell is a large part, or whose slab launches could not
// fill the GPU, one launch with all three axes.
//  * The fields a sweep updates in place are loaded and stored
//    non-temporally (read and written once per sweep by their owner).
//  * The z-differentiated fields live in per-thread register windows of 2N
//    values that slide with k, so each of their planes is read once.
//  * Every value a plane needs (y stencil taps, material, the
//    read-modify-write fields, CPML memory variables and coefficients) is
//    gathered into a register "plane" before any arithmetic, and the next
//    plane is fetched while the current one is computed: each wave keeps
//    ~25 independent loads in flight instead of waiting on one branch at a
//    time.  Every lane, idle ones included, loads at an in-bounds column, so
//    the loads need no per-component predicates.
//  * x taps, and the material at p+1, come from the neighbouring lanes (DPP
//    wave shifts), the two wave-edge values from scalar loads; the y taps'
//    centres come from the z windows.  Both sweeps then fit 4 waves per SIMD
//    at 200^3 (120 / 124 VGPRs against 137 / 147 with 16-byte x-tap loads):
//    measured on MI355X in one process, bit-identical, 200^3 0.2855 ->
//    0.2734 ms per step, 600x600x400 4.342 -> 4.125 ms.  The same shifts with
//    the edge values from vector loads ran 0.298-0.343 ms: a wait for those
//    loads is a wait for the whole prefetched plane (vmcnt counts in order,
//    and the plane's lane-predicated slab loads leave the compiler no count
//    but 0), and each adds a vector-memory instruction.
//  * Blocks are renumbered so that the blocks of one neighbourhood run on the
//    same XCD (MI355X dispatches consecutive workgroups round-robin over 8
//    XCDs, each with its own L2) and halo lines are shared within one L2.
//  * The z slab of the CPML is a wave-uniform condition; x and y slabs are
//    per lane.
// Arithmetic is term for term the reference's (build with -ffp-contract=off).
#include <algorithm>

#include "ssg3d_common.h"

namespace sw {

namespace {

// Columns are numbered by their flat index q = j*nx + i within a plane,
// from f0 (the box's first column rounded down to a multiple of 32) to fn:
// a wave's 64 columns then start on a 128-byte line of every array (the
// planes of a grid with nx*ny a multiple of 32 all start on a line), so a
// row of 64 floats costs 2 line requests instead of 3.  Columns outside the
// box in that range (the x edges of each row) are idle lanes.
struct Box {
    int i0, j0, w, h, k0, k1, ncb;  // ncb: column blocks of the box
    int f0, fn, nx;                 // flat column range [f0, fn)
};

struct Geo {
    Box box;
    int nblock;
};

// i, j: the lane's cell column, -1 on idle lanes; qi, qj: the column of the
// lane's flat index, idle lanes included (they keep loading for their
// neighbours' x taps).
struct Cell {
    int i, j, qi, qj, k0, k1;
};

__device__ __forceinline__ Cell box_cell(const Geo &geo)
{
    const int b = xcd_logical_block(blockIdx.x, geo.nblock);
    const Box &B = geo.box;
    // logical order: groups of kGroup column blocks, each group's z chunks
    // bottom to top, its column blocks fastest within a chunk (the last
    // group may be narrower)
    const int nzc = (B.k1 - B.k0 + kZChunk - 1) / kZChunk;
    const int gi = b / (kGroup * nzc), r = b - gi * kGroup * nzc;
    const int gw = min(kGroup, B.ncb - gi * kGroup);
    const int zb = r / gw, cb = gi * kGroup + r - zb * gw;
    const int q = B.f0 + cb * kThreads + (int)threadIdx.x;
    const int j = q / B.nx, i = q - j * B.nx;
    Cell c;
    c.qi = i;
    c.qj = j;
    if (q >= B.fn || j < B.j0 || j >= B.j0 + B.h || i < B.i0 || i >= B.i0 + B.w) {
        c.i = -1; c.j = -1;
    } else {
        c.i = i;
        c.j = j;
    }
    c.k0 = B.k0 + zb * kZChunk;
    c.k1 = min(c.k0 + kZChunk, B.k1);
    return c;
}


// LARGE: arrays above 4 GiB; the block rebases its full-grid pointers to its
// first plane (otherwise kb is 0 and the rebasing compiles away, which keeps
// the scalar registers of the common case free).
template <int N, int CLS, int PM, bool LARGE>
__global__ void __launch_bounds__(kThreads) k_velocity(sw_ssg3d g, Geo geo, Fold3D fo)
{
    const Cell cl = box_cell(geo);
    const int i = cl.i, j = cl.j;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // Idle lanes run the loop without storing: they load at their own
    // column (its row moved into [N, ny-N) if outside, where no updated
    // cell uses it), so every lane can serve its neighbours' x taps.
    const bool live = i >= 0;
    const int li = cl.qi, lj = min(max(cl.qj, N), ny - N - 1);
    const int SY = 4 * nx;
    const uint32_t SZ = 4u * (uint32_t)(nx * ny);
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    float c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = g.c[t];
    const bool vx_col = live && i < nx - N - 1 && j >= N + 1;
    const bool vy_col = live && j < ny - N - 1 && i >= N + 1;
    const bool vz_col = live && i >= N + 1 && j >= N + 1;
    // CPML axes of this instance (none for the split-field class, whose
    // kernels here run only its split-free interior)
    constexpr int CPM = CLS == SW_CLASS_SPML ? 0 : PM;
    Slab sl = slab_setup<CPM>(g, li, lj);
    if (!live) sl.xs = sl.ys = -1;
    const int kb = LARGE ? cl.k0 - N : 0;  // base plane of this block's offsets
    const Res R = make_res(g, kb);
    const Coef cx = coef_of(g.px, li, sl.xs >= 0);
    const Coef cy = coef_of(g.py, lj, sl.ys >= 0);
    const uint32_t pc = 4u * (uint32_t)(lj * nx + li);
    // windows: wxz[t] = sxz[k-N+t], wyz[t] = syz[k-N+t], wzz[t] = szz[k-N+1+t]
    float wxz[2 * N], wyz[2 * N], wzz[2 * N];
    const int k0 = cl.k0;
#pragma unroll
    for (int t = 0; t < 2 * N - 1; ++t) {
        wxz[t] = ld(R.sxz, pc + (uint32_t)(k0 - N + t - kb) * SZ);
        wyz[t] = ld(R.syz, pc + (uint32_t)(k0 - N + t - kb) * SZ);
        wzz[t] = ld(R.szz, pc + (uint32_t)(k0 - N + 1 + t - kb) * SZ);
    }
    // the register plane is double-buffered: plane k+1 is fetched while k is computed
    VPlane<N> P;
    load_vplane<N, CPM>(P, g, R, pc, k0, kb, SY, SZ, sl);
    const Edge e0 = wave_edge(pc);  // lane 0's and lane 63's plane offsets
    for (int k = k0; k < cl.k1; ++k) {
        const uint32_t p = pc + (uint32_t)(k - kb) * SZ;
        const Edge e{e0.lo + (uint32_t)(k - kb) * SZ, e0.hi + (uint32_t)(k - kb) * SZ};
        wxz[2 * N - 1] = P.wxz;
        wyz[2 * N - 1] = P.wyz;
        wzz[2 * N - 1] = P.wzz;
        VPlane<N> Q;
        if (k + 1 < cl.k1) load_vplane<N, CPM>(Q, g, R, pc, k + 1, kb, SY, SZ, sl);
        // x taps and the p+1 densities from the neighbouring lanes (all lanes active)
        P.syz_y[N] = wyz[N];  // syz at p, from the window
        float sxx_x[2 * N], sxy_x[2 * N], sxz_x[2 * N];
        xtaps_fm_lanes<N>(sxx_x, P.sxx, R.sxx, e);
        xtaps_bm_lanes<N>(sxy_x, P.sxy_y[N], R.sxy, e);
        xtaps_bm_lanes<N>(sxz_x, wxz[N], R.sxz, e);
        const float rho_x = next_x(P.rho[0], R.rho, e);                           // p+1
        const float rho_xy = next_x(P.rho[1], R.rho, Edge{e.lo + SY, e.hi + SY});  // p+1+SY
        const uint32_t ox = sl.rx + (uint32_t)k * sl.sx, oy = sl.ry + (uint32_t)k * sl.sy;
        const uint32_t oz = (uint32_t)P.zs * sl.sz + pc;
        // vx at (i+1/2, j, k): src/Elastic3D_SSG_CPML.py:164-178
        if (vx_col && k >= N + 1) {
            float dsxx = diff_fm<N>(sxx_x, c) / dx;
            float dsxy = diff_bm<N>(P.sxy_y, c) / dy;
            float dsxz = diff_bm<N>(wxz, c) / dz;
            if (sl.xs >= 0) {
                dsxx = cpml_step(dsxx, P.mx[0], cx.ah, cx.bh, cx.kh);
                st(R.mx, 0 * sl.nxs + ox, P.mx[0]);
            }
            if (sl.ys >= 0) {
                dsxy = cpml_step(dsxy, P.my[0], cy.a, cy.b, cy.k);
                st(R.my, 0 * sl.nys + oy, P.my[0]);
            }
            if (P.zs >= 0) {
                dsxz = cpml_step(dsxz, P.mz[0], P.cz.a, P.cz.b, P.cz.k);
                st(R.mz, 0 * sl.nzs + oz, P.mz[0]);
            }
            if (CLS == SW_CLASS_SPML)  // src/Elastic3D_SSG_SPML.py:203: x+y+z, times 1/rho
                st_nt(R.vx, p, P.v[0] + (dsxx + dsxy + dsxz) * dt * (1.f / P.rho[0]));
            else
                st_nt(R.vx, p, P.v[0] + (dsxx + dsxz + dsxy) * dt / P.rho[0]);
        }
        // vy at (i, j+1/2, k): :180-195
        if (vy_col && k >= N + 1) {
            const float rho_half_y = 0.5f * (P.rho[0] + P.rho[1]);
            float dsxy = diff_bm<N>(sxy_x, c) / dx;
            float dsyy = diff_fm<N>(P.syy, c) / dy;
            float dsyz = diff_bm<N>(wyz, c) / dz;
            if (sl.xs >= 0) {
                dsxy = cpml_step(dsxy, P.mx[1], cx.a, cx.b, cx.k);
                st(R.mx, 1 * sl.nxs + ox, P.mx[1]);
            }
            if (sl.ys >= 0) {
                dsyy = cpml_step(dsyy, P.my[1], cy.ah, cy.bh, cy.kh);
                st(R.my, 1 * sl.nys + oy, P.my[1]);
            }
            if (P.zs >= 0) {
                dsyz = cpml_step(dsyz, P.mz[1], P.cz.a, P.cz.b, P.cz.k);
                st(R.mz, 1 * sl.nzs + oz, P.mz[1]);
            }
            if (CLS == SW_CLASS_SPML)  // :231
                st_nt(R.vy, p, P.v[1] + (dsxy + dsyy + dsyz) * dt * (1.f / rho_half_y));
            else
                st_nt(R.vy, p, P.v[1] + (dsxy + dsyy + dsyz) * dt / rho_half_y);
        }
        // vz at (i, j, k+1/2): :197-212 (Tensor :228-243)
        if (vz_col && k < nz - N - 1) {
            const float rho_avg = 0.25f * (P.rho[0] + P.rho[1] + rho_xy + rho_x);
            float dsxz = diff_bm<N>(sxz_x, c) / dx;
            float dsyz = diff_bm<N>(P.syz_y, c) / dy;
            float dszz = diff_fm<N>(wzz, c) / dz;
            if (sl.xs >= 0) {
                dsxz = cpml_step(dsxz, P.mx[2], cx.a, cx.b, cx.k);
                st(R.mx, 2 * sl.nxs + ox, P.mx[2]);
            }
            if (sl.ys >= 0) {
                dsyz = cpml_step(dsyz, P.my[2], cy.a, cy.b, cy.k);
                st(R.my, 2 * sl.nys + oy, P.my[2]);
            }
            if (P.zs >= 0) {
                dszz = cpml_step(dszz, P.mz[2], P.cz.ah, P.cz.bh, P.cz.kh);
                st(R.mz, 2 * sl.nzs + oz, P.mz[2]);
            }
            if (CLS == SW_CLASS_SPML)  // :257
                st_nt(R.vz, p, P.v[2] + (dsxz + dsyz + dszz) * dt * (1.f / rho_avg));
            else if (CLS == SW_CLASS_TENSOR)
                st_nt(R.vz, p, P.v[2] + (dsxz + dsyz + dszz) * dt / rho_avg);
            else
                st_nt(R.vz, p, P.v[2] + (dszz + dsxz + dsyz) * dt / rho_avg);
        }
        win_shift<N>(wxz);
        win_shift<N>(wyz);
        win_shift<N>(wzz);
        P = Q;
    }
    // folded force source: after the loop (no other thread of this sweep
    // reads the cell's velocity), which keeps it out of the loop's registers
    if (fo.force && i == fo.ix && j == fo.iy && fo.iz >= cl.k0 && fo.iz < cl.k1)
        force_src3d(g, fo.force, fo.amp[fo.it_force], ((long)fo.iz * ny + j) * nx + i);
}

// -------------------------------------------------------------- stress sweep
template <int N>
struct SPlane {
    float vx_y[2 * N];  // fm along y (Dy3fm, sxy); centres from the windows,
    float vy_y[2 * N];  // bm along y (Dy3bm, normal)  x taps from the lanes
    float vz_y[2 * N];  // fm along y (Dy3fm, syz)
    float lam;          // p (p+1 from the next lane)
    float mu[3];        // p, p+SY, p+SZ (p+1 from the next lane)
    float s[6];         // sxx syy szz sxy sxz syz at p
    float wvx, wvy, wvz; // next window entries: vx, vy at k+N; vz at k+N-1
    float mx[3], my[3], mz[3];
    Coef cz;
    int zs;
};

template <int N, int PM>
__device__ __forceinline__ void load_splane(SPlane<N> &P, const sw_ssg3d &g, const Res &R, uint32_t pc, int k, int kb,
                                            int SY, uint32_t SZ, const Slab &sl)
{
    const uint32_t vb = pc - (uint32_t)(N * 4 + N * SY) + (uint32_t)(k - kb) * SZ;
    const int c0 = N * 4 + N * SY;
    // y taps without their centres: the z windows hold vx, vy, vz at p
    taps_fm_nc<N>(P.vx_y, R.vx, vb, c0, SY);
    taps_bm_nc<N>(P.vy_y, R.vy, vb, c0, SY);
    taps_fm_nc<N>(P.vz_y, R.vz, vb, c0, SY);
    P.lam = ldu(R.lam, vb, c0);
    P.mu[0] = ldu(R.mu, vb, c0);
    P.mu[1] = ldu(R.mu, vb, c0 + SY);
    P.mu[2] = ldu(R.mu, vb, c0 + (int)SZ);
    P.s[0] = ldu_nt(R.sxx, vb, c0);
    P.s[1] = ldu_nt(R.syy, vb, c0);
    P.s[2] = ldu_nt(R.szz, vb, c0);
    P.s[3] = ldu_nt(R.sxy, vb, c0);
    P.s[4] = ldu_nt(R.sxz, vb, c0);
    P.s[5] = ldu_nt(R.syz, vb, c0);
    P.wvx = ldu(R.vx, vb, c0 + N * (int)SZ);
    P.wvy = ldu(R.vy, vb, c0 + N * (int)SZ);
    P.wvz = ldu(R.vz, vb, c0 + (N - 1) * (int)SZ);
    P.zs = (PM & 4) ? slab_index(k, g.pz.lo, g.pz.hi) : -1;
    P.cz = coef_of(g.pz, k, P.zs >= 0);
    const uint32_t ox = sl.rx + (uint32_t)k * sl.sx, oy = sl.ry + (uint32_t)k * sl.sy;
    const uint32_t oz = (uint32_t)P.zs * sl.sz + pc;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        // x: dvx_dx, dvy_dx, dvz_dx; y: dvy_dy, dvx_dy, dvz_dy; z: dvz_dz, dvx_dz, dvy_dz
        P.mx[m] = sl.xs >= 0 ? ldu(R.mx, ox, (int)((3 + m) * sl.nxs)) : 0.f;
        P.my[m] = sl.ys >= 0 ? ldu(R.my, oy, (int)((3 + m) * sl.nys)) : 0.f;
        P.mz[m] = P.zs >= 0 ? ldu(R.mz, oz, (int)((3 + m) * sl.nzs)) : 0.f;
    }
}

template <int N, int CLS, int PM, bool LARGE>
__global__ void __launch_bounds__(kThreads) k_stress(sw_ssg3d g, Geo geo, Fold3D fo)
{
    if ((int)blockIdx.x >= geo.nblock) {
        gather3d(g, fo, (int)blockIdx.x - geo.nblock, (int)gridDim.x - geo.nblock);
        return;
    }
    const Cell cl = box_cell(geo);
    const int i = cl.i, j = cl.j;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // idle lanes load at their own (row-clamped) column and store nothing
    const bool live = i >= 0;
    const int li = cl.qi, lj = min(max(cl.qj, N), ny - N - 1);
    const int SY = 4 * nx;
    const uint32_t SZ = 4u * (uint32_t)(nx * ny);
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    float c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = g.c[t];
    const bool nrm_col = live && i >= N + 1 && j >= N + 1;
    const bool sxy_col = live && i >= N + 1 && j >= N + 1;
    const bool sxz_col = live && i < nx - N - 1;
    const bool syz_col = live && j < ny - N - 1;
    constexpr int CPM = CLS == SW_CLASS_SPML ? 0 : PM;
    Slab sl = slab_setup<CPM>(g, li, lj);
    if (!live) sl.xs = sl.ys = -1;
    const int kb = LARGE ? cl.k0 - N : 0;  // base plane of this block's offsets
    const Res R = make_res(g, kb);
    const Coef cx = coef_of(g.px, li, sl.xs >= 0);
    const Coef cy = coef_of(g.py, lj, sl.ys >= 0);
    const uint32_t pc = 4u * (uint32_t)(lj * nx + li);
    // windows: wvz[t] = vz[k-N+t] (backward), wvx[t] = vx[k-N+1+t], wvy[t] = vy[k-N+1+t] (forward)
    float wvz[2 * N], wvx[2 * N], wvy[2 * N];
    const int k0 = cl.k0;
#pragma unroll
    for (int t = 0; t < 2 * N - 1; ++t) {
        wvz[t] = ld(R.vz, pc + (uint32_t)(k0 - N + t - kb) * SZ);
        wvx[t] = ld(R.vx, pc + (uint32_t)(k0 - N + 1 + t - kb) * SZ);
        wvy[t] = ld(R.vy, pc + (uint32_t)(k0 - N + 1 + t - kb) * SZ);
    }
    SPlane<N> P;
    load_splane<N, CPM>(P, g, R, pc, k0, kb, SY, SZ, sl);
    const Edge e0 = wave_edge(pc);
    for (int k = k0; k < cl.k1; ++k) {
        const uint32_t p = pc + (uint32_t)(k - kb) * SZ;
        const Edge e{e0.lo + (uint32_t)(k - kb) * SZ, e0.hi + (uint32_t)(k - kb) * SZ};
        wvz[2 * N - 1] = P.wvz;
        wvx[2 * N - 1] = P.wvx;
        wvy[2 * N - 1] = P.wvy;
        SPlane<N> Q;
        if (k + 1 < cl.k1) load_splane<N, CPM>(Q, g, R, pc, k + 1, kb, SY, SZ, sl);
        // centres of the y taps from the windows (vx, vy, vz at p)
        P.vx_y[N - 1] = wvx[N - 1];
        P.vy_y[N] = wvy[N - 1];
        P.vz_y[N - 1] = wvz[N];
        // x taps and the p+1 moduli from the neighbouring lanes (all lanes active)
        float vx_x[2 * N], vy_x[2 * N], vz_x[2 * N];
        xtaps_bm_lanes<N>(vx_x, wvx[N - 1], R.vx, e);
        xtaps_fm_lanes<N>(vy_x, wvy[N - 1], R.vy, e);
        xtaps_fm_lanes<N>(vz_x, wvz[N], R.vz, e);
        const float lam_x = next_x(P.lam, R.lam, e);   // p+1
        const float mu_x = next_x(P.mu[0], R.mu, e);  // p+1
        const uint32_t ox = sl.rx + (uint32_t)k * sl.sx, oy = sl.ry + (uint32_t)k * sl.sy;
        const uint32_t oz = (uint32_t)P.zs * sl.sz + pc;
        // normal stresses at (i, j, k): src/Elastic3D_SSG_CPML.py:247-266
        if (nrm_col && k >= N + 1) {
            // CPML/Tensor average lambda, mu over (i, i+1); SPML takes the
            // point values (src/Elastic3D_SSG_SPML.py:300-302)
            const float lam_h = CLS == SW_CLASS_SPML ? P.lam : 0.5f * (P.lam + lam_x);
            const float mu_h = CLS == SW_CLASS_SPML ? P.mu[0] : 0.5f * (P.mu[0] + mu_x);
            const float lpm = lam_h + 2.f * mu_h;
            float dvxx = diff_bm<N>(vx_x, c) / dx;
            float dvyy = diff_bm<N>(P.vy_y, c) / dy;
            float dvzz = diff_bm<N>(wvz, c) / dz;
            if (sl.xs >= 0) {
                dvxx = cpml_step(dvxx, P.mx[0], cx.a, cx.b, cx.k);
                st(R.mx, 3 * sl.nxs + ox, P.mx[0]);
            }
            if (sl.ys >= 0) {
                dvyy = cpml_step(dvyy, P.my[0], cy.a, cy.b, cy.k);
                st(R.my, 3 * sl.nys + oy, P.my[0]);
            }
            if (P.zs >= 0) {
                dvzz = cpml_step(dvzz, P.mz[0], P.cz.a, P.cz.b, P.cz.k);
                st(R.mz, 3 * sl.nzs + oz, P.mz[0]);
            }
            st_nt(R.sxx, p, P.s[0] + (lpm * dvxx + lam_h * (dvzz + dvyy)) * dt);
            st_nt(R.syy, p, P.s[1] + (lpm * dvyy + lam_h * (dvxx + dvzz)) * dt);
            st_nt(R.szz, p, P.s[2] + (lpm * dvzz + lam_h * (dvxx + dvyy)) * dt);
        }
        // sxy at (i+1/2, j+1/2, k): :267-279
        if (sxy_col) {
            const float mu_a = 0.25f * (P.mu[1] + P.mu[0] + P.mu[0] + mu_x);
            float dvyx = diff_fm<N>(vy_x, c) / dx;
            float dvxy = diff_fm<N>(P.vx_y, c) / dy;
            if (sl.xs >= 0) {
                dvyx = cpml_step(dvyx, P.mx[1], cx.ah, cx.bh, cx.kh);
                st(R.mx, 4 * sl.nxs + ox, P.mx[1]);
            }
            if (sl.ys >= 0) {
                dvxy = cpml_step(dvxy, P.my[1], cy.ah, cy.bh, cy.kh);
                st(R.my, 4 * sl.nys + oy, P.my[1]);
            }
            st_nt(R.sxy, p, P.s[3] + mu_a * (dvxy + dvyx) * dt);
        }
        // sxz at (i+1/2, j, k+1/2): :281-293
        if (sxz_col && k < nz - N - 1) {
            const float mu_a = 0.25f * (P.mu[0] + mu_x + P.mu[0] + P.mu[1]);
            float dvzx = diff_fm<N>(vz_x, c) / dx;
            float dvxz = diff_fm<N>(wvx, c) / dz;
            if (sl.xs >= 0) {
                dvzx = cpml_step(dvzx, P.mx[2], cx.ah, cx.bh, cx.kh);
                st(R.mx, 5 * sl.nxs + ox, P.mx[2]);
            }
            if (P.zs >= 0) {
                dvxz = cpml_step(dvxz, P.mz[1], P.cz.ah, P.cz.bh, P.cz.kh);
                st(R.mz, 4 * sl.nzs + oz, P.mz[1]);
            }
            st_nt(R.sxz, p, P.s[4] + mu_a * (dvxz + dvzx) * dt);
        }
        // syz at (i, j+1/2, k+1/2): :295-307
        if (syz_col && k < nz - N - 1) {
            const float mu_a = 0.25f * (P.mu[1] + P.mu[0] + P.mu[2] + P.mu[0]);
            float dvyz = diff_fm<N>(wvy, c) / dz;
            float dvzy = diff_fm<N>(P.vz_y, c) / dy;
            if (sl.ys >= 0) {
                dvzy = cpml_step(dvzy, P.my[2], cy.ah, cy.bh, cy.kh);
                st(R.my, 5 * sl.nys + oy, P.my[2]);
            }
            if (P.zs >= 0) {
                dvyz = cpml_step(dvyz, P.mz[2], P.cz.ah, P.cz.bh, P.cz.kh);
                st(R.mz, 5 * sl.nzs + oz, P.mz[2]);
            }
            st_nt(R.syz, p, P.s[5] + mu_a * (dvyz + dvzy) * dt);
        }
        win_shift<N>(wvz);
        win_shift<N>(wvx);
        win_shift<N>(wvy);
        P = Q;
    }
    // folded next-step stress source, after the loop (no other thread of
    // this sweep reads stresses)
    if (fo.stress && i == fo.ix && j == fo.iy && fo.iz >= cl.k0 && fo.iz < cl.k1)
        stress_src3d(g, fo.stress, fo.mt, fo.amp[fo.it_next], ((long)fo.iz * ny + j) * nx + i);
}


// Cells a sweep updates: [N, n-N) on every axis, in one launch (PM = 7) or
// in five launches whose boxes all keep full-width x rows (long coalesced
// segments):
//   PM = 1: planes and rows outside the z and y slabs (x slab per lane);
//   PM = 3: rows inside a y slab, planes outside the z slabs (2 boxes);
//   PM = 7: planes inside a z slab (2 boxes).
// The mask PM names the axes whose CPML the kernel instantiates; dropping z
// and y from the bulk of the grid frees the registers of their memory
// variables and coefficients, so more waves are in flight there.
struct Plan {
    int n = 0;
    int pm[5];
    Geo g[5];
};

// Chunks of kZChunk planes on every grid: the many short blocks spread
// evenly over the CUs and keep the last round of a launch short, and the
// 2N-1 window planes a chunk re-reads are those its group of column blocks
// read one chunk earlier (kGroup), still in L2.  Measured on MI355X, ms per
// step: in plane-by-plane block order with chunks of 2 / 3 / 4 / 8 / 16
// planes, 200^3 0.309 / 0.310 / 0.307 / 0.318 / -, 600x600x400 - / - / 4.66 /
// 4.74 / 5.08 (5.03 with 64-plane chunks); in groups of 32 column blocks
// with chunks of 2 / 4 / 6 / 8 planes, 200^3 0.309 / 0.305 / 0.313 / 0.317,
// 600x600x400 4.77 (4) / 4.92 (8); with line-aligned columns and chunks of
// 4 / 6 / 8 planes, 200^3 0.289 / 0.289 / 0.298, 600x600x400 4.42 / 4.47 /
// 4.49, 200^3 split-field class 0.425 / 0.436 / 0.450; with the x taps from
// neighbouring lanes 3-plane chunks (kGroup has the figures).
void add_box(Plan &p, int nx, int pm, int i0, int i1, int j0, int j1, int k0, int k1)
{
    if (i1 <= i0 || j1 <= j0 || k1 <= k0) return;
    Geo &geo = p.g[p.n];
    p.pm[p.n++] = pm;
    Box &b = geo.box;
    b.i0 = i0; b.w = i1 - i0;
    b.j0 = j0; b.h = j1 - j0;
    b.k0 = k0; b.k1 = k1;
    b.nx = nx;
    b.f0 = (j0 * nx + i0) & ~31;
    b.fn = (j1 - 1) * nx + i1;
    b.ncb = (b.fn - b.f0 + kThreads - 1) / kThreads;
    geo.nblock = (int)((long)b.ncb * ((k1 - k0 + kZChunk - 1) / kZChunk));
}

Plan make_plan(const sw_ssg3d &g, int N)
{
    Plan p;
    const int x0 = N, x1 = g.nx - N, y0 = N, y1 = g.ny - N, z0 = N, z1 = g.nz - N;
    const int yl = std::max(y0, g.py.lo), yh = std::min(y1, g.py.hi);
    const int zl = std::max(z0, g.pz.lo), zh = std::min(z1, g.pz.hi);
    // Several launches pay off only when the bulk box dominates and every
    // slab launch fills the GPU: each extra launch adds a ramp-up and a tail,
    // and a launch of fewer blocks than compute units leaves CUs idle for its
    // whole length.  Measured on MI355X, ms per step: 200^3 with a 20-cell
    // CPML (bulk 64 %) one launch 0.289 against 0.355 in five; 600x600x400
    // (bulk 84 %, slab launches of 1,200-7,000 blocks) five 4.34 against 4.55;
    // its 600x600x54 shard (y slabs of 611 blocks) 0.563 against 0.576; the
    // 168^3 grid of a 4-cell CPML (bulk 95 %, slab launches of 105-120
    // blocks) one launch 0.206 against 0.256 per learned-boundary step.
    // With the x taps from neighbouring lanes (4 waves per SIMD in the
    // PM=7 kernels, 5 in PM=1): 200^3 one 0.274 against 0.333 in five,
    // 600x600x400 five 4.132 against 4.233, the shard 0.534 either way.
    const double bulk = (yl < yh && zl < zh) ? (double)(yh - yl) * (zh - zl) / ((double)(y1 - y0) * (z1 - z0)) : 0.0;
    const int mode = launch_plan_mode();
    if (mode != SW_PLAN_ONE && (mode == SW_PLAN_BOXES ? bulk > 0.0 : bulk >= 0.75)) {
        add_box(p, g.nx, 1, x0, x1, yl, yh, zl, zh);
        add_box(p, g.nx, 3, x0, x1, y0, yl, zl, zh);
        add_box(p, g.nx, 3, x0, x1, yh, y1, zl, zh);
        add_box(p, g.nx, 7, x0, x1, y0, y1, z0, zl);
        add_box(p, g.nx, 7, x0, x1, y0, y1, zh, z1);
        bool fills = true;
        for (int t = 1; t < p.n; ++t) fills = fills && p.g[t].nblock >= compute_units();
        if (mode == SW_PLAN_BOXES || fills) return p;
        p = Plan();
    }
    add_box(p, g.nx, 7, x0, x1, y0, y1, z0, z1);
    return p;
}

template <int N, int CLS, bool LARGE>
void launch_velocity(const sw_ssg3d &g, const Plan &p, const Fold3D &fo, hipStream_t st)
{
    for (int t = 0; t < p.n; ++t) {
        const dim3 grd(p.g[t].nblock), blk(kThreads);
        if (p.pm[t] == 0)
            hipLaunchKernelGGL((k_velocity<N, CLS, 0, LARGE>), grd, blk, 0, st, g, p.g[t], fo);
        else if (p.pm[t] == 1)
            hipLaunchKernelGGL((k_velocity<N, CLS, 1, LARGE>), grd, blk, 0, st, g, p.g[t], fo);
        else if (p.pm[t] == 3)
            hipLaunchKernelGGL((k_velocity<N, CLS, 3, LARGE>), grd, blk, 0, st, g, p.g[t], fo);
        else
            hipLaunchKernelGGL((k_velocity<N, CLS, 7, LARGE>), grd, blk, 0, st, g, p.g[t], fo);
    }
}

// The first launch also carries the folded receiver gather (fo.gather extra
// blocks); the others get a copy without it.
template <int N, int CLS, bool LARGE>
void launch_stress(const sw_ssg3d &g, const Plan &p, const Fold3D &fo, hipStream_t st)
{
    Fold3D rest = fo;
    rest.record = 0;
    rest.gather = 0;
    for (int t = 0; t < p.n; ++t) {
        const Fold3D &f = t == 0 ? fo : rest;
        const dim3 grd(p.g[t].nblock + (f.record ? f.gather : 0)), blk(kThreads);
        if (p.pm[t] == 0)
            hipLaunchKernelGGL((k_stress<N, CLS, 0, LARGE>), grd, blk, 0, st, g, p.g[t], f);
        else if (p.pm[t] == 1)
            hipLaunchKernelGGL((k_stress<N, CLS, 1, LARGE>), grd, blk, 0, st, g, p.g[t], f);
        else if (p.pm[t] == 3)
            hipLaunchKernelGGL((k_stress<N, CLS, 3, LARGE>), grd, blk, 0, st, g, p.g[t], f);
        else
            hipLaunchKernelGGL((k_stress<N, CLS, 7, LARGE>), grd, blk, 0, st, g, p.g[t], f);
    }
}

}  // namespace

// Arrays above 4 GiB take the rebasing kernels.
inline bool large_grid(const sw_ssg3d &g) { return 4L * g.nx * g.ny * g.nz > (1L << 32); }

// The split-free interior of the split-field class in the stress sweep: one
// PM = 0 launch (plain updates, no split values); the shell around it is
// launched by launch_split_shell_stress (ssg3d_split.hip).
Plan make_interior_plan(const sw_ssg3d &g, int N)
{
    Plan p;
    const SplitBox b = split_box(g, N);
    add_box(p, g.nx, 0, b.xl, b.xh, b.yl, b.yh, b.zl, b.zh);
    return p;
}

template <int N, int CLS>
void launch_velocity_any(const sw_ssg3d &g, const Plan &p, const Fold3D &fo, hipStream_t st)
{
    if (large_grid(g))
        launch_velocity<N, CLS, true>(g, p, fo, st);
    else
        launch_velocity<N, CLS, false>(g, p, fo, st);
}

template <int N, int CLS>
void launch_stress_any(const sw_ssg3d &g, const Plan &p, const Fold3D &fo, hipStream_t st)
{
    if (large_grid(g))
        launch_stress<N, CLS, true>(g, p, fo, st);
    else
        launch_stress<N, CLS, false>(g, p, fo, st);
}

int launch_ssg3d_velocity_fold(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st)
{
    // split-field class: its slabs and full-width middle rows (ssg3d_split.hip)
    if (g->cls == SW_CLASS_SPML) return launch_split_shell_velocity(g, fo, st);
    SW_DISPATCH_ORDER(g->order, {
        const Plan p = make_plan(*g, N);
        if (g->cls == SW_CLASS_TENSOR)
            launch_velocity_any<N, SW_CLASS_TENSOR>(*g, p, fo, st);
        else
            launch_velocity_any<N, SW_CLASS_CPML>(*g, p, fo, st);
    });
    return launch_status("ssg3d velocity");
}

// Split-field PML class (src/Elastic3D_SSG_SPML.py:160-393): the shell launch
// (which carries the folded receiver gather) and the interior launch.
int launch_ssg3d_stress_fold(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st)
{
    if (g->cls == SW_CLASS_SPML) {
        int rc = launch_split_shell_stress(g, fo, st);
        if (rc) return rc;
        Fold3D rest = fo;
        rest.record = 0;
        rest.gather = 0;
        SW_DISPATCH_ORDER(g->order, {
            const Plan p = make_interior_plan(*g, N);
            launch_stress_any<N, SW_CLASS_SPML>(*g, p, rest, st);
        });
        return launch_status("ssg3d stress (split-field PML interior)");
    }
    SW_DISPATCH_ORDER(g->order, {
        const Plan p = make_plan(*g, N);
        launch_stress_any<N, SW_CLASS_CPML>(*g, p, fo, st);
    });
    return launch_status("ssg3d stress");
}

int ssg3d_plan_masks(const sw_ssg3d *g, int32_t *pm, int32_t max)
{
    int n = 0;
    if (g->cls == SW_CLASS_SPML) {  // shell (flags of all three axes), interior
        const int32_t m[2] = {7, 0};
        for (int t = 0; t < 2 && t < max; ++t) pm[t] = m[t];
        return 2;
    }
    SW_DISPATCH_ORDER(g->order, {
        const Plan p = make_plan(*g, N);
        n = p.n;
        for (int t = 0; t < p.n && t < max; ++t) pm[t] = p.pm[t];
    });
    return n;
}

int launch_ssg3d_velocity(const sw_ssg3d *g, hipStream_t st) { return launch_ssg3d_velocity_fold(g, Fold3D{}, st); }
int launch_ssg3d_stress(const sw_ssg3d *g, hipStream_t st) { return launch_ssg3d_stress_fold(g, Fold3D{}, st); }
int launch_ssg3d_velocity_spml(const sw_ssg3d *g, hipStream_t st) { return launch_ssg3d_velocity_fold(g, Fold3D{}, st); }
int launch_ssg3d_stress_spml(const sw_ssg3d *g, hipStream_t st) { return launch_ssg3d_stress_fold(g, Fold3D{}, st); }

}  // namespace sw
