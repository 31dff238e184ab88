// Split-field PML of the 3-D sweeps (src/Elastic3D_SSG_SPML.py:160-393).
//
// The reference updates a field component from Crank-Nicolson split values
// (one per derivative direction, damped by that direction's d) wherever the
// component's own in-PML test holds, and plainly elsewhere.  The split values
// live in shell storage (sw_internal.h: SplitShell): only the cells outside
// the split-free box, where every test is false, have a slot, and a slot's
// 24 components sit in groups of up to four floats within its 64-slot block,
// so a thread moves its 9 (velocity) or 15 (stress) split values with 3 or 5
// vector accesses, a wave's accesses to them cover one contiguous block, and
// a group is written back whole (components the step leaves unchanged with
// their old values).  Split values are read once and written once per sweep,
// so they are loaded and stored non-temporally and do not displace the
// stencil's field lines from L2.
//
// Launches (boxes of updated cells, the block order of the plain sweeps):
//  * stress sweep: one launch over the shell, in four boxes (planes below and
//    above the split-free box; in its planes the rows outside it, one box
//    with a row gap; in its rows the columns outside it, one box with a
//    column gap), then the split-free interior with the plain kernels of
//    ssg3d.hip.  Its split values are loaded when a plane starts, not with
//    the prefetched plane, which keeps the kernel at 124 registers (4 waves
//    per SIMD at order 4).
//  * velocity sweep: the z and y slabs of the shell, then the middle rows at
//    full width, whose x-strip lanes take the split updates and whose other
//    lanes the plain ones (full rows: no line is shared by two launches).
//    The slab kernel prefetches the split values with the plane; the
//    middle-rows kernel prefetches the old velocities instead and loads the
//    strip lanes' split values when a plane starts.
// Measured on MI355X, 200^3 with a 20-cell layer, ms per sweep (velocity /
// stress): 0.191 / 0.277 with 24 full-grid split arrays and one launch over
// everything; 0.199 / 0.293 shell and interior launches, one split array per
// component; 0.193 / 0.289 grouped split storage; 0.182 / 0.263 with
// non-temporal split accesses and late split loads (stress); 0.180 / 0.270
// with the middle rows at full width in both sweeps (the stress kernel holds
// 142 registers there, 3 waves per SIMD); interior and shell in one launch
// 0.225 / 0.337.
// A component whose own test is false at a shell cell (the innermost layer
// of a slab, where only the other stagger's flag is set) takes the plain
// update, loading its old value then.
// Arithmetic is term for term the reference's (build with -ffp-contract=off).
#include <climits>

#include "ssg3d_common.h"

namespace sw {

namespace {

// One box of shell cells: columns [i0, i0+w) and rows [j0, j0+h) with the
// gaps applied (a column at or past gi moves right by gwi, a row at or past
// gj up by gwj), planes [k0, k1).  Split slot of cell (i, j, k):
//   C + jj*pitch + ii + k*kstride  (mod 2^32), ii = i < si ? i : i - di,
//   jj = (j < sj ? j : j - dj) - oj.
struct SBox {
    int i0, w, j0, h, k0, k1, ncb;
    int gi, gwi, gj, gwj;
    int si, di, sj, dj, oj;
    uint32_t C, pitch, kstride;
    // boxes of full rows (no column gap) number their columns by flat index
    // v = j*nx + i over rows [j0, j0+h) before the row gap, from f0 (rounded
    // down to a multiple of 64) to fn: a wave's columns then start on a
    // 128-byte line of every field and, in the z and y slabs, on a 64-slot
    // block of the split storage; lanes on the x edges of a row idle
    int flat, nx, f0, fn;
};

struct ShellGeo {
    SBox box[4];
    int start[4];  // first logical block of each box
    int nbox, nblock;
};

// Column (i, j) and z chunk [k0, k1) of logical block b, thread t, in the
// block order of the plain sweeps (groups of kGroup column blocks marching
// through their z chunks), and the split slot of its plane k: q0 + k*qk.
struct SCell {
    int i, j, k0, k1;
    uint32_t q0, qk;
};

__host__ __device__ inline SCell shell_cell(const ShellGeo &geo, int b, int t)
{
    int s = 0;
    for (int u = 1; u < 4; ++u)
        if (u < geo.nbox && b >= geo.start[u]) s = u;
    const SBox &B = geo.box[s];
    const int lb = b - geo.start[s];
    const int nzc = (B.k1 - B.k0 + kZChunk - 1) / kZChunk;
    const int grp = lb / (kGroup * nzc), r = lb - grp * kGroup * nzc;
    const int left = B.ncb - grp * kGroup;
    const int gw = left < kGroup ? left : kGroup;
    const int zb = r / gw, cb = grp * kGroup + r - zb * gw;
    const int c = cb * kThreads + t;
    SCell o;
    o.k0 = B.k0 + zb * kZChunk;
    o.k1 = o.k0 + kZChunk < B.k1 ? o.k0 + kZChunk : B.k1;
    int i, j;
    bool idle;
    if (B.flat) {
        const int v = B.f0 + c;
        j = v / B.nx;
        i = v - j * B.nx;
        idle = v >= B.fn || j < B.j0 || j >= B.j0 + B.h || i < B.i0 || i >= B.i0 + B.w;
    } else {
        idle = c >= B.w * B.h;
        i = B.i0 + c % B.w;
        j = B.j0 + c / B.w;
        if (i >= B.gi) i += B.gwi;
    }
    if (idle) {
        o.i = -1; o.j = -1; o.q0 = 0; o.qk = 0;
        return o;
    }
    if (j >= B.gj) j += B.gwj;
    o.i = i;
    o.j = j;
    const int ii = i < B.si ? i : i - B.di;
    const int jj = (j < B.sj ? j : j - B.dj) - B.oj;
    o.q0 = B.C + (uint32_t)jj * B.pitch + (uint32_t)ii;
    o.qk = B.kstride;
    return o;
}

// Box sets of the launches: the whole shell (stress sweep), its z and y
// slabs only, and the middle rows at full width (velocity sweep).
enum { kShell = 0, kSlabs = 1, kMiddleRows = 2 };

ShellGeo make_shell_geo(const sw_ssg3d &g, int N, int part)
{
    const SplitBox b = split_box(g, N);
    const SplitShell sh = split_shell(g);
    const uint32_t nx = (uint32_t)g.nx, PL = (uint32_t)g.nx * (uint32_t)g.ny;
    ShellGeo geo{};
    int nblock = 0;
    auto add = [&](int x0, int x1, int gi, int gwi, int y0, int y1, int gj, int gwj, int k0, int k1, SBox slot) {
        const int w = x1 - x0 - gwi, h = y1 - y0 - gwj;
        if (w <= 0 || h <= 0 || k1 <= k0) return;
        SBox &B = geo.box[geo.nbox];
        B = slot;
        B.i0 = x0; B.w = w; B.gi = gwi ? gi : INT_MAX; B.gwi = gwi;
        B.j0 = y0; B.h = h; B.gj = gwj ? gj : INT_MAX; B.gwj = gwj;
        B.k0 = k0; B.k1 = k1;
        B.flat = gwi == 0;
        B.nx = g.nx;
        B.f0 = (y0 * g.nx + x0) & ~63;
        B.fn = (y0 + h - 1) * g.nx + x0 + w;
        B.ncb = B.flat ? (B.fn - B.f0 + kThreads - 1) / kThreads : (w * h + kThreads - 1) / kThreads;
        geo.start[geo.nbox++] = nblock;
        nblock += B.ncb * ((k1 - k0 + kZChunk - 1) / kZChunk);
    };
    if (part == kShell || part == kSlabs) {
        SBox zslab{};
        zslab.si = INT_MAX; zslab.sj = INT_MAX;
        zslab.pitch = nx; zslab.kstride = PL;
        zslab.C = 0;
        add(b.x0, b.x1, 0, 0, b.y0, b.y1, 0, 0, b.z0, b.zl, zslab);
        zslab.C = 0u - (uint32_t)(sh.zhi - sh.zlo) * PL;
        add(b.x0, b.x1, 0, 0, b.y0, b.y1, 0, 0, b.zh, b.z1, zslab);
        SBox yslab{};
        yslab.si = INT_MAX;
        yslab.sj = sh.ylo; yslab.dj = sh.yhi - sh.ylo;
        yslab.pitch = nx; yslab.kstride = (uint32_t)sh.wy * nx;
        yslab.C = (uint32_t)sh.OZ - (uint32_t)sh.zlo * yslab.kstride;
        add(b.x0, b.x1, 0, 0, b.y0, b.y1, b.yl, b.yh - b.yl, b.zl, b.zh, yslab);
    }
    if (part == kShell || part == kMiddleRows) {
        // the x part of the storage; in the middle-rows box the split-free
        // columns get slots too, which the kernel never uses
        SBox xslab{};
        xslab.si = sh.xlo; xslab.di = sh.xhi - sh.xlo;
        xslab.sj = INT_MAX; xslab.oj = sh.ylo;
        xslab.pitch = (uint32_t)sh.wx; xslab.kstride = (uint32_t)sh.my * (uint32_t)sh.wx;
        xslab.C = (uint32_t)sh.OY - (uint32_t)sh.zlo * xslab.kstride;
        if (part == kShell)
            add(b.x0, b.x1, b.xl, b.xh - b.xl, b.yl, b.yh, 0, 0, b.zl, b.zh, xslab);
        else
            add(b.x0, b.x1, 0, 0, b.yl, b.yh, 0, 0, b.zl, b.zh, xslab);
    }
    geo.nblock = nblock;
    return geo;
}

// Crank-Nicolson step of one split component, q = (n q + incr) / d
// (src/Elastic3D_SSG_SPML.py:199-204 and alike).
__device__ __forceinline__ float cn(float &q, float n, float d, float incr)
{
    q = (n * q + incr) / d;
    return q;
}

// The split block of slot s and the slot's lane in it.
struct SplitRef {
    float *blk;
    uint32_t l;
};

__device__ __forceinline__ SplitRef split_ref(float *split, uint32_t s)
{
    return SplitRef{split + (size_t)(s >> 6) * kSplitBlockFloats, s & 63u};
}

typedef float f4a __attribute__((ext_vector_type(4)));
typedef float f2a __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void ld4(float *q, const float *p)
{
    const f4a v = __builtin_nontemporal_load(reinterpret_cast<const f4a *>(p));
    q[0] = v.x; q[1] = v.y; q[2] = v.z; q[3] = v.w;
}

__device__ __forceinline__ void st4(float *p, const float *q)
{
    __builtin_nontemporal_store(f4a{q[0], q[1], q[2], q[3]}, reinterpret_cast<f4a *>(p));
}

__device__ __forceinline__ void ld2(float *q, const float *p)
{
    const f2a v = __builtin_nontemporal_load(reinterpret_cast<const f2a *>(p));
    q[0] = v.x; q[1] = v.y;
}

__device__ __forceinline__ void st2(float *p, const float *q)
{
    __builtin_nontemporal_store(f2a{q[0], q[1]}, reinterpret_cast<f2a *>(p));
}

__device__ __forceinline__ float ld1(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st1(float *p, float v) { __builtin_nontemporal_store(v, p); }

// ------------------------------------------------------------ velocity sweep
template <int N, bool MID>
struct VSPlane {
    float sxx[2 * N];   // fm taps along x (Dx3fm, vx)
    float sxy_x[2 * N]; // bm taps along x (Dx3bm, vy)
    float sxy_y[2 * N]; // bm taps along y (Dy3bm, vx)
    float syy[2 * N];   // fm taps along y (Dy3fm, vy)
    float sxz_x[2 * N]; // bm taps along x (Dx3bm, vz)
    float syz_y[2 * N]; // bm taps along y (Dy3bm, vz)
    float rho[4];       // p, p+SY, p+1+SY, p+1
    float wxz, wyz, wzz; // next window entries: sxz, syz at k+N-1, szz at k+N
    float q[MID ? 1 : 9]; // slabs: vx_x vx_y vx_z vy_x vy_y vy_z vz_x vz_y vz_z
    float v[MID ? 3 : 1]; // middle rows: vx, vy, vz at p
};

template <int N, bool MID>
__device__ __forceinline__ void load_vsplane(VSPlane<N, MID> &P, const Res &R, const SplitRef &q, uint32_t pc, int k,
                                             int kb, int SY, uint32_t SZ)
{
    const uint32_t vb = pc - (uint32_t)(N * 4 + N * SY) + (uint32_t)(k - kb) * SZ;
    const int c0 = N * 4 + N * SY;
    xtaps_fm<N>(P.sxx, R.sxx, vb, c0);
    xtaps_bm<N>(P.sxy_x, R.sxy, vb, c0);
    taps_bm<N>(P.sxy_y, R.sxy, vb, c0, SY);
    taps_fm<N>(P.syy, R.syy, vb, c0, SY);
    xtaps_bm<N>(P.sxz_x, R.sxz, vb, c0);
    taps_bm<N>(P.syz_y, R.syz, vb, c0, SY);
    P.rho[0] = ldu(R.rho, vb, c0);
    P.rho[1] = ldu(R.rho, vb, c0 + SY);
    P.rho[2] = ldu(R.rho, vb, c0 + 4 + SY);
    P.rho[3] = ldu(R.rho, vb, c0 + 4);
    P.wxz = ldu(R.sxz, vb, c0 + (N - 1) * (int)SZ);
    P.wyz = ldu(R.syz, vb, c0 + (N - 1) * (int)SZ);
    P.wzz = ldu(R.szz, vb, c0 + N * (int)SZ);
    if (MID) {
        P.v[0] = ldu(R.vx, vb, c0);
        P.v[1] = ldu(R.vy, vb, c0);
        P.v[2] = ldu(R.vz, vb, c0);
    } else {
        ld4(P.q, q.blk + 4 * q.l);
        ld4(P.q + 4, q.blk + 256 + 4 * q.l);
        P.q[8] = ld1(q.blk + 512 + q.l);
    }
}

template <int N, bool LARGE, bool MID>
__global__ void __launch_bounds__(kThreads) k_velocity_shell(sw_ssg3d g, ShellGeo geo, Fold3D fo)
{
    const SCell cl = shell_cell(geo, xcd_logical_block(blockIdx.x, geo.nblock), (int)threadIdx.x);
    const int i = cl.i, j = cl.j;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i < N || i >= nx - N || j < N || j >= ny - N) return;
    const int SY = 4 * nx;
    const uint32_t SZ = 4u * (uint32_t)(nx * ny);
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    const float hdt = 0.5f * dt;
    float c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = g.c[t];
    const bool vx_col = i < nx - N - 1 && j >= N + 1;
    const bool vy_col = j < ny - N - 1 && i >= N + 1;
    const bool vz_col = i >= N + 1 && j >= N + 1;
    // in-PML flags and damping of the column's integer and half points
    const bool xi = g.px.in[i] != 0, xh = g.px.in_h[i] != 0;
    const bool yi = g.py.in[j] != 0, yh = g.py.in_h[j] != 0;
    const float dxi = g.px.d[i], dxh = g.px.d_h[i], dyi = g.py.d[j], dyh = g.py.d_h[j];
    const int kb = LARGE ? cl.k0 - N : 0;  // base plane of this block's field offsets
    const Res R = make_res(g, kb);
    const uint32_t pc = 4u * (uint32_t)(j * nx + i);
    // windows: wxz[t] = sxz[k-N+t], wyz[t] = syz[k-N+t], wzz[t] = szz[k-N+1+t]
    float wxz[2 * N], wyz[2 * N], wzz[2 * N];
    const int k0 = cl.k0;
#pragma unroll
    for (int t = 0; t < 2 * N - 1; ++t) {
        wxz[t] = ld(R.sxz, pc + (uint32_t)(k0 - N + t - kb) * SZ);
        wyz[t] = ld(R.syz, pc + (uint32_t)(k0 - N + t - kb) * SZ);
        wzz[t] = ld(R.szz, pc + (uint32_t)(k0 - N + 1 + t - kb) * SZ);
    }
    const bool strip = MID ? (i < g.px.lo || i >= g.px.hi) : true;
    VSPlane<N, MID> P;
    load_vsplane<N, MID>(P, R, split_ref(g.split, cl.q0 + (uint32_t)k0 * cl.qk), pc, k0, kb, SY, SZ);
    for (int k = k0; k < cl.k1; ++k) {
        float qm[9];
        if (MID && strip) {
            const SplitRef s0 = split_ref(g.split, cl.q0 + (uint32_t)k * cl.qk);
            ld4(qm, s0.blk + 4 * s0.l);
            ld4(qm + 4, s0.blk + 256 + 4 * s0.l);
            qm[8] = ld1(s0.blk + 512 + s0.l);
        }
        float *const qq = MID ? qm : P.q;
        VSPlane<N, MID> Q;
        if (k + 1 < cl.k1)
            load_vsplane<N, MID>(Q, R, split_ref(g.split, cl.q0 + (uint32_t)(k + 1) * cl.qk), pc, k + 1, kb, SY, SZ);
        const bool zi = g.pz.in[k] != 0, zh = g.pz.in_h[k] != 0;
        const float dzi = g.pz.d[k], dzh = g.pz.d_h[k];
        const SplitRef sq = split_ref(g.split, cl.q0 + (uint32_t)k * cl.qk);
        const uint32_t p = pc + (uint32_t)(k - kb) * SZ;
        wxz[2 * N - 1] = P.wxz;
        wyz[2 * N - 1] = P.wyz;
        wzz[2 * N - 1] = P.wzz;
        // vx at (i+1/2, j, k): src/Elastic3D_SSG_SPML.py:182-205, x+y+z times 1/rho
        if (vx_col && k >= N + 1) {
            const float dsxx = diff_fm<N>(P.sxx, c) / dx;
            const float dsxy = diff_bm<N>(P.sxy_y, c) / dy;
            const float dsxz = diff_bm<N>(wxz, c) / dz;
            const float rho_over = 1.f / P.rho[0];
            if (strip && (xh | yi | zi)) {
                const float a = cn(qq[0], 1.f - hdt * dxh, 1.f + hdt * dxh, dsxx * dt * rho_over);
                const float b = cn(qq[1], 1.f - hdt * dyi, 1.f + hdt * dyi, dsxy * dt * rho_over);
                const float e = cn(qq[2], 1.f - hdt * dzi, 1.f + hdt * dzi, dsxz * dt * rho_over);
                st_nt(R.vx, p, a + b + e);
            } else {
                st_nt(R.vx, p, (MID ? P.v[0] : ld(R.vx, p)) + (dsxx + dsxy + dsxz) * dt * rho_over);
            }
        }
        // vy at (i, j+1/2, k): :209-233
        if (vy_col && k >= N + 1) {
            const float rho_half_y = 0.5f * (P.rho[0] + P.rho[1]);
            const float dsxy = diff_bm<N>(P.sxy_x, c) / dx;
            const float dsyy = diff_fm<N>(P.syy, c) / dy;
            const float dsyz = diff_bm<N>(wyz, c) / dz;
            const float rho_over = 1.f / rho_half_y;
            if (strip && (xi | yh | zi)) {
                const float a = cn(qq[3], 1.f - hdt * dxi, 1.f + hdt * dxi, dt * dsxy * rho_over);
                const float b = cn(qq[4], 1.f - hdt * dyh, 1.f + hdt * dyh, dt * dsyy * rho_over);
                const float e = cn(qq[5], 1.f - hdt * dzi, 1.f + hdt * dzi, dt * dsyz * rho_over);
                st_nt(R.vy, p, a + b + e);
            } else {
                st_nt(R.vy, p, (MID ? P.v[1] : ld(R.vy, p)) + (dsxy + dsyy + dsyz) * dt * rho_over);
            }
        }
        // vz at (i, j, k+1/2): :235-259
        if (vz_col && k < nz - N - 1) {
            const float rho_avg = 0.25f * (P.rho[0] + P.rho[1] + P.rho[2] + P.rho[3]);
            const float dsxz = diff_bm<N>(P.sxz_x, c) / dx;
            const float dsyz = diff_bm<N>(P.syz_y, c) / dy;
            const float dszz = diff_fm<N>(wzz, c) / dz;
            const float rho_over = 1.f / rho_avg;
            if (strip && (xi | yi | zh)) {
                const float a = cn(qq[6], 1.f - hdt * dxi, 1.f + hdt * dxi, dt * dsxz * rho_over);
                const float b = cn(qq[7], 1.f - hdt * dyi, 1.f + hdt * dyi, dt * dsyz * rho_over);
                const float e = cn(qq[8], 1.f - hdt * dzh, 1.f + hdt * dzh, dt * dszz * rho_over);
                st_nt(R.vz, p, a + b + e);
            } else {
                st_nt(R.vz, p, (MID ? P.v[2] : ld(R.vz, p)) + (dsxz + dsyz + dszz) * dt * rho_over);
            }
        }
        // the split groups, updated or not
        if (strip) {
            st4(sq.blk + 4 * sq.l, qq);
            st4(sq.blk + 256 + 4 * sq.l, qq + 4);
            st1(sq.blk + 512 + sq.l, qq[8]);
        }
        win_shift<N>(wxz);
        win_shift<N>(wyz);
        win_shift<N>(wzz);
        P = Q;
    }
    // folded force source, after the loop (no other thread of this sweep
    // reads the cell's velocity)
    if (fo.force && i == fo.ix && j == fo.iy && fo.iz >= cl.k0 && fo.iz < cl.k1)
        force_src3d(g, fo.force, fo.amp[fo.it_force], ((long)fo.iz * ny + j) * nx + i);
}

// -------------------------------------------------------------- stress sweep
template <int N>
struct SSPlane {
    float vx_x[2 * N];  // bm along x (Dx3bm, normal)
    float vx_y[2 * N];  // fm along y (Dy3fm, sxy)
    float vy_x[2 * N];  // fm along x (Dx3fm, sxy)
    float vy_y[2 * N];  // bm along y (Dy3bm, normal)
    float vz_x[2 * N];  // fm along x (Dx3fm, sxz)
    float vz_y[2 * N];  // fm along y (Dy3fm, syz)
    float lam;          // p
    float mu[4];        // p, p+1, p+SY, p+SZ
    float wvx, wvy, wvz; // next window entries: vx, vy at k+N; vz at k+N-1
    // sxx_x sxx_y sxx_z syy_x syy_y syy_z sxy_x sxy_y sxz_x sxz_z syz_y syz_z
    // szz_x szz_y szz_z (split components 9..23)
};

template <int N>
__device__ __forceinline__ void load_ssplane(SSPlane<N> &P, const Res &R, const SplitRef &q, uint32_t pc, int k,
                                             int kb, int SY, uint32_t SZ)
{
    const uint32_t vb = pc - (uint32_t)(N * 4 + N * SY) + (uint32_t)(k - kb) * SZ;
    const int c0 = N * 4 + N * SY;
    xtaps_bm<N>(P.vx_x, R.vx, vb, c0);
    taps_fm<N>(P.vx_y, R.vx, vb, c0, SY);
    xtaps_fm<N>(P.vy_x, R.vy, vb, c0);
    taps_bm<N>(P.vy_y, R.vy, vb, c0, SY);
    xtaps_fm<N>(P.vz_x, R.vz, vb, c0);
    taps_fm<N>(P.vz_y, R.vz, vb, c0, SY);
    P.lam = ldu(R.lam, vb, c0);
    P.mu[0] = ldu(R.mu, vb, c0);
    P.mu[1] = ldu(R.mu, vb, c0 + 4);
    P.mu[2] = ldu(R.mu, vb, c0 + SY);
    P.mu[3] = ldu(R.mu, vb, c0 + (int)SZ);
    P.wvx = ldu(R.vx, vb, c0 + N * (int)SZ);
    P.wvy = ldu(R.vy, vb, c0 + N * (int)SZ);
    P.wvz = ldu(R.vz, vb, c0 + (N - 1) * (int)SZ);
}

template <int N, bool LARGE>
__global__ void __launch_bounds__(kThreads) k_stress_shell(sw_ssg3d g, ShellGeo geo, Fold3D fo)
{
    if ((int)blockIdx.x >= geo.nblock) {
        gather3d(g, fo, (int)blockIdx.x - geo.nblock, (int)gridDim.x - geo.nblock);
        return;
    }
    const SCell cl = shell_cell(geo, xcd_logical_block(blockIdx.x, geo.nblock), (int)threadIdx.x);
    const int i = cl.i, j = cl.j;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i < N || i >= nx - N || j < N || j >= ny - N) return;
    const int SY = 4 * nx;
    const uint32_t SZ = 4u * (uint32_t)(nx * ny);
    const float dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    const float hdt = 0.5f * dt;
    float c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = g.c[t];
    const bool nrm_col = i >= N + 1 && j >= N + 1;
    const bool sxy_col = i >= N + 1 && j >= N + 1;
    const bool sxz_col = i < nx - N - 1;
    const bool syz_col = j < ny - N - 1;
    const bool xi = g.px.in[i] != 0, xh = g.px.in_h[i] != 0;
    const bool yi = g.py.in[j] != 0, yh = g.py.in_h[j] != 0;
    const float dxi = g.px.d[i], dxh = g.px.d_h[i], dyi = g.py.d[j], dyh = g.py.d_h[j];
    const int kb = LARGE ? cl.k0 - N : 0;
    const Res R = make_res(g, kb);
    const uint32_t pc = 4u * (uint32_t)(j * nx + i);
    // windows: wvz[t] = vz[k-N+t] (backward), wvx[t] = vx[k-N+1+t], wvy[t] = vy[k-N+1+t] (forward)
    float wvz[2 * N], wvx[2 * N], wvy[2 * N];
    const int k0 = cl.k0;
#pragma unroll
    for (int t = 0; t < 2 * N - 1; ++t) {
        wvz[t] = ld(R.vz, pc + (uint32_t)(k0 - N + t - kb) * SZ);
        wvx[t] = ld(R.vx, pc + (uint32_t)(k0 - N + 1 + t - kb) * SZ);
        wvy[t] = ld(R.vy, pc + (uint32_t)(k0 - N + 1 + t - kb) * SZ);
    }
    SSPlane<N> P;
    load_ssplane<N>(P, R, split_ref(g.split, cl.q0 + (uint32_t)k0 * cl.qk), pc, k0, kb, SY, SZ);
    for (int k = k0; k < cl.k1; ++k) {
        const SplitRef sq = split_ref(g.split, cl.q0 + (uint32_t)k * cl.qk);
        struct { float q[15]; } PQ;
        ld4(PQ.q, sq.blk + 576 + 4 * sq.l);
        ld4(PQ.q + 4, sq.blk + 832 + 4 * sq.l);
        ld4(PQ.q + 8, sq.blk + 1088 + 4 * sq.l);
        ld2(PQ.q + 12, sq.blk + 1344 + 2 * sq.l);
        PQ.q[14] = ld1(sq.blk + 1472 + sq.l);
        SSPlane<N> Q;
        if (k + 1 < cl.k1)
            load_ssplane<N>(Q, R, split_ref(g.split, cl.q0 + (uint32_t)(k + 1) * cl.qk), pc, k + 1, kb, SY, SZ);
        const bool zi = g.pz.in[k] != 0, zh = g.pz.in_h[k] != 0;
        const float dzi = g.pz.d[k], dzh = g.pz.d_h[k];
        const uint32_t p = pc + (uint32_t)(k - kb) * SZ;
        wvz[2 * N - 1] = P.wvz;
        wvx[2 * N - 1] = P.wvx;
        wvy[2 * N - 1] = P.wvy;
        // normal stresses at (i, j, k) with the point lambda, mu
        // (src/Elastic3D_SSG_SPML.py:296-333)
        if (nrm_col && k >= N + 1) {
            const float lam_h = P.lam, mu_h = P.mu[0];
            const float lpm = lam_h + 2.f * mu_h;
            const float dvxx = diff_bm<N>(P.vx_x, c) / dx;
            const float dvyy = diff_bm<N>(P.vy_y, c) / dy;
            const float dvzz = diff_bm<N>(wvz, c) / dz;
            if (xi | yi | zi) {
                const float xn = 1.f - hdt * dxi, yn = 1.f - hdt * dyi, zn = 1.f - hdt * dzi;
                const float xd = 1.f + hdt * dxi, yd = 1.f + hdt * dyi, zd = 1.f + hdt * dzi;
                const float a0 = cn(PQ.q[0], xn, xd, dt * lpm * dvxx);
                const float a1 = cn(PQ.q[1], yn, yd, dt * lam_h * dvyy);
                const float a2 = cn(PQ.q[2], zn, zd, dt * lam_h * dvzz);
                const float b0 = cn(PQ.q[3], xn, xd, dt * lam_h * dvxx);
                const float b1 = cn(PQ.q[4], yn, yd, dt * lpm * dvyy);
                const float b2 = cn(PQ.q[5], zn, zd, dt * lam_h * dvzz);
                const float c0 = cn(PQ.q[12], xn, xd, dt * lam_h * dvxx);
                const float c1 = cn(PQ.q[13], yn, yd, dt * lam_h * dvyy);
                const float c2 = cn(PQ.q[14], zn, zd, dt * lpm * dvzz);
                st_nt(R.sxx, p, a0 + a1 + a2);
                st_nt(R.syy, p, b0 + b1 + b2);
                st_nt(R.szz, p, c0 + c1 + c2);
            } else {
                st_nt(R.sxx, p, ld(R.sxx, p) + (lpm * dvxx + lam_h * (dvzz + dvyy)) * dt);
                st_nt(R.syy, p, ld(R.syy, p) + (lpm * dvyy + lam_h * (dvxx + dvzz)) * dt);
                st_nt(R.szz, p, ld(R.szz, p) + (lpm * dvzz + lam_h * (dvxx + dvyy)) * dt);
            }
        }
        // sxy at (i+1/2, j+1/2, k): :334-350
        if (sxy_col) {
            const float mu_a = 0.25f * (P.mu[2] + P.mu[0] + P.mu[0] + P.mu[1]);
            const float dvyx = diff_fm<N>(P.vy_x, c) / dx;
            const float dvxy = diff_fm<N>(P.vx_y, c) / dy;
            if (xh | yh | zi) {
                const float a = cn(PQ.q[6], 1.f - hdt * dxh, 1.f + hdt * dxh, dt * mu_a * dvyx);
                const float b = cn(PQ.q[7], 1.f - hdt * dyh, 1.f + hdt * dyh, dt * mu_a * dvxy);
                st_nt(R.sxy, p, a + b);
            } else {
                st_nt(R.sxy, p, ld(R.sxy, p) + mu_a * (dvxy + dvyx) * dt);
            }
        }
        // sxz at (i+1/2, j, k+1/2): :352-371
        if (sxz_col && k < nz - N - 1) {
            const float mu_a = 0.25f * (P.mu[0] + P.mu[1] + P.mu[0] + P.mu[2]);
            const float dvzx = diff_fm<N>(P.vz_x, c) / dx;
            const float dvxz = diff_fm<N>(wvx, c) / dz;
            if (xh | yi | zh) {
                const float a = cn(PQ.q[8], 1.f - hdt * dxh, 1.f + hdt * dxh, dt * mu_a * dvzx);
                const float b = cn(PQ.q[9], 1.f - hdt * dzh, 1.f + hdt * dzh, dt * mu_a * dvxz);
                st_nt(R.sxz, p, a + b);
            } else {
                st_nt(R.sxz, p, ld(R.sxz, p) + mu_a * (dvxz + dvzx) * dt);
            }
        }
        // syz at (i, j+1/2, k+1/2): :373-391
        if (syz_col && k < nz - N - 1) {
            const float mu_a = 0.25f * (P.mu[2] + P.mu[0] + P.mu[3] + P.mu[0]);
            const float dvyz = diff_fm<N>(wvy, c) / dz;
            const float dvzy = diff_fm<N>(P.vz_y, c) / dy;
            if (xi | yh | zh) {
                const float a = cn(PQ.q[10], 1.f - hdt * dyh, 1.f + hdt * dyh, dt * mu_a * dvzy);
                const float b = cn(PQ.q[11], 1.f - hdt * dzh, 1.f + hdt * dzh, dt * mu_a * dvyz);
                st_nt(R.syz, p, a + b);
            } else {
                st_nt(R.syz, p, ld(R.syz, p) + mu_a * (dvyz + dvzy) * dt);
            }
        }
        st4(sq.blk + 576 + 4 * sq.l, PQ.q);
        st4(sq.blk + 832 + 4 * sq.l, PQ.q + 4);
        st4(sq.blk + 1088 + 4 * sq.l, PQ.q + 8);
        st2(sq.blk + 1344 + 2 * sq.l, PQ.q + 12);
        st1(sq.blk + 1472 + sq.l, PQ.q[14]);
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

inline bool large_grid(const sw_ssg3d &g) { return 4L * g.nx * g.ny * g.nz > (1L << 32); }

}  // namespace

// Velocity sweep: the shell's z and y slabs, then the middle rows (the x
// strips with the split-free columns between them).
int launch_split_shell_velocity(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st)
{
    SW_DISPATCH_ORDER(g->order, {
        const bool big = large_grid(*g);
        const ShellGeo slabs = make_shell_geo(*g, N, kSlabs), mid = make_shell_geo(*g, N, kMiddleRows);
        if (slabs.nblock > 0) {
            if (big)
                hipLaunchKernelGGL((k_velocity_shell<N, true, false>), dim3(slabs.nblock), dim3(kThreads), 0, st, *g,
                                   slabs, fo);
            else
                hipLaunchKernelGGL((k_velocity_shell<N, false, false>), dim3(slabs.nblock), dim3(kThreads), 0, st,
                                   *g, slabs, fo);
        }
        if (mid.nblock > 0) {
            if (big)
                hipLaunchKernelGGL((k_velocity_shell<N, true, true>), dim3(mid.nblock), dim3(kThreads), 0, st, *g,
                                   mid, fo);
            else
                hipLaunchKernelGGL((k_velocity_shell<N, false, true>), dim3(mid.nblock), dim3(kThreads), 0, st, *g,
                                   mid, fo);
        }
    });
    return launch_status("ssg3d velocity (split-field PML)");
}

// Stress sweep: the whole shell in one launch (with fo's receiver gather);
// the split-free interior is launched by ssg3d.hip.
int launch_split_shell_stress(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st)
{
    SW_DISPATCH_ORDER(g->order, {
        const ShellGeo geo = make_shell_geo(*g, N, kShell);
        const int nb = geo.nblock + (fo.record ? fo.gather : 0);
        if (nb > 0) {
            if (large_grid(*g))
                hipLaunchKernelGGL((k_stress_shell<N, true>), dim3(nb), dim3(kThreads), 0, st, *g, geo, fo);
            else
                hipLaunchKernelGGL((k_stress_shell<N, false>), dim3(nb), dim3(kThreads), 0, st, *g, geo, fo);
        }
    });
    return launch_status("ssg3d stress (split-field PML shell)");
}

// Every (cell, slot) pair whose split values a sweep updates, in its launches'
// logical block order: cell = (k*ny + j)*nx + i, slot its index in the shell
// storage.  sweep 0: the velocity sweep (slab launch, then the x-strip
// columns of the middle-rows launch); 1: the stress sweep's shell launch.
// Host-only: the same shell_cell the kernels run.
long split_shell_map(const sw_ssg3d *g, int sweep, int64_t *cell, int64_t *slot, long max)
{
    long n = 0;
    SW_DISPATCH_ORDER(g->order, {
        const int parts[2] = {sweep == 0 ? kSlabs : kShell, sweep == 0 ? kMiddleRows : -1};
        for (int part : parts) {
            if (part < 0) continue;
            const ShellGeo geo = make_shell_geo(*g, N, part);
            for (int b = 0; b < geo.nblock; ++b)
                for (int t = 0; t < kThreads; ++t) {
                    const SCell c = shell_cell(geo, b, t);
                    if (c.i < N || c.i >= g->nx - N || c.j < N || c.j >= g->ny - N) continue;
                    if (part == kMiddleRows && c.i >= g->px.lo && c.i < g->px.hi) continue;  // split-free column
                    for (int k = c.k0; k < c.k1; ++k) {
                        if (n < max) {
                            cell[n] = ((int64_t)k * g->ny + c.j) * g->nx + c.i;
                            slot[n] = (int64_t)(uint32_t)(c.q0 + (uint32_t)k * c.qk);
                        }
                        ++n;
                    }
                }
        }
    });
    return n;
}

}  // namespace sw
