// Device helpers of the 3-D sweep kernels (ssg3d.hip, ssg3d_split.hip):
// addressing, stencil taps (loaded, or shifted in from neighbouring lanes),
// staggered differences, CPML slabs and the velocity-sweep register plane.  They sit in an anonymous namespace, so a translation unit that
// includes them keeps its own copy.
#pragma once

#include "sw_internal.h"

namespace sw {

namespace {

// A launch covers one box of cells: columns [i0, i0+w) x [j0, j0+h) over
// planes [k0, k1).  The box is cut into blocks of kThreads consecutive
// columns in x-fastest order times chunks of kZChunk planes.  The plain
// sweeps count columns by flat index j*nx + i from a 128-byte line boundary
// (lanes on the x edges of a row idle); the split-field shell counts only
// the shell's own columns.
constexpr int kThreads = 256;

// Blocks run in groups of kGroup column blocks (8,192 columns) that march
// through all z chunks before the next group starts.  With short chunks a
// chunk's window planes are the ones the same group read one chunk earlier,
// still in the XCD's L2, while a group is wide enough that its y halo rows
// are a small part of its reads.  Measured on MI355X, ms per step with groups
// of 4 / 8 / 16 / 24 / 32 / 48 / 64 / 96 column blocks: 200^3 0.327 / 0.322 /
// 0.322 / 0.313 / 0.305 / 0.308 / 0.305 / 0.307, against 0.318 in
// plane-by-plane order; 600x600x400 4.77 (32), 4.73 (64), 4.91 plane by plane.
// With line-aligned columns, groups of 16 / 32 / 64: 200^3 0.301 / 0.289 /
// 0.290, 600x600x400 4.57 / 4.42 / 4.38, 200^3 split-field class 0.431 /
// 0.425 / 0.422; groups of 32 / 64 / 128 / 256 on 600x600x400 4.429 /
// 4.385 / 4.374 / 4.414 and on its 600x600x54 shard 0.577 / 0.581 / 0.577 /
// 0.589 (one group size for every grid: the gains elsewhere stay near 1 %).
// With the x taps from neighbouring lanes (4 waves per SIMD), chunks of 3 / 4
// / 5 planes in groups of 32 / 48 / 64 column blocks, ms per step (one
// process, bit-identical): 200^3 4x32 0.2744, 3x32 0.2723, 4x64 0.2749,
// 3x64 0.2717, 3x48 0.2732, 4x48 0.2778, 5x32 0.2783 (another box); 600x600x400
// 4.141 / 4.136 / 4.069 / 4.052 / 4.090 / 4.097; split-field class 0.4161 /
// 0.4043 / 0.4125 / 0.4024 / 0.4043 / 0.4151; 600x600x54 shard 0.5341 /
// 0.5303 / 0.5335 / 0.5312 / 0.5312 / 0.5340: 3-plane chunks in groups of 64.
constexpr int kGroup = 64;

// Fields are addressed by 32-bit byte offsets from their base pointers: one
// offset register per column serves every field, and the wave-uniform tap
// deltas fold into the load instructions.  A block rebases the full-grid
// pointers to the first plane its z chunk reads (make_res), so their offsets
// only span the chunk's planes plus the stencil halo, whatever the array
// size; the host checks that this span, and each memory slab, stays below
// 4 GiB.
__device__ __forceinline__ float ld(const float *__restrict__ base, uint32_t off)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + off);
}

__device__ __forceinline__ void st(float *__restrict__ base, uint32_t off, float v)
{
    *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + off) = v;
}

// Load at byte offset v + d, d a wave-uniform non-negative delta.
__device__ __forceinline__ float ldu(const float *__restrict__ base, uint32_t v, int d)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + v + (uint32_t)d);
}

// The fields a sweep updates in place (the velocities in the velocity sweep,
// the stresses in the stress sweep) are read once and written once per sweep
// by the thread that owns the cell: their old values are loaded and their
// new values stored non-temporally, which keeps them from displacing the
// stencil's lines in L2 (measured on MI355X, ms per step, bit-identical:
// 200^3 0.2894 -> 0.2854, 600x600x400 4.434 -> 4.345, split-field class
// 0.4257 -> 0.4182; non-temporal loads with ordinary stores ran 0.2999, and
// the CPML memory variables loaded non-temporally too 0.2865 and 4.485).
__device__ __forceinline__ float ldu_nt(const float *__restrict__ base, uint32_t v, int d)
{
    return __builtin_nontemporal_load(reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + v + (uint32_t)d));
}

__device__ __forceinline__ void st_nt(float *__restrict__ base, uint32_t off, float v)
{
    __builtin_nontemporal_store(v, reinterpret_cast<float *>(reinterpret_cast<char *>(base) + off));
}

// Base pointers of all the arrays a 3-D sweep touches.
struct Res {
    float *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *sxz, *syz, *mx, *my, *mz;
    const float *rho, *lam, *mu;
};

// Full-grid pointers rebased to plane kb.  The memory slabs keep their own
// base (their 32-bit offsets span a whole slab, which the host keeps below
// 4 GiB).
__device__ __forceinline__ Res make_res(const sw_ssg3d &g, int kb)
{
    const long pl = (long)kb * g.nx * g.ny;
    Res r;
    r.vx = g.vx + pl; r.vy = g.vy + pl; r.vz = g.vz + pl;
    r.sxx = g.sxx + pl; r.syy = g.syy + pl; r.szz = g.szz + pl;
    r.sxy = g.sxy + pl; r.sxz = g.sxz + pl; r.syz = g.syz + pl;
    r.rho = g.rho + pl; r.lam = g.lam + pl; r.mu = g.mu + pl;
    r.mx = g.mem_x; r.my = g.mem_y; r.mz = g.mem_z;
    return r;
}

// Tap arrays: a "bm" array holds f[p + (t-N)*st] (offsets -N..N-1), an "fm"
// array f[p + (t-N+1)*st] (offsets -N+1..N), t = 0..2N-1.  All loads of a
// plane are addressed from vb = byte offset of (i-N, j-N, k); c0 is the
// uniform delta of (i, j, k) from there and st the tap stride, in bytes.
template <int N>
__device__ __forceinline__ void taps_bm(float *w, const float *__restrict__ f, uint32_t vb, int c0, int st)
{
#pragma unroll
    for (int t = 0; t < 2 * N; ++t) w[t] = ldu(f, vb, c0 + (t - N) * st);
}

template <int N>
__device__ __forceinline__ void taps_fm(float *w, const float *__restrict__ f, uint32_t vb, int c0, int st)
{
#pragma unroll
    for (int t = 0; t < 2 * N; ++t) w[t] = ldu(f, vb, c0 + (t - N + 1) * st);
}

// The same without the centre (offset 0: w[N] of a bm array, w[N-1] of an
// fm array), for fields whose z window already holds it; the kernel fills
// it in before use.
template <int N>
__device__ __forceinline__ void taps_bm_nc(float *w, const float *__restrict__ f, uint32_t vb, int c0, int st)
{
#pragma unroll
    for (int t = 0; t < 2 * N; ++t)
        if (t != N) w[t] = ldu(f, vb, c0 + (t - N) * st);
}

template <int N>
__device__ __forceinline__ void taps_fm_nc(float *w, const float *__restrict__ f, uint32_t vb, int c0, int st)
{
#pragma unroll
    for (int t = 0; t < 2 * N; ++t)
        if (t != N - 1) w[t] = ldu(f, vb, c0 + (t - N + 1) * st);
}

// x taps of the split-field shell kernels (ssg3d_split.hip), whose lanes do
// not always hold consecutive columns: 2N consecutive floats, one 16-byte (or
// 8-byte) load per 4 (or 2) taps instead of one load each.  The addresses
// are only 4-byte aligned, which global vector loads accept.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
struct __attribute__((packed, aligned(4))) F4U { f4v v; };
struct __attribute__((packed, aligned(4))) F2U { f2v v; };

template <int M>
__device__ __forceinline__ void ld_run(float *w, const float *__restrict__ f, uint32_t v, int d)
{
    const char *b = reinterpret_cast<const char *>(f) + v + (uint32_t)d;
    int t = 0;
#pragma unroll
    for (; t + 4 <= M; t += 4) {
        const f4v x = reinterpret_cast<const F4U *>(b + 4 * t)->v;
        w[t] = x.x; w[t + 1] = x.y; w[t + 2] = x.z; w[t + 3] = x.w;
    }
#pragma unroll
    for (; t + 2 <= M; t += 2) {
        const f2v x = reinterpret_cast<const F2U *>(b + 4 * t)->v;
        w[t] = x.x; w[t + 1] = x.y;
    }
#pragma unroll
    for (; t < M; ++t) w[t] = *reinterpret_cast<const float *>(b + 4 * t);
}

template <int N>
__device__ __forceinline__ void xtaps_bm(float *w, const float *__restrict__ f, uint32_t vb, int c0)
{
    ld_run<2 * N>(w, f, vb, c0 - N * 4);
}

template <int N>
__device__ __forceinline__ void xtaps_fm(float *w, const float *__restrict__ f, uint32_t vb, int c0)
{
    ld_run<2 * N>(w, f, vb, c0 + (1 - N) * 4);
}

// x taps from the neighbouring lanes (the plain sweeps).  A wave's 64 lanes
// hold 64 consecutive columns of a row-major plane, so the x neighbour at
// distance d of lane l sits in lane l - d or l + d, except past lane 0 or 63.
// Each lane keeps its own value; wave_shr:1 (DPP control 0x138) gives lane l
// the value of lane l - 1 and wave_shl:1 (0x130) that of lane l + 1, and the
// lane without a source (lane 0, resp. 63) takes the instruction's `old`
// operand.  The values lanes 0 and 63 lack are wave-uniform, so `old` is a
// scalar: the field at lane 0's offset minus 4d or lane 63's plus 4d, read
// through the scalar cache (s_load, no vector-memory instruction; the sweep
// never writes the arrays it reads this way).  Every lane of the wave must
// be active at the shifts, and every lane whose value a cell uses holds its
// own column's value (the kernels keep idle lanes running).  This replaces
// each field's 16-byte x-tap load and the p+1 loads of the material, and the
// register plane no longer carries the x taps.
__device__ __forceinline__ float from_prev_lane(float v, float old)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ float from_next_lane(float v, float old)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// Wave-uniform load at byte offset off through the scalar cache.
typedef const __attribute__((address_space(4))) float sfloat;
__device__ __forceinline__ float lds_u(const float *__restrict__ base, uint32_t off)
{
    return *(sfloat *)(reinterpret_cast<const char *>(base) + off);  // address-space cast
}

// Edge offsets of a wave: the byte offsets of lane 0's and lane 63's cells
// (call with every lane active).  Both columns have N rows of the plane on
// either side, so the N floats before and after them are inside the array.
struct Edge {
    uint32_t lo, hi;
};

__device__ __forceinline__ Edge wave_edge(uint32_t v)
{
    return Edge{(uint32_t)__builtin_amdgcn_readfirstlane(v), (uint32_t)__builtin_amdgcn_readlane(v, 63)};
}

// bm taps f[q + t - N], fm taps f[q + t - N + 1], t = 0..2N-1, of field f at
// the plane whose edge offsets are e; own = f[q].
template <int N>
__device__ __forceinline__ void xtaps_bm_lanes(float *w, float own, const float *__restrict__ f, Edge e)
{
    w[N] = own;
#pragma unroll
    for (int t = N - 1; t >= 0; --t) w[t] = from_prev_lane(w[t + 1], lds_u(f, e.lo - 4u * (N - t)));
#pragma unroll
    for (int t = N + 1; t < 2 * N; ++t) w[t] = from_next_lane(w[t - 1], lds_u(f, e.hi + 4u * (t - N)));
}

template <int N>
__device__ __forceinline__ void xtaps_fm_lanes(float *w, float own, const float *__restrict__ f, Edge e)
{
    w[N - 1] = own;
#pragma unroll
    for (int t = N - 2; t >= 0; --t) w[t] = from_prev_lane(w[t + 1], lds_u(f, e.lo - 4u * (N - 1 - t)));
#pragma unroll
    for (int t = N; t < 2 * N; ++t) w[t] = from_next_lane(w[t - 1], lds_u(f, e.hi + 4u * (t - N + 1)));
}

// f[q + 1] of a value the lane holds as own = f[q].
__device__ __forceinline__ float next_x(float own, const float *__restrict__ f, Edge e)
{
    return from_next_lane(own, lds_u(f, e.hi + 4u));
}

// Staggered differences from tap arrays, accumulated exactly as
// src/Differential3D.py:7-44 does (0, then c[id]*(difference) in id order).
template <int N>
__device__ __forceinline__ float diff_bm(const float *w, const float *c)
{
    float v = 0.f;
#pragma unroll
    for (int id = 0; id < N; ++id) v += c[id] * (w[N + id] - w[N - id - 1]);
    return v;
}

template <int N>
__device__ __forceinline__ float diff_fm(const float *w, const float *c)
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

template <int M>
__device__ __forceinline__ void copy_n(float *d, const float *s)
{
#pragma unroll
    for (int t = 0; t < M; ++t) d[t] = s[t];
}

// Byte offsets of one column into the CPML slab arrays.
struct Slab {
    int xs, ys;
    uint32_t rx, ry;      // x / y slab offset at k = 0
    uint32_t sx, sy, sz;  // per-k strides of the x / y slab arrays; plane size of the z slab
    uint32_t nxs, nys, nzs;
};

template <int PM>
__device__ __forceinline__ Slab slab_setup(const sw_ssg3d &g, int i, int j)
{
    Slab s;
    const int wx = g.px.lo + (g.nx - g.px.hi);
    const int wy = g.py.lo + (g.ny - g.py.hi);
    const int wz = g.pz.lo + (g.nz - g.pz.hi);
    s.xs = (PM & 1) ? slab_index(i, g.px.lo, g.px.hi) : -1;
    s.ys = (PM & 2) ? slab_index(j, g.py.lo, g.py.hi) : -1;
    s.rx = 4u * (uint32_t)(j * wx + s.xs);
    s.ry = 4u * (uint32_t)(s.ys * g.nx + i);
    s.sx = 4u * (uint32_t)(g.ny * wx);
    s.sy = 4u * (uint32_t)(wy * g.nx);
    s.sz = 4u * (uint32_t)(g.ny * g.nx);
    s.nxs = s.sx * (uint32_t)g.nz;
    s.nys = s.sy * (uint32_t)g.nz;
    s.nzs = s.sz * (uint32_t)wz;
    return s;
}

// CPML coefficients of one point: integer (a, b, k) and half (ah, bh, kh).
struct Coef {
    float a, b, k, ah, bh, kh;
};

__device__ __forceinline__ Coef coef_of(const sw_axis &ax, int i, bool in_slab)
{
    Coef c{0.f, 1.f, 1.f, 0.f, 1.f, 1.f};
    if (in_slab) {
        c.a = ax.a[i]; c.b = ax.b[i]; c.k = ax.k[i];
        c.ah = ax.a_h[i]; c.bh = ax.b_h[i]; c.kh = ax.k_h[i];
    }
    return c;
}

// m = b*m + a*D; D = D/k + m  (src/Elastic3D_SSG_CPML.py:172-177)
__device__ __forceinline__ float cpml_step(float D, float &m, float a, float b, float k)
{
    m = b * m + a * D;
    return D / k + m;
}

// ------------------------------------------------------------ velocity sweep
template <int N>
struct VPlane {
    float sxx;          // sxx at p (its x taps come from the neighbouring lanes)
    float sxy_y[2 * N]; // bm taps along y (Dy3bm, vx)
    float syy[2 * N];   // fm taps along y (Dy3fm, vy)
    float syz_y[2 * N]; // bm taps along y (Dy3bm, vz); centre from the window
    float rho[2];       // p, p+SY (p+1, p+1+SY from the next lane)
    float v[3];
    float wxz, wyz, wzz; // next window entries: sxz, syz at k+N-1, szz at k+N
    float mx[3], my[3], mz[3];
    Coef cz;
    int zs;
};

// pc: byte offset of (i, j) in a plane; SY, SZ: byte strides; R rebased to
// plane kb.
template <int N, int PM>
__device__ __forceinline__ void load_vplane(VPlane<N> &P, const sw_ssg3d &g, const Res &R, uint32_t pc, int k, int kb,
                                            int SY, uint32_t SZ, const Slab &sl)
{
    // vb: byte offset of (i-N, j-N, k); every load below is vb + a uniform delta
    const uint32_t vb = pc - (uint32_t)(N * 4 + N * SY) + (uint32_t)(k - kb) * SZ;
    const int c0 = N * 4 + N * SY;
    P.sxx = ldu(R.sxx, vb, c0);
    taps_bm<N>(P.sxy_y, R.sxy, vb, c0, SY);
    taps_fm<N>(P.syy, R.syy, vb, c0, SY);
    taps_bm_nc<N>(P.syz_y, R.syz, vb, c0, SY);  // centre: the syz window
    P.rho[0] = ldu(R.rho, vb, c0);
    P.rho[1] = ldu(R.rho, vb, c0 + SY);
    P.v[0] = ldu_nt(R.vx, vb, c0);
    P.v[1] = ldu_nt(R.vy, vb, c0);
    P.v[2] = ldu_nt(R.vz, vb, c0);
    P.wxz = ldu(R.sxz, vb, c0 + (N - 1) * (int)SZ);
    P.wyz = ldu(R.syz, vb, c0 + (N - 1) * (int)SZ);
    P.wzz = ldu(R.szz, vb, c0 + N * (int)SZ);
    P.zs = (PM & 4) ? slab_index(k, g.pz.lo, g.pz.hi) : -1;
    P.cz = coef_of(g.pz, k, P.zs >= 0);
    const uint32_t ox = sl.rx + (uint32_t)k * sl.sx, oy = sl.ry + (uint32_t)k * sl.sy;
    const uint32_t oz = (uint32_t)P.zs * sl.sz + pc;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        P.mx[m] = sl.xs >= 0 ? ldu(R.mx, ox, (int)(m * sl.nxs)) : 0.f;
        P.my[m] = sl.ys >= 0 ? ldu(R.my, oy, (int)(m * sl.nys)) : 0.f;
        P.mz[m] = P.zs >= 0 ? ldu(R.mz, oz, (int)(m * sl.nzs)) : 0.f;
    }
}

// ------------------------------------------------------------ receivers
// Velocity receivers of this step, gathered by the nb blocks appended to the
// first stress launch: data[(it*ncomp + comp)*nrec + r].
__device__ void gather3d(const sw_ssg3d &g, const Fold3D &fo, int b, int nb)
{
    const sw_receivers &r = fo.rec;
    const int n = r.nrec * r.ncomp;
    float *out = r.data + (long)fo.it_rec * n;
    for (int t = b * kThreads + (int)threadIdx.x; t < n; t += nb * kThreads) {
        const int comp = t / r.nrec, q = t - comp * r.nrec;
        const int code = r.comp[comp];
        out[t] = (code == 0 ? g.vx : code == 1 ? g.vy : g.vz)[r.idx[q]];
    }
}

}  // namespace

}  // namespace sw
