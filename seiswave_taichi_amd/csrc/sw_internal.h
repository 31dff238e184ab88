// Internal helpers shared by the libseiswave translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/seiswave.h"

namespace sw {

// Planes per z chunk of the 3-D sweeps: a block marches its columns over one
// chunk, and its 32-bit offsets span the chunk plus the stencil halo.
constexpr int kZChunk = 3;

constexpr int kXCD = 8;  // XCDs per MI355X; any value keeps the renumbering a bijection

// Logical block of dispatched block b out of nb.  MI355X hands consecutive
// workgroups of a launch round-robin to its 8 XCDs, each with its own L2; the
// renumbering gives XCD x a contiguous range of logical blocks, so blocks
// that share halo lines share an L2.
__device__ __forceinline__ int xcd_logical_block(int b, int nb)
{
    const int q = nb / kXCD, r = nb % kXCD;
    const int x = b % kXCD, s = b / kXCD;
    return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

// Thread tile of the 2-D kernels: kTile2DX columns by kTile2DY rows.  With
// the XCD renumbering below, on MI355X (16-shot 800x300 gradient, s; 400^2
// step, us): 32x8 0.919, 10.4; 64x4 0.816, 10.1; 128x2 0.833, 10.7; 256x1
// 0.946, 10.0.
constexpr int kTile2DX = 64, kTile2DY = 4;

// Block coordinates of a launch grid of 3 dimensions with the XCD
// renumbering applied to its x-fastest linear order (a 2-D tile grid per
// shot, shots along z): the tiles above and below a tile, whose rows its
// z stencil reads, run on the same XCD.  On MI355X the forward sweeps of 16
// shots of 800x300 ran 59 against 46 Gcells/s in dispatch order, a 400^2
// step 10.0 against 11.6 us.
struct Blk {
    int x, y, z;
};

__device__ __forceinline__ Blk xcd_block3()
{
    Blk r;
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int b = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
    const int l = xcd_logical_block(b, gx * gy * (int)gridDim.z);
    const int t = l / gx;
    r.x = l - t * gx;
    r.z = t / gy;
    r.y = t - r.z * gy;
    return r;
}

// Per-thread message of the last failure (read back by sw_last_error).
void set_error(const char *fmt, ...);

// Compute units of the current device, looked up once per device (256, the
// MI355X's count, where the query fails, e.g. on a host without a GPU).
int compute_units();

// The launch plan mode set by sw_set_launch_plan (SW_PLAN_*).
int launch_plan_mode();

#define SW_TRY_HIP(expr)                                                              \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            ::sw::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                            __FILE__, __LINE__);                                      \
            return SW_EHIP;                                                           \
        }                                                                             \
    } while (0)

#define SW_REQUIRE(cond, ...)                                                         \
    do {                                                                              \
        if (!(cond)) {                                                                \
            ::sw::set_error(__VA_ARGS__);                                             \
            return SW_EINVAL;                                                         \
        }                                                                             \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Checks the launch that was just enqueued.
inline int launch_status(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return SW_EHIP;
    }
    return SW_OK;
}

// Validation shared by the 3-D entry points.
int check_ssg3d(const sw_ssg3d *g);  // class-specific boundary arrays follow g->cls
int check_ssg2d(const sw_ssg2d *g);

// Slab width of an axis: indices [0,lo) and [hi,n) carry memory variables.
inline int slab_width(const sw_axis &a, int n) { return a.lo + (n - a.hi); }

// Index of a point inside its axis slab, or -1 outside the slab.
__device__ __forceinline__ int slab_index(int i, int lo, int hi)
{
    return i < lo ? i : (i >= hi ? lo + (i - hi) : -1);
}

// Shell storage of the split-field PML.  The reference keeps 24 full-grid
// split arrays but reads and writes them only where a component's in-PML
// test holds, i.e. outside the split-free box [xlo,xhi) x [ylo,yhi) x
// [zlo,zhi) (the axes' lo/hi).  Each component here holds only the S cells
// of the shell, in three parts:
//   z part  [wz][ny][nx]  planes k < zlo or k >= zhi (plane index zs(k));
//   y part  [mz][wy][nx]  planes zlo <= k < zhi, rows j < ylo or j >= yhi;
//   x part  [mz][my][wx]  those planes, rows ylo <= j < yhi, columns outside
//                         [xlo, xhi);
// with w_d = n_d - (hi_d - lo_d), m_d = hi_d - lo_d and slab index
// s(i) = i < lo ? i : i - (hi - lo).  An axis with lo = hi has no split-free
// range: its slab is the whole axis.  Cells are numbered ("slots") through
// the three parts in that order.  The 24 components of 64 consecutive slots
// share one block of 1536 floats, in groups that a sweep thread moves with
// one 16-, 8- or 4-byte access each (float offsets in the block, lane
// l = slot mod 64):
//   comps 0-3 at 0 + 4l, 4-7 at 256 + 4l, 8 at 512 + l        (velocity sweep)
//   comps 9-12 at 576 + 4l, 13-16 at 832 + 4l, 17-20 at 1088 + 4l,
//   21-22 at 1344 + 2l, 23 at 1472 + l                          (stress sweep)
constexpr int kSplitBlock = 64, kSplitBlockFloats = 24 * kSplitBlock;

struct SplitShell {
    int xlo, xhi, ylo, yhi, zlo, zhi;
    int wx, wy, wz, my, mz;
    long S, OZ, OY;  // cells per component; first cell of the y and x parts
};

inline SplitShell split_shell(const sw_ssg3d &g)
{
    SplitShell s;
    s.xlo = g.px.lo; s.xhi = g.px.hi;
    s.ylo = g.py.lo; s.yhi = g.py.hi;
    s.zlo = g.pz.lo; s.zhi = g.pz.hi;
    s.wx = g.nx - (s.xhi - s.xlo);
    s.wy = g.ny - (s.yhi - s.ylo);
    s.wz = g.nz - (s.zhi - s.zlo);
    s.my = s.yhi - s.ylo;
    s.mz = s.zhi - s.zlo;
    s.OZ = (long)s.wz * g.ny * g.nx;
    s.OY = s.OZ + (long)s.mz * s.wy * g.nx;
    s.S = s.OY + (long)s.mz * s.my * s.wx;
    return s;
}

// Updated cells [x0,x1) x [y0,y1) x [z0,z1) (= [N, n-N) per axis) and the
// split-free box clamped into them, [xl,xh) x [yl,yh) x [zl,zh) (empty on an
// axis without a split-free range).  The shell launch covers the updated
// cells outside that box: planes [z0,zl) and [zh,z1); in [zl,zh) the rows
// outside [yl,yh); in those rows the columns outside [xl,xh).
struct SplitBox {
    int x0, x1, y0, y1, z0, z1;
    int xl, xh, yl, yh, zl, zh;
};

inline SplitBox split_box(const sw_ssg3d &g, int N)
{
    auto clamp = [](int v, int a, int b) { return v < a ? a : (v > b ? b : v); };
    SplitBox b;
    b.x0 = N; b.x1 = g.nx - N;
    b.y0 = N; b.y1 = g.ny - N;
    b.z0 = N; b.z1 = g.nz - N;
    b.xl = clamp(g.px.lo, b.x0, b.x1); b.xh = clamp(g.px.hi, b.xl, b.x1);
    b.yl = clamp(g.py.lo, b.y0, b.y1); b.yh = clamp(g.py.hi, b.yl, b.y1);
    b.zl = clamp(g.pz.lo, b.z0, b.z1); b.zh = clamp(g.pz.hi, b.zl, b.z1);
    return b;
}

// ---------------------------------------------------------------------------
// Staggered first derivatives (src/Differential3D.py:7-44): the accumulator
// starts at 0 and adds c[id]*(difference) in id order, so the rounding matches
// the reference term for term (the library is built with -ffp-contract=off).
template <int N>
__device__ __forceinline__ float d_fm(const float *__restrict__ f, long p, long st, const float *c)
{
    float v = 0.f;
#pragma unroll
    for (int id = 0; id < N; ++id) v += c[id] * (f[p + (id + 1) * st] - f[p - id * st]);
    return v;
}

template <int N>
__device__ __forceinline__ float d_bm(const float *__restrict__ f, long p, long st, const float *c)
{
    float v = 0.f;
#pragma unroll
    for (int id = 0; id < N; ++id) v += c[id] * (f[p + id * st] - f[p - (id + 1) * st]);
    return v;
}

// CPML memory-variable step and derivative correction
// (src/Elastic3D_SSG_CPML.py:172-177): m = b*m + a*D;  D = D/k + m.
__device__ __forceinline__ float cpml_apply(float D, float *m, float a, float b, float k)
{
    float mm = b * (*m) + a * D;
    *m = mm;
    return D / k + mm;
}

// Shot batches of the 2-D path: shot b of a batch keeps its state (and its
// adjoint) at the descriptor's pointers plus b * stride floats; the model,
// PML coefficients and flags are shared.
__device__ __forceinline__ void shift_state(sw_ssg2d &g, long off)
{
    g.vx += off; g.vz += off; g.sxx += off; g.szz += off; g.sxz += off; g.mem += off;
}

__device__ __forceinline__ void shift_adj(sw_adj2d &q, long off)
{
    q.vx += off; q.vz += off; q.sxx += off; q.szz += off; q.sxz += off; q.mem += off; q.tmp += off;
    q.g_b1 += off; q.g_b4 += off; q.g_lamh += off; q.g_muh += off; q.g_muz += off;
}

// Point sources of the 3-D step at cell p, shared by the injection kernel and
// the sweeps that fold the injection in: a force adds dt*amp/rho[p] to a
// velocity after the velocity sweep (src/Elastic3D_SSG_CPML.py:245,
// src/Elastic3D_SSG_SPML.py:294); a moment tensor adds amp*(-M_ab) to the six
// stresses before it (src/Elastic3D_SSG_Tensor.py:186-193), an explosive
// source amp*(-1/sqrt(3)) to the normal ones.
__device__ __forceinline__ void force_src3d(const sw_ssg3d &g, int kind, float amp, long p)
{
    if (kind == SW_SRC_FORCE_VZ)
        g.vz[p] += g.dt * amp / g.rho[p];
    else
        g.vx[p] += g.dt * amp / g.rho[p];
}

__device__ __forceinline__ void stress_src3d(const sw_ssg3d &g, int kind, const float *mt, float amp, long p)
{
    if (kind == SW_SRC_MOMENT) {
        g.sxx[p] += amp * (-mt[0]);
        g.syy[p] += amp * (-mt[1]);
        g.szz[p] += amp * (-mt[2]);
        g.sxy[p] += amp * (-mt[3]);
        g.sxz[p] += amp * (-mt[4]);
        g.syz[p] += amp * (-mt[5]);
    } else {
        const float e = amp * (-0.57735026918962576f);
        g.sxx[p] += e;
        g.syy[p] += e;
        g.szz[p] += e;
    }
}

// Work of a 3-D step folded into its sweeps, as Fold2D: a force source of
// this step (velocity-sweep thread of the source cell, after its update), a
// moment / explosive source of the NEXT step (stress-sweep thread of the
// source cell, after its update) and the velocity receivers of this step
// (gathered by `gather` extra blocks appended to the first stress launch).
struct Fold3D {
    int force, it_force;
    int stress, it_next;
    const float *amp;
    float mt[6];
    int ix, iy, iz;
    int record, it_rec, gather;
    sw_receivers rec;
};

// Split-field PML shell of the 3-D sweeps (ssg3d_split.hip): one launch per
// sweep over the updated cells outside the split-free box (the stress launch
// carries fo's receiver gather).
int launch_split_shell_velocity(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st);
int launch_split_shell_stress(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st);

// Point sources of the 2-D step at cell p, shared by the injection kernel and
// the sweeps that fold the injection in (the same expressions, so the same
// rounded values): a force (src/Elastic2D_CPML.py:170, :245) adds
// dt*amp/rho[p] to a velocity after the velocity sweep; an explosive or
// moment-tensor source adds amp*(-M) to the stresses before it (the 2-D
// moment tensor is mt[0], mt[2], mt[4] = Mxx, Mzz, Mxz).
__device__ __forceinline__ void force_src2d(sw_ssg2d &g, int kind, float amp, long p)
{
    if (kind == SW_SRC_FORCE_VZ)
        g.vz[p] += g.dt * amp / g.rho[p];
    else
        g.vx[p] += g.dt * amp / g.rho[p];
}

__device__ __forceinline__ void stress_src2d(sw_ssg2d &g, int kind, float m0, float m2, float m4, float amp, long p)
{
    if (kind == SW_SRC_MOMENT) {
        g.sxx[p] += amp * (-m0);
        g.szz[p] += amp * (-m2);
        g.sxz[p] += amp * (-m4);
    } else {
        const float e = amp * (-0.70710678118654752f);
        g.sxx[p] += e;
        g.szz[p] += e;
    }
}

// Work of a 2-D step folded into its two sweeps (on small grids the separate
// one-thread and one-block launches cost a quarter of a step):
//  * force: a force source of this step, added by the velocity-sweep thread
//    of the source cell after its own update (no other thread of that sweep
//    touches the cell);
//  * stress: an explosive / moment source of the NEXT step, added by the
//    stress-sweep thread of the source cell after its own update (no other
//    thread of that sweep reads stresses), i.e. right before that step;
//  * record: the velocity receivers of this step, gathered by the first
//    block of each shot in the stress sweep (velocities are final by then).
// Shot b's source cell is (ix[b], iz[b]); its traces start b * ds floats
// after rec.data.  A zero field switches a part off.
struct Fold2D {
    int force, it_force;
    int stress, it_next;
    const float *amp;
    float m0, m2, m4;
    int ix[SW_MAX_BATCH], iz[SW_MAX_BATCH];
    int record, it_rec;
    sw_receivers rec;
    long ds;
};

}  // namespace sw

// Dispatch on the stencil half-length N (accuracy) with a compile-time constant.
#define SW_DISPATCH_ORDER(order, ...)                                                 \
    switch (order) {                                                                  \
    case 1: { constexpr int N = 1; __VA_ARGS__; } break;                              \
    case 2: { constexpr int N = 2; __VA_ARGS__; } break;                              \
    case 3: { constexpr int N = 3; __VA_ARGS__; } break;                              \
    case 4: { constexpr int N = 4; __VA_ARGS__; } break;                              \
    case 5: { constexpr int N = 5; __VA_ARGS__; } break;                              \
    case 6: { constexpr int N = 6; __VA_ARGS__; } break;                              \
    case 7: { constexpr int N = 7; __VA_ARGS__; } break;                              \
    case 8: { constexpr int N = 8; __VA_ARGS__; } break;                              \
    default: ::sw::set_error("order %d outside 1..%d", (int)(order), SW_MAX_ORDER);    \
        return SW_EINVAL;                                                             \
    }
