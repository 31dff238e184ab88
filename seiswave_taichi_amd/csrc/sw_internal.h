// Internal helpers shared by the libseiswave translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/seiswave.h"

namespace sw {

// Per-thread message of the last failure (read back by sw_last_error).
void set_error(const char *fmt, ...);

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
int check_ssg3d(const sw_ssg3d *g, bool spml);
int check_ssg2d(const sw_ssg2d *g);

// Slab width of an axis: indices [0,lo) and [hi,n) carry memory variables.
inline int slab_width(const sw_axis &a, int n) { return a.lo + (n - a.hi); }

// Index of a point inside its axis slab, or -1 outside the slab.
__device__ __forceinline__ int slab_index(int i, int lo, int hi)
{
    return i < lo ? i : (i >= hi ? lo + (i - hi) : -1);
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
