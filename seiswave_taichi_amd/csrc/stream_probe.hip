// Measurement probes (no counterpart in the reference).
//
// Streaming probe: a kernel that moves exactly the bytes of a sweep's access
// mix (nin arrays read, nout arrays written, every element once) and does no
// stencil work, so its rate is the practical HBM ceiling the sweeps are
// measured against.
//
// Matrix-core probe: chains of f32 16x16x4 MFMA products, the instruction of
// the learned boundary's correction kernel, with four independent
// accumulators per wave and nothing else, so its rate is the practical
// matrix-core ceiling that kernel is measured against.
//
// Each thread handles 4 consecutive floats per array (16-byte loads and
// stores) and walks the arrays with a grid stride; the host picks the grid.
// The streaming probe's loads and stores are non-temporal, the fastest form
// of a pure stream measured on MI355X (11 inputs and 6 outputs of 8 M floats:
// 5.80 TB/s, against 5.00 TB/s with ordinary accesses and 4.96 TB/s with two
// elements per thread in flight, same run).
#include "sw_internal.h"

namespace sw {

namespace {

constexpr int kMaxIn = 20, kMaxOut = 21;

struct StreamArgs {
    const float *in[kMaxIn];
    float *out[kMaxOut];
};

typedef float f4 __attribute__((ext_vector_type(4)));

template <int NIN, int NOUT>
__global__ void __launch_bounds__(256) k_stream(StreamArgs a, long n4)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n4; t += stride) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < NIN; ++q) acc += __builtin_nontemporal_load(&reinterpret_cast<const f4 *>(a.in[q])[t]);
#pragma unroll
        for (int o = 0; o < NOUT; ++o)
            __builtin_nontemporal_store(acc * (float)(o + 1), &reinterpret_cast<f4 *>(a.out[o])[t]);
    }
}

template <int NIN, int NOUT>
void launch(const StreamArgs &a, long n4, int nblock, hipStream_t st)
{
    hipLaunchKernelGGL((k_stream<NIN, NOUT>), dim3(nblock), dim3(256), 0, st, a, n4);
}

// Each wave issues `steps` rounds of 4 independent MFMA products; one float
// per lane of the result is written so the products are kept.
__global__ void __launch_bounds__(256) k_mfma_probe(float *out, int steps)
{
    typedef float mf4 __attribute__((ext_vector_type(4)));
    mf4 acc[4] = {mf4{0.f, 0.f, 0.f, 0.f}, mf4{0.f, 0.f, 0.f, 0.f}, mf4{0.f, 0.f, 0.f, 0.f},
                  mf4{0.f, 0.f, 0.f, 0.f}};
    // lane-distinct operands (lane l supplies A = 1 + l/100, B = 1 - l/200), so
    // the result also shows the operand layout
    const int l = (int)threadIdx.x % 64;
    const float a = 1.f + 0.01f * (float)l, b = 1.f - 0.005f * (float)l;
    for (int i = 0; i < steps; ++i) {
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[k], 0, 0, 0);
    }
    float v = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) v += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = v;
}

}  // namespace

int launch_mfma_probe(float *out, int nblock, int steps, hipStream_t st)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(nblock), dim3(256), 0, st, out, steps);
    return launch_status("matrix-core probe");
}

int launch_stream_probe(const float *const *in, int nin, float *const *out, int nout, long n, int nblock,
                        hipStream_t st)
{
    StreamArgs a{};
    for (int q = 0; q < nin; ++q) a.in[q] = in[q];
    for (int o = 0; o < nout; ++o) a.out[o] = out[o];
    const long n4 = n / 4;
    // the mixes of the two sweeps (velocity: 10 read, 3 written; stress: 11
    // read, 6 written), of their split-field shell cells (velocity: 6
    // stresses, rho and 9 split values read, 3 velocities and the 9 split
    // values written; stress: 3 velocities, lambda, mu and 15 split values
    // read, 6 stresses and the 15 split values written) and of a copy
    if (nin == 10 && nout == 3)
        launch<10, 3>(a, n4, nblock, st);
    else if (nin == 11 && nout == 6)
        launch<11, 6>(a, n4, nblock, st);
    else if (nin == 16 && nout == 12)
        launch<16, 12>(a, n4, nblock, st);
    else if (nin == 20 && nout == 21)
        launch<20, 21>(a, n4, nblock, st);
    else if (nin == 1 && nout == 1)
        launch<1, 1>(a, n4, nblock, st);
    else {
        set_error("stream probe: mix %d read / %d written not instantiated (10/3, 11/6, 16/12, 20/21, 1/1)", nin,
                  nout);
        return SW_EINVAL;
    }
    return launch_status("stream probe");
}

}  // namespace sw
