// Learned thin absorbing boundary: a per-boundary-column MLP that corrects
// the fields next to a thin CPML (no counterpart in the reference; it
// replaces most of a wide CPML).
//
// Geometry.  The "interior" is the box of cells outside every CPML slab.  A
// band is the D = 4 outermost interior cells along one face normal; each band
// column (one lateral position of one face) is one MLP sample.  Faces are
// numbered x-lo, x-hi, y-lo, y-hi, z-lo, z-hi; x faces span the full lateral
// y-z extent, y faces exclude the x bands, z faces exclude the x and y bands,
// so every band cell belongs to exactly one column.
//
// Features of a column: the 9 fields (vx vy vz sxx syy szz sxy sxz syz) at
// the D band cells after the current step (cur), then cur minus the same 36
// values after the previous step (as corrected), i.e. 72 = 9 x 4 x 2; band
// cell d = 0 touches the CPML.  Output: 36 corrections added to cur.
//
// Network of one face (params, row-major):
//   W1[72][64] b1[64] W2[64][64] b2[64] W3[64][36] b3[36] in_scale[72] out_scale[36]
//   u = x / in_scale,  s = RMS(u) (over the 72 entries),
//   y = (W3^T tanh(W2^T tanh(W1^T (u / s) + b1) + b2) + b3) * out_scale * s
// (y = 0 where s = 0).  The training entry points see samples already divided
// by s: their forward pass is the bracket applied to x / in_scale.
#include <algorithm>
#include <mutex>

#include "sw_internal.h"

namespace sw {

namespace {

constexpr int kD = 4, kF = 9, kIn = 2 * kD * kF, kOut = kD * kF, kH = 64;
constexpr int kParams = kIn * kH + kH + kH * kH + kH + kH * kOut + kOut + kIn + kOut;

struct BandGeo {
    int ox, oy, oz, ex, ey, ez;
    int ncol[6];
    int col0[7];
};

BandGeo band_geo(const sw_band &b)
{
    BandGeo q;
    q.ox = b.ox; q.oy = b.oy; q.oz = b.oz; q.ex = b.ex; q.ey = b.ey; q.ez = b.ez;
    const int ex2 = b.ex - 2 * kD, ey2 = b.ey - 2 * kD;
    q.ncol[0] = q.ncol[1] = b.ey * b.ez;
    q.ncol[2] = q.ncol[3] = ex2 * b.ez;
    q.ncol[4] = q.ncol[5] = ex2 * ey2;
    q.col0[0] = 0;
    for (int f = 0; f < 6; ++f) q.col0[f + 1] = q.col0[f] + q.ncol[f];
    return q;
}

// Cell (i, j, k) of band depth d of column `a, b` of face f.
__device__ __forceinline__ void band_cell(const BandGeo &q, int f, int a, int b, int d, int &i, int &j, int &k)
{
    switch (f) {
    case 0: i = q.ox + d; j = q.oy + a; k = q.oz + b; break;
    case 1: i = q.ox + q.ex - 1 - d; j = q.oy + a; k = q.oz + b; break;
    case 2: i = q.ox + kD + a; j = q.oy + d; k = q.oz + b; break;
    case 3: i = q.ox + kD + a; j = q.oy + q.ey - 1 - d; k = q.oz + b; break;
    case 4: i = q.ox + kD + a; j = q.oy + kD + b; k = q.oz + d; break;
    default: i = q.ox + kD + a; j = q.oy + kD + b; k = q.oz + q.ez - 1 - d; break;
    }
}

__device__ __forceinline__ int face_of(const BandGeo &q, int c)
{
    int f = 0;
#pragma unroll
    for (int t = 1; t < 6; ++t) f += (c >= q.col0[t]);
    return f;
}

__device__ __forceinline__ int lateral_a_extent(const BandGeo &q, int f)
{
    return f < 2 ? q.ey : q.ex - 2 * kD;
}

struct Fields9 {
    float *f[kF];
};

Fields9 fields9(const sw_ssg3d &g)
{
    return Fields9{{g.vx, g.vy, g.vz, g.sxx, g.syy, g.szz, g.sxy, g.sxz, g.syz}};
}

// Current band values of every column: out[c][d*9 + field].
__global__ void k_band_gather(Fields9 F, int nx, int ny, BandGeo q, float *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= q.col0[6]) return;
    const int f = face_of(q, c);
    const int r = c - q.col0[f];
    const int na = lateral_a_extent(q, f);
    const int a = r % na, b = r / na;
    for (int d = 0; d < kD; ++d) {
        int i, j, k;
        band_cell(q, f, a, b, d, i, j, k);
        const long p = ((long)k * ny + j) * nx + i;
#pragma unroll
        for (int t = 0; t < kF; ++t) out[(long)c * kOut + d * kF + t] = F.f[t][p];
    }
}

// Fused inference: gather, MLP, correct the fields, keep the corrected band.
//
// Each wave works alone on tiles of 16 columns of one face, and a tile never
// leaves the wave's registers.  Lane (li, lg) = lane li + 16 lg holds band
// depth lg of column li of the tile.  The layers run on the matrix cores as
// f32 16x16x4 products D[m][n] = sum_k A[m][k] B[k][n], with the weights as A
// (m = output neuron within a block of 16) and the activations as B (n = the
// tile's column): lane (li, lg) supplies A[li][lg] and B[lg][li] and receives
// D[4 lg + r][li], r = 0..3.
//  * Layer 1 takes its 72 inputs in the order (m, depth): k-step m (0..17)
//    feeds input m of each depth (field m at the depth for m < 9, the change
//    of field m - 9 for m >= 9), which is exactly what lane (li, lg) gathered
//    for its own depth lg; W1's rows are permuted to match when the face's
//    weights are staged in LDS.
//  * A layer's output leaves lane (li, lg) with neurons 16 nb + 4 lg + r of
//    column li; a 4x4 transpose over the lane groups lg and r (two
//    permlane32 and two permlane16 swaps per block of 4 registers) gives it
//    neurons 16 nb + 4 r + lg, the B operands of the next layer's k-steps
//    16 nb + 4 r.
//  * Layer 3's 36 outputs are laid out so that lane (li, lg) receives the 9
//    corrections of its own depth: slot 16 nb + 4 g + r is output
//    9 g + 4 nb + r (zero weights where 4 nb + r >= 9).
// LDS holds only the face's permuted weights (52 KB), so no barrier follows
// the one after staging them; while one wave issues matrix products the
// other waves of its SIMD gather, evaluate tanh or scatter.  A wave gathers
// its next tile while it computes the current one, and its first tile while
// the block stages the weights (110 VGPRs).  One block of 12 waves per CU
// (3 per SIMD; about three tiles per wave: the 8,553 tiles of the 200^3 case
// on 238 blocks) keeps the gathers behind the matrix products.  Measured on
// MI355X at 200^3: 53.8 us per launch, against 59.3 us for 10-wave blocks at
// 5 waves per SIMD without the prefetch (at most two tiles per wave) and
// 82 us for 8- or 16-wave blocks with it.  Each block serves one face: the
// host gives every face a share of the blocks in proportion to its tiles.
constexpr int kWaveCols = 16;
constexpr int kCorrectWaves = 12;
constexpr int kCorrectThreads = 64 * kCorrectWaves;
typedef float mf4 __attribute__((ext_vector_type(4)));

// tanh from the hardware exponential and reciprocal: 1 - 2 / (e^{2x} + 1);
// absolute error ~1e-7 (relative error grows only where tanh itself is tiny).
// (A division here compiles to the full IEEE sequence, ten instructions.)
__device__ __forceinline__ float fast_tanh(float x)
{
    return 1.f - 2.f * __builtin_amdgcn_rcpf(__expf(2.f * x) + 1.f);
}

// Staged weights of one face (floats).  W1P[m][g][li][nb] = W1[in1(m, g)][16 nb + li];
// W2P[k][li][nb] = W2[k][16 nb + li]; W3P[k][li][nb] = W3[k][out3(nb, li)] (zero
// for nb = 3 and for the padding slots); B3P, OSP[16 nb + 4 g + r] = b3,
// out_scale of output 9 g + 4 nb + r (zero for padding); ISI = 1 / in_scale.
constexpr int kSW1 = 0, kSB1 = kSW1 + kIn * kH, kSW2 = kSB1 + kH, kSB2 = kSW2 + kH * kH, kSW3 = kSB2 + kH;
constexpr int kSB3 = kSW3 + kH * 64, kSOS = kSB3 + 48, kSIS = kSOS + 48, kStaged = kSIS + kIn;
// parameter offsets (params layout above)
constexpr int kPB1 = kIn * kH, kPW2 = kPB1 + kH, kPB2 = kPW2 + kH * kH, kPW3 = kPB2 + kH, kPB3 = kPW3 + kH * kOut;
constexpr int kPIS = kPB3 + kOut, kPOS = kPIS + kIn;
static_assert(kPOS + kOut == kParams, "parameter layout");

__device__ __forceinline__ int in1_of(int m, int g) { return m < kF ? kF * g + m : kOut + kF * g + (m - kF); }

// Blocks of each face: blocks [blk0[f], blk0[f+1]) take the ntile[f] tiles of
// face f in contiguous, near-equal ranges; the waves of a block take its
// tiles in turn.
struct BlockMap {
    int blk0[7];
    int ntile[6];
};

// One lane's band cell of a tile: column li of the tile, depth lg.
struct BandIO {
    int col;
    bool live;
    long pc;
    float cur[kF], pv[kF];
};

__device__ __forceinline__ void fetch_band(BandIO &b, const Fields9 &F, int nx, int ny, const BandGeo &q, int f,
                                           const float *prev, int tile, int li, int lg)
{
    const int r0 = tile * kWaveCols + li;  // column within the face
    b.live = r0 < q.ncol[f];
    const int r = b.live ? r0 : q.ncol[f] - 1;  // a dead lane reads the face's last column
    b.col = q.col0[f] + r;
    const int na = lateral_a_extent(q, f);
    int i, j, k;
    band_cell(q, f, r % na, r / na, lg, i, j, k);
    b.pc = ((long)k * ny + j) * nx + i;
    float cv[kF], pv[kF];
#pragma unroll
    for (int u = 0; u < kF; ++u) {
        cv[u] = F.f[u][b.pc];
        pv[u] = prev[(long)b.col * kOut + lg * kF + u];
    }
#pragma unroll
    for (int u = 0; u < kF; ++u) {
        b.cur[u] = b.live ? cv[u] : 0.f;
        b.pv[u] = b.live ? pv[u] : 0.f;
    }
}

// Rows of 4 registers x 4 lane groups transposed (see above).
__device__ __forceinline__ void transpose4(float *h)
{
    const auto s02 = __builtin_amdgcn_permlane32_swap(__float_as_uint(h[0]), __float_as_uint(h[2]), false, false);
    const auto s13 = __builtin_amdgcn_permlane32_swap(__float_as_uint(h[1]), __float_as_uint(h[3]), false, false);
    const auto t01 = __builtin_amdgcn_permlane16_swap(s02[0], s13[0], false, false);
    const auto t23 = __builtin_amdgcn_permlane16_swap(s02[1], s13[1], false, false);
    h[0] = __uint_as_float(t01[0]);
    h[1] = __uint_as_float(t01[1]);
    h[2] = __uint_as_float(t23[0]);
    h[3] = __uint_as_float(t23[1]);
}

__global__ void __launch_bounds__(kCorrectThreads) __attribute__((amdgpu_waves_per_eu(3))) k_band_correct(Fields9 F, int nx, int ny, BandGeo q, BlockMap bm,
                                                                  const float *params, float *prev)
{
    __shared__ __attribute__((aligned(16))) float w[kStaged];
    const int b = (int)blockIdx.x;
    int f = 0;
#pragma unroll
    for (int u = 1; u < 6; ++u) f += (b >= bm.blk0[u]);
    const int nb_f = bm.blk0[f + 1] - bm.blk0[f], lb = b - bm.blk0[f];
    const int t0 = (int)(((long)lb * bm.ntile[f]) / nb_f);
    const int t1 = (int)(((long)(lb + 1) * bm.ntile[f]) / nb_f);
    const int wv = (int)threadIdx.x / 64, ln = (int)threadIdx.x % 64;
    const int li = ln % 16, lg = ln / 16;
    BandIO io;
    int tile = t0 + wv;
    if (tile < t1) fetch_band(io, F, nx, ny, q, f, prev, tile, li, lg);
    {
        // All of a thread's loads are issued before its first LDS store, so
        // the ~26 L2 reads per thread overlap instead of running one by one.
        const float *P = params + (long)f * kParams;
        constexpr int kPer = (kStaged + kCorrectThreads - 1) / kCorrectThreads;
        float x[kPer];
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int v = (int)threadIdx.x + t * kCorrectThreads;
            int src = -1;  // parameter index, or -1 for a zero (padding)
            bool inv = false;
            if (v < kSB1) {  // W1P
                const int nb = v & 3, li = (v >> 2) & 15, g = (v >> 6) & 3, m = v >> 8;
                src = in1_of(m, g) * kH + 16 * nb + li;
            } else if (v < kSW2) {
                src = kPB1 + (v - kSB1);
            } else if (v < kSB2) {  // W2P
                const int e = v - kSW2, nb = e & 3, li = (e >> 2) & 15, k = e >> 6;
                src = kPW2 + k * kH + 16 * nb + li;
            } else if (v < kSW3) {
                src = kPB2 + (v - kSB2);
            } else if (v < kSB3) {  // W3P
                const int e = v - kSW3, nb = e & 3, li = (e >> 2) & 15, k = e >> 6;
                const int u = 4 * nb + (li & 3);
                if (nb < 3 && u < kF) src = kPW3 + k * kOut + kF * (li >> 2) + u;
            } else if (v < kSIS) {  // B3P, OSP
                const int e = (v - kSB3) % 48, nb = e >> 4, g = (e >> 2) & 3, r = e & 3, u = 4 * nb + r;
                if (u < kF) src = (v < kSOS ? kPB3 : kPOS) + kF * g + u;
            } else if (v < kStaged) {  // 1 / in_scale: the inputs are scaled with one multiply each
                src = kPIS + (v - kSIS);
                inv = true;
            }
            const float y = src >= 0 ? P[src] : 0.f;
            x[t] = inv ? 1.f / y : y;
        }
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int v = (int)threadIdx.x + t * kCorrectThreads;
            if (v < kStaged) w[v] = x[t];
        }
    }
    __syncthreads();
    const float *W1P = w + kSW1, *B1 = w + kSB1, *W2P = w + kSW2, *B2 = w + kSB2, *W3P = w + kSW3;
    const float *B3P = w + kSB3, *OSP = w + kSOS, *ISI = w + kSIS;
    for (; tile < t1; tile += kCorrectWaves) {
        BandIO nio;
        const bool more = tile + kCorrectWaves < t1;
        if (more) fetch_band(nio, F, nx, ny, q, f, prev, tile + kCorrectWaves, li, lg);
        // the staged weights are read where they are used: without this the
        // compiler hoists all of them out of the tile loop into registers
        asm volatile("" ::: "memory");
        // 1. inputs of depth lg of column li and the column's scale s, from
        // the four depths' partial sums added in depth order
        float xin[2 * kF];
        float ss = 0.f;
#pragma unroll
        for (int u = 0; u < kF; ++u) {
            xin[u] = io.cur[u] * ISI[kF * lg + u];
            xin[kF + u] = (io.cur[u] - io.pv[u]) * ISI[kOut + kF * lg + u];
            ss += xin[u] * xin[u] + xin[kF + u] * xin[kF + u];
        }
        const float sc = sqrtf((__shfl(ss, li) + __shfl(ss, li + 16) + __shfl(ss, li + 32) + __shfl(ss, li + 48)) /
                               (float)kIn);
        const float ainv = sc > 0.f ? 1.f / sc : 0.f;
        // 2. layer 1: h1 = tanh(W1^T (u / s) + b1)
        float h[4][4];
        {
            mf4 acc[4];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) acc[nb] = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < 2 * kF; ++m) {
                const float bv = xin[m] * ainv;
                const mf4 av = *reinterpret_cast<const mf4 *>(W1P + ((m * 4 + lg) * 16 + li) * 4);
#pragma unroll
                for (int nb = 0; nb < 4; ++nb)
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[nb], bv, acc[nb], 0, 0, 0);
            }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const mf4 bias = *reinterpret_cast<const mf4 *>(B1 + 16 * nb + 4 * lg);
#pragma unroll
                for (int r = 0; r < 4; ++r) h[nb][r] = fast_tanh(acc[nb][r] + bias[r]);
                transpose4(h[nb]);
            }
        }
        // 3. layer 2: h2 = tanh(W2^T h1 + b2)
        {
            mf4 acc[4];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) acc[nb] = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb)
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    const int k0 = 16 * kb + 4 * s4;
                    const mf4 av = *reinterpret_cast<const mf4 *>(W2P + ((k0 + lg) * 16 + li) * 4);
#pragma unroll
                    for (int nb = 0; nb < 4; ++nb)
                        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[nb], h[kb][s4], acc[nb], 0, 0, 0);
                }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const mf4 bias = *reinterpret_cast<const mf4 *>(B2 + 16 * nb + 4 * lg);
#pragma unroll
                for (int r = 0; r < 4; ++r) h[nb][r] = fast_tanh(acc[nb][r] + bias[r]);
                transpose4(h[nb]);
            }
        }
        // 4. layer 3: y = (W3^T h2 + b3) * out_scale * s, the 9 outputs of
        // depth lg in slots (nb, r) with 4 nb + r < 9
        mf4 acc3[3];
#pragma unroll
        for (int nb = 0; nb < 3; ++nb) acc3[nb] = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int k0 = 16 * kb + 4 * s4;
                const mf4 av = *reinterpret_cast<const mf4 *>(W3P + ((k0 + lg) * 16 + li) * 4);
#pragma unroll
                for (int nb = 0; nb < 3; ++nb)
                    acc3[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[nb], h[kb][s4], acc3[nb], 0, 0, 0);
            }
        // 5. add the correction to depth lg of column li, keep the band
        if (io.live) {
#pragma unroll
            for (int u = 0; u < kF; ++u) {
                const int nb = u / 4, r = u % 4, slot = 16 * nb + 4 * lg + r;
                const float y = sc > 0.f ? (acc3[nb][r] + B3P[slot]) * OSP[slot] * sc : 0.f;
                const float v = io.cur[u] + y;
                F.f[u][io.pc] = v;
                prev[(long)io.col * kOut + lg * kF + u] = v;
            }
        }
        if (more) io = nio;
    }
}

// ------------------------------------------------------------ training
// Each thread sums one output in row order (the oracle's order); the loops
// are unrolled by 8 so that 8 rows' loads are in flight while the sums keep
// that order (measured on MI355X: a 6-face training 9.9 -> 2.2 s, parameters
// bit-identical).
// Dense layer Z = A W + b (A [B][nin], W [nin][nout]); act 1 applies tanh.
// The first layer divides its input by in_scale (when given).
__global__ void k_dense(const float *A, const float *W, const float *bias, const float *in_scale, float *Z, int B,
                        int nin, int nout, int act)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * nout) return;
    const int o = (int)(t % nout);
    const long row = t / nout;
    float acc = bias[o];
#pragma unroll 8
    for (int i = 0; i < nin; ++i) {
        const float a = in_scale ? A[row * nin + i] / in_scale[i] : A[row * nin + i];
        acc += a * W[(long)i * nout + o];
    }
    Z[t] = act ? tanhf(acc) : acc;
}

// dZ of a tanh layer from the gradient of its output: dZ = (dH W_next^T) * (1 - H^2),
// with dH given as the next layer's dZ [B][nnext] and W_next [n][nnext].
__global__ void k_dense_back(const float *dnext, const float *Wnext, const float *H, float *dZ, int B, int n,
                             int nnext)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * n) return;
    const int i = (int)(t % n);
    const long row = t / n;
    float acc = 0.f;
#pragma unroll 8
    for (int o = 0; o < nnext; ++o) acc += dnext[row * nnext + o] * Wnext[(long)i * nnext + o];
    const float h = H[t];
    dZ[t] = acc * (1.f - h * h);
}

// dW[i][o] = sum_b A[b][i] dZ[b][o] (A scaled by 1/in_scale when given); db[o] = sum_b dZ[b][o].
__global__ void k_dense_wgrad(const float *A, const float *in_scale, const float *dZ, float *dW, float *db, int B,
                              int nin, int nout)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nin * nout) {
        const int i = t / nout, o = t % nout;
        const float sc = in_scale ? 1.f / in_scale[i] : 1.f;
        float acc = 0.f;
#pragma unroll 8
        for (int b = 0; b < B; ++b) acc += A[(long)b * nin + i] * dZ[(long)b * nout + o];
        dW[t] = acc * sc;
    } else if (t < nin * nout + nout) {
        const int o = t - nin * nout;
        float acc = 0.f;
#pragma unroll 8
        for (int b = 0; b < B; ++b) acc += dZ[(long)b * nout + o];
        db[o] = acc;
    }
}

// Mean squared error in normalised output units: L = mean((Y - T/out_scale)^2);
// dY = 2 (Y - T/out_scale) / (B*nout); loss_parts[b] = row sum of squares.
__global__ void k_mse(const float *Y, const float *T, const float *out_scale, float *dY, float *loss_rows, int B,
                      int nout)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float acc = 0.f;
    const float inv = 2.f / ((float)B * (float)nout);
    for (int o = 0; o < nout; ++o) {
        const float e = Y[(long)b * nout + o] - T[(long)b * nout + o] / out_scale[o];
        dY[(long)b * nout + o] = e * inv;
        acc += e * e;
    }
    loss_rows[b] = acc;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int learned_band_columns(const sw_band *b)
{
    return band_geo(*b).col0[6];
}

int launch_band_gather(const sw_ssg3d *g, const sw_band *b, float *out, hipStream_t st)
{
    const BandGeo q = band_geo(*b);
    hipLaunchKernelGGL(k_band_gather, dim3(nblk(q.col0[6], 256)), dim3(256), 0, st, fields9(*g), g->nx, g->ny, q,
                       out);
    return launch_status("band gather");
}

int launch_band_correct(const sw_ssg3d *g, const sw_band *b, const float *params, float *prev, hipStream_t st)
{
    const BandGeo q = band_geo(*b);
    BlockMap bm;
    long total = 0;
    for (int f = 0; f < 6; ++f) {
        bm.ntile[f] = (q.ncol[f] + kWaveCols - 1) / kWaveCols;
        total += bm.ntile[f];
    }
    if (total == 0) return SW_OK;
    // about three tiles per wave, at most one block per CU (one round of
    // resident blocks: at most 256 on MI355X)
    const long want = std::max(1L, std::min((long)compute_units(), (total + 3 * kCorrectWaves - 1) / (3 * kCorrectWaves)));
    bm.blk0[0] = 0;
    for (int f = 0; f < 6; ++f) {
        long nbf = bm.ntile[f] ? (want * bm.ntile[f] + total - 1) / total : 0;
        nbf = std::min<long>(std::max<long>(nbf, bm.ntile[f] ? 1 : 0), bm.ntile[f]);
        bm.blk0[f + 1] = bm.blk0[f] + (int)nbf;
    }
    hipLaunchKernelGGL(k_band_correct, dim3(bm.blk0[6]), dim3(kCorrectThreads), 0, st, fields9(*g), g->nx, g->ny, q,
                       bm, params, prev);
    return launch_status("band correction");
}

int launch_mlp_forward(const float *params, const float *X, int B, float *H1, float *H2, float *Y, hipStream_t st)
{
    const float *W1 = params, *B1 = W1 + kIn * kH, *W2 = B1 + kH, *B2 = W2 + kH * kH, *W3 = B2 + kH;
    const float *B3 = W3 + kH * kOut, *IS = B3 + kOut;
    hipLaunchKernelGGL(k_dense, dim3(nblk((long)B * kH, 256)), dim3(256), 0, st, X, W1, B1, IS, H1, B, kIn, kH, 1);
    hipLaunchKernelGGL(k_dense, dim3(nblk((long)B * kH, 256)), dim3(256), 0, st, H1, W2, B2, (const float *)nullptr,
                       H2, B, kH, kH, 1);
    hipLaunchKernelGGL(k_dense, dim3(nblk((long)B * kOut, 256)), dim3(256), 0, st, H2, W3, B3,
                       (const float *)nullptr, Y, B, kH, kOut, 0);
    return launch_status("mlp forward");
}

int launch_mlp_loss_backward(const float *params, const float *X, const float *H1, const float *H2, const float *Y,
                             const float *T, int B, float *grad, float *work, float *loss_rows, hipStream_t st)
{
    const float *W1 = params, *B1 = W1 + kIn * kH, *W2 = B1 + kH, *B2 = W2 + kH * kH, *W3 = B2 + kH;
    const float *B3 = W3 + kH * kOut, *IS = B3 + kOut, *OS = IS + kIn;
    (void)B1; (void)B2;
    float *gW1 = grad, *gB1 = gW1 + kIn * kH, *gW2 = gB1 + kH, *gB2 = gW2 + kH * kH, *gW3 = gB2 + kH;
    float *gB3 = gW3 + kH * kOut;
    float *dY = work, *dZ2 = dY + (long)B * kOut, *dZ1 = dZ2 + (long)B * kH;
    hipLaunchKernelGGL(k_mse, dim3(nblk(B, 256)), dim3(256), 0, st, Y, T, OS, dY, loss_rows, B, kOut);
    hipLaunchKernelGGL(k_dense_wgrad, dim3(nblk(kH * kOut + kOut, 256)), dim3(256), 0, st, H2,
                       (const float *)nullptr, dY, gW3, gB3, B, kH, kOut);
    hipLaunchKernelGGL(k_dense_back, dim3(nblk((long)B * kH, 256)), dim3(256), 0, st, dY, W3, H2, dZ2, B, kH, kOut);
    hipLaunchKernelGGL(k_dense_wgrad, dim3(nblk(kH * kH + kH, 256)), dim3(256), 0, st, H1, (const float *)nullptr,
                       dZ2, gW2, gB2, B, kH, kH);
    hipLaunchKernelGGL(k_dense_back, dim3(nblk((long)B * kH, 256)), dim3(256), 0, st, dZ2, W2, H1, dZ1, B, kH, kH);
    hipLaunchKernelGGL(k_dense_wgrad, dim3(nblk(kIn * kH + kH, 256)), dim3(256), 0, st, X, IS, dZ1, gW1, gB1, B, kIn,
                       kH);
    return launch_status("mlp backward");
}

int launch_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, int t,
                hipStream_t st)
{
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return launch_status("adam");
}

int mlp_param_count() { return kParams; }

// ------------------------------------------------ allocation-free training step
// Scratch of sw_mlp_loss_backward_noalloc: the dY, dZ2 and dZ1 rows
// (kOut + 2 kH floats per sample) of up to kScratchRows samples, 10.7 MB held
// on every device the library runs on until the process ends.  In host code
// `scratch` names a host-side stand-in; kernels get its device address from
// mlp_scratch_lookup, one lookup per device.
constexpr int kScratchRows = SW_MLP_NOALLOC_MAX_ROWS;
constexpr long kScratchFloats = (long)kScratchRows * (kOut + 2 * kH);
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device (looked up once per
// device, with that device current).
int mlp_scratch_lookup(float **out)
{
    int dev = 0;
    SW_TRY_HIP(hipGetDevice(&dev));
    SW_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the scratch table holds", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        SW_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<float *>(p);
    }
    *out = g_scratch[dev];
    return SW_OK;
}

int launch_mlp_loss_backward_noalloc(const float *params, const float *X, const float *H1, const float *H2,
                                     const float *Y, const float *T, int B, float *grad, float *loss_rows,
                                     hipStream_t st)
{
    SW_REQUIRE(B <= kScratchRows, "batch of %d samples above the %d the allocation-free scratch holds", B,
               kScratchRows);
    float *work = nullptr;
    int rc = mlp_scratch_lookup(&work);
    if (rc) return rc;
    // every scratch row is written (k_mse, k_dense_back) before it is read,
    // so the previous call's contents need no clearing
    return launch_mlp_loss_backward(params, X, H1, H2, Y, T, B, grad, work, loss_rows, st);
}

}  // namespace sw
