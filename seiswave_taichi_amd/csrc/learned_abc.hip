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
// A block of 256 threads works on tiles of 64 columns of one face at a time:
// the features go to LDS, and each layer is a 64 x n_out product on the
// matrix cores (f32 16x16x4 MFMA, full f32 arithmetic), each wave owning 16
// columns; the weights of the face sit in LDS and activations are re-staged
// in LDS between layers.  Blocks take contiguous ranges of face-major tiles
// and reload weights only when the face changes.
constexpr int kCorrectThreads = 256;
constexpr int kTile = 64;
typedef float mf4 __attribute__((ext_vector_type(4)));
constexpr int kW1End = kIn * kH;                  // end of W1 in a face's parameters
constexpr int kW2Begin = kIn * kH + kH;           // W2 follows b1
constexpr int kW2End = kW2Begin + kH * kH;
static_assert(kParams % 4 == 0 && kW1End % 4 == 0 && kW2Begin % 4 == 0, "16-byte weight loads");

// tanh from the hardware exponential: 1 - 2 / (e^{2x} + 1); absolute error
// ~1e-7 (relative error grows only where tanh itself is tiny).
__device__ __forceinline__ float fast_tanh(float x)
{
    return 1.f - __fdividef(2.f, __expf(2.f * x) + 1.f);
}
constexpr int kXS = kIn + 1;  // padded LDS row strides (odd: conflict-free column access)
constexpr int kHS = kH + 1;

struct TileMap {
    int tile0[7];  // first tile of each face (face-major order)
};

// One thread's share of a tile's band: depth gd of column gc (see below).
struct BandIO {
    int f, gc, gd, col;
    bool live;
    long pc;
    float cur[kF], pv[kF];
};

// Global accesses (gather and scatter) use thread (gc, gd): on x faces the 4
// band cells of a column are consecutive in x, so 4 neighbouring lanes take
// the 4 depths of one column (16-byte runs); on y and z faces consecutive
// lanes take consecutive columns, which are consecutive in x.
__device__ __forceinline__ void fetch_band(BandIO &b, const Fields9 &F, int nx, int ny, const BandGeo &q,
                                           const TileMap &tm, const float *prev, int tile)
{
    int f = 0;
#pragma unroll
    for (int u = 1; u < 6; ++u) f += (tile >= tm.tile0[u]);
    const bool xface = f < 2;
    b.f = f;
    b.gc = xface ? (int)threadIdx.x >> 2 : (int)threadIdx.x % kTile;
    b.gd = xface ? (int)threadIdx.x & 3 : (int)threadIdx.x / kTile;
    const int r = (tile - tm.tile0[f]) * kTile + b.gc;  // column within the face
    b.live = r < q.ncol[f];
    b.col = q.col0[f] + r;
    b.pc = 0;
    if (b.live) {
        const int na = lateral_a_extent(q, f);
        int i, j, k;
        band_cell(q, f, r % na, r / na, b.gd, i, j, k);
        b.pc = ((long)k * ny + j) * nx + i;
    }
#pragma unroll
    for (int u = 0; u < kF; ++u) {
        b.cur[u] = b.live ? F.f[u][b.pc] : 0.f;
        b.pv[u] = b.live ? prev[(long)b.col * kOut + b.gd * kF + u] : 0.f;
    }
}

__global__ void __launch_bounds__(kCorrectThreads) k_band_correct(Fields9 F, int nx, int ny, BandGeo q, TileMap tm,
                                                                  const float *params, float *prev)
{
    __shared__ __attribute__((aligned(16))) float w[kParams];
    __shared__ float xs[kTile * kXS];  // features, then h2
    __shared__ float hs[kTile * kHS];  // h1, then the correction
    __shared__ float red[4][kTile];
    __shared__ float scale[kTile];
    const int ntile = tm.tile0[6];
    const int t0 = (int)(((long)blockIdx.x * ntile) / gridDim.x);
    const int t1 = (int)(((long)(blockIdx.x + 1) * ntile) / gridDim.x);
    const int c = threadIdx.x % kTile, grp = threadIdx.x / kTile;
    int loaded = -1;
    BandIO io, nx_io;
    if (t0 < t1) fetch_band(io, F, nx, ny, q, tm, prev, t0);
    for (int tile = t0; tile < t1; ++tile) {
        const int f = io.f;
        __syncthreads();  // the previous tile is done with w, xs and hs
        if (f != loaded) {
            const float *P = params + (long)f * kParams;
            // W1 and W2 (64 outputs each) are stored permuted, [row][li][nb] =
            // W[row][16 nb + li], so one 16-byte LDS read gives a lane its
            // B operands of all 4 output blocks; the rest as in params.
            // 16-byte global loads, 4 in flight per thread.
            const mf4 *P4 = reinterpret_cast<const mf4 *>(P);
#pragma unroll 4
            for (int v = threadIdx.x; v < kParams / 4; v += blockDim.x) {
                const mf4 x = P4[v];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int u = 4 * v + e;
                    int dst = u;
                    if (u < kW1End || (u >= kW2Begin && u < kW2End)) {
                        const int base = u < kW1End ? 0 : kW2Begin;
                        const int row = (u - base) / kH, col = (u - base) % kH;
                        dst = base + row * kH + 4 * (col % 16) + col / 16;
                    }
                    w[dst] = x[e];
                }
            }
            loaded = f;
            __syncthreads();
        }
        const float *W1 = w, *B1 = W1 + kIn * kH, *W2 = B1 + kH, *B2 = W2 + kH * kH, *W3 = B2 + kH;
        const float *B3 = W3 + kH * kOut, *IS = B3 + kOut, *OS = IS + kIn;
        // 1. features of band depth gd of column gc (fetched ahead)
        {
            float ss = 0.f;
#pragma unroll
            for (int u = 0; u < kF; ++u) {
                const int o = io.gd * kF + u;
                const float x0 = io.cur[u] / IS[o];
                const float x1 = (io.cur[u] - io.pv[u]) / IS[kOut + o];
                xs[io.gc * kXS + o] = x0;
                xs[io.gc * kXS + kOut + o] = x1;
                ss += x0 * x0 + x1 * x1;
            }
            red[io.gd][io.gc] = ss;
        }
        __syncthreads();
        if (grp == 0) scale[c] = sqrtf((red[0][c] + red[1][c] + red[2][c] + red[3][c]) / (float)kIn);
        __syncthreads();
        // the next tile's band loads overlap this tile's layers
        if (tile + 1 < t1) fetch_band(nx_io, F, nx, ny, q, tm, prev, tile + 1);
        // 2-4. The three layers as 16x16x4 f32 matrix-core products: wave wv
        // owns tile rows (columns) 16 wv .. 16 wv + 15; lane (li, lk) feeds
        // A[li][lk] = act[16 wv + li][k0 + lk] and B[lk][li] = W[k0 + lk][n0 + li],
        // and receives D[4 lk + r][li], r = 0..3.  Rows are private to a wave.
        const int wv = (int)threadIdx.x / 64, ln = (int)threadIdx.x % 64;
        const int li = ln % 16, lk = ln / 16;
        const int arow = 16 * wv + li;  // A row of this lane
        const float ainv = scale[arow] > 0.f ? 1.f / scale[arow] : 0.f;
        // layer 1: h1 = tanh(W1^T (u / s) + b1), 64 outputs in 4 blocks of 16
        {
            mf4 acc[4];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) acc[nb] = mf4{0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < kIn; k0 += 4) {
                const float av = xs[arow * kXS + k0 + lk] * ainv;
                const mf4 bv = *reinterpret_cast<const mf4 *>(W1 + (k0 + lk) * kH + 4 * li);
#pragma unroll
                for (int nb = 0; nb < 4; ++nb)
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[nb], acc[nb], 0, 0, 0);
            }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    hs[(16 * wv + 4 * lk + r) * kHS + 16 * nb + li] = fast_tanh(acc[nb][r] + B1[16 * nb + li]);
        }
        // layers 2 and 3 read only this wave's rows: no block barrier
        __builtin_amdgcn_wave_barrier();
        // layer 2: h2 = tanh(W2^T h1 + b2), staged in xs
        {
            mf4 acc[4];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) acc[nb] = mf4{0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < kH; k0 += 4) {
                const float av = hs[arow * kHS + k0 + lk];
                const mf4 bv = *reinterpret_cast<const mf4 *>(W2 + (k0 + lk) * kH + 4 * li);
#pragma unroll
                for (int nb = 0; nb < 4; ++nb)
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[nb], acc[nb], 0, 0, 0);
            }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    xs[(16 * wv + 4 * lk + r) * kXS + 16 * nb + li] = fast_tanh(acc[nb][r] + B2[16 * nb + li]);
        }
        __builtin_amdgcn_wave_barrier();
        // layer 3: y = (W3^T h2 + b3) * out_scale * s, 36 outputs in 3 blocks
        // of 16 (the last 12 columns of B are zero), staged in hs
        {
            mf4 acc[3];
#pragma unroll
            for (int nb = 0; nb < 3; ++nb) acc[nb] = mf4{0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < kH; k0 += 4) {
                const float av = xs[arow * kXS + k0 + lk];
#pragma unroll
                for (int nb = 0; nb < 3; ++nb) {
                    const int o = 16 * nb + li;
                    const float bv = o < kOut ? W3[(k0 + lk) * kOut + o] : 0.f;
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[nb], 0, 0, 0);
                }
            }
#pragma unroll
            for (int nb = 0; nb < 3; ++nb) {
                const int o = 16 * nb + li;
                if (o < kOut)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * wv + 4 * lk + r;
                        const float sr = scale[row];
                        hs[row * kHS + o] = sr > 0.f ? (acc[nb][r] + B3[o]) * OS[o] * sr : 0.f;
                    }
            }
        }
        __syncthreads();
        // 5. add the correction to band depth gd of column gc, keep the band
        if (io.live) {
#pragma unroll
            for (int u = 0; u < kF; ++u) {
                const int o = io.gd * kF + u;
                const float v = io.cur[u] + hs[io.gc * kHS + o];
                F.f[u][io.pc] = v;
                prev[(long)io.col * kOut + o] = v;
            }
        }
        io = nx_io;
    }
}

// ------------------------------------------------------------ training
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
        for (int b = 0; b < B; ++b) acc += A[(long)b * nin + i] * dZ[(long)b * nout + o];
        dW[t] = acc * sc;
    } else if (t < nin * nout + nout) {
        const int o = t - nin * nout;
        float acc = 0.f;
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
    TileMap tm;
    tm.tile0[0] = 0;
    for (int f = 0; f < 6; ++f) tm.tile0[f + 1] = tm.tile0[f] + (q.ncol[f] + kTile - 1) / kTile;
    if (tm.tile0[6] == 0) return SW_OK;
    // two resident blocks per CU (80 KB of LDS each), a few tiles per block
    const int nblock = std::min(tm.tile0[6], 512);
    hipLaunchKernelGGL(k_band_correct, dim3(nblock), dim3(kCorrectThreads), 0, st, fields9(*g), g->nx, g->ny, q, tm,
                       params, prev);
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
