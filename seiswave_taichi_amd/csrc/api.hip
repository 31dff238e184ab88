// C ABI of libseiswave.so: argument checks, source injection, receiver
// gather, multi-step drivers and the free-surface helpers.
#include "sw_internal.h"

#include <atomic>
#include <climits>
#include <cstring>
#include <mutex>

namespace sw {

namespace {
thread_local char g_err[512] = "";
std::atomic<int> g_plan_mode{SW_PLAN_AUTO};
}

int compute_units()
{
    static std::mutex mu;
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();  // no device here: clear the sticky error
        return 256;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (cus[dev] == 0) {
        int v = 0;
        cus[dev] = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0 ? v : 256;
    }
    return cus[dev];
}

int launch_plan_mode() { return g_plan_mode.load(std::memory_order_relaxed); }

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int launch_ssg3d_velocity(const sw_ssg3d *g, hipStream_t st);
int launch_ssg3d_stress(const sw_ssg3d *g, hipStream_t st);
int launch_ssg3d_velocity_spml(const sw_ssg3d *g, hipStream_t st);
int launch_ssg3d_stress_spml(const sw_ssg3d *g, hipStream_t st);
int ssg3d_plan_masks(const sw_ssg3d *g, int32_t *pm, int32_t max);
long split_shell_map(const sw_ssg3d *g, int sweep, int64_t *cell, int64_t *slot, long max);
int launch_ssg3d_velocity_fold(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st);
int launch_ssg3d_stress_fold(const sw_ssg3d *g, const Fold3D &fo, hipStream_t st);
int launch_ssg2d_velocity(const sw_ssg2d *g, hipStream_t st);
int launch_ssg2d_stress(const sw_ssg2d *g, hipStream_t st);
int launch_adj2d_stress(const sw_ssg2d *f, const sw_adj2d *q, hipStream_t st);
int learned_band_columns(const sw_band *b);
int launch_band_gather(const sw_ssg3d *g, const sw_band *b, float *out, hipStream_t st);
int launch_band_correct(const sw_ssg3d *g, const sw_band *b, const float *params, float *prev, hipStream_t st);
int launch_mlp_forward(const float *params, const float *X, int B, float *H1, float *H2, float *Y, hipStream_t st);
int launch_mlp_loss_backward(const float *params, const float *X, const float *H1, const float *H2, const float *Y,
                             const float *T, int B, float *grad, float *work, float *loss_rows, hipStream_t st);
int launch_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, int t,
                hipStream_t st);
int mlp_param_count();
int launch_stream_probe(const float *const *in, int nin, float *const *out, int nout, long n, int nblock,
                        hipStream_t st);
int launch_mfma_probe(float *out, int nblock, int steps, hipStream_t st);
int mlp_scratch_lookup(float **out);
int launch_mlp_loss_backward_noalloc(const float *params, const float *X, const float *H1, const float *H2,
                                     const float *Y, const float *T, int B, float *grad, float *loss_rows,
                                     hipStream_t st);
int launch_adj2d_velocity(const sw_ssg2d *f, const sw_adj2d *q, hipStream_t st);
int launch_ssg2d_velocity_batch(const sw_ssg2d *g, int nbatch, long bs, const Fold2D &fo, hipStream_t st);
int launch_ssg2d_stress_batch(const sw_ssg2d *g, int nbatch, long bs, const Fold2D &fo, hipStream_t st);
int launch_adj2d_stress_batch(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long fs, long as, const float *sv,
                              long ss, hipStream_t st);
int launch_adj2d_velocity_batch(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long fs, long as,
                                const float *sv, long ss, hipStream_t st);
int launch_ssg2d_sweeps_saved(const sw_ssg2d *g, int nbatch, long bs, bool stress, float *sv, long ss,
                              const Fold2D &fo, hipStream_t st);
int launch_adj2d_fused(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long as, int pass, const float *sv, long ss,
                       float *tv, long ts, hipStream_t st);
int launch_adj2d_inject_batch(const sw_adj2d *q, const sw_receivers *r, int it, int nbatch, long as, long ds,
                              hipStream_t st);
int launch_adj2d_model_gradient_batch(const sw_ssg2d *f, const sw_adj2d *q, int nbatch, long as, const float *vp,
                                      const float *vs, float *gvp, float *gvs, float *grho, hipStream_t st);
int launch_adj2d_inject(const sw_adj2d *q, const sw_receivers *r, int it, hipStream_t st);
int launch_adj2d_model_gradient(const sw_ssg2d *f, const sw_adj2d *q, const float *vp, const float *vs, float *gvp,
                                float *gvs, float *grho, hipStream_t st);

static int check_axis(const sw_axis &a, int n, const char *name, bool need_cpml, bool need_split)
{
    SW_REQUIRE(a.lo >= 0 && a.lo <= a.hi && a.hi <= n, "axis %s: slab bounds lo=%d hi=%d outside [0,%d]",
               name, a.lo, a.hi, n);
    if (need_cpml)
        SW_REQUIRE(a.a && a.b && a.k && a.a_h && a.b_h && a.k_h, "axis %s: CPML coefficient arrays missing", name);
    if (need_split)
        SW_REQUIRE(a.d && a.d_h && a.in && a.in_h, "axis %s: split-field PML arrays missing", name);
    return SW_OK;
}

int check_ssg3d(const sw_ssg3d *g)
{
    SW_REQUIRE(g != nullptr, "null descriptor");
    SW_REQUIRE(g->cls == SW_CLASS_CPML || g->cls == SW_CLASS_TENSOR || g->cls == SW_CLASS_SPML,
               "unknown 3-D class %d", g->cls);
    // the split-field class keeps split arrays and damping profiles, the
    // others CPML memory slabs and coefficients
    const bool spml = g->cls == SW_CLASS_SPML;
    SW_REQUIRE(g->order >= 1 && g->order <= SW_MAX_ORDER, "order %d outside 1..%d", g->order, SW_MAX_ORDER);
    const int need = 2 * g->order + 2;
    SW_REQUIRE(g->nx >= need && g->ny >= need && g->nz >= need,
               "grid %dx%dx%d too small for order %d (each axis needs >= %d)", g->nx, g->ny, g->nz,
               g->order, need);
    // The sweeps address a z chunk (kZChunk planes) and its stencil halo
    // with 32-bit byte offsets from pointers rebased per block, the z memory
    // slab with 32-bit offsets from its base, and receivers by int32 cells.
    const long cells = (long)g->nx * g->ny * g->nz, plane = 4L * g->nx * g->ny;
    SW_REQUIRE(cells <= INT32_MAX, "grid %dx%dx%d: %ld cells, above the 2^31-1 an int32 cell index reaches",
               g->nx, g->ny, g->nz, cells);
    SW_REQUIRE(((long)kZChunk + 2 * g->order + 1) * plane <= (1L << 32),
               "planes of %dx%d cells too large: a z chunk of the sweeps must stay below 4 GiB", g->nx, g->ny);
    // the six memory variables of each axis share one 32-bit offset range
    const long wx = g->px.lo + g->nx - g->px.hi, wy = g->py.lo + g->ny - g->py.hi, wz = g->pz.lo + g->nz - g->pz.hi;
    SW_REQUIRE(24L * wx * g->ny * g->nz <= (1L << 32) && 24L * wy * g->nx * g->nz <= (1L << 32) &&
                   24L * wz * g->nx * g->ny <= (1L << 32),
               "CPML memory slabs of the %dx%dx%d grid above 4 GiB per axis", g->nx, g->ny, g->nz);
    SW_REQUIRE(g->rho && g->lam && g->mu, "model arrays missing");
    SW_REQUIRE(g->vx && g->vy && g->vz && g->sxx && g->syy && g->szz && g->sxy && g->sxz && g->syz,
               "wavefield arrays missing");
    int rc;
    if ((rc = check_axis(g->px, g->nx, "x", !spml, spml)) != SW_OK) return rc;
    if ((rc = check_axis(g->py, g->ny, "y", !spml, spml)) != SW_OK) return rc;
    if ((rc = check_axis(g->pz, g->nz, "z", !spml, spml)) != SW_OK) return rc;
    if (spml) {
        SW_REQUIRE(g->split != nullptr, "split-field arrays missing");
        // the sweeps move the split values with 16-byte accesses
        SW_REQUIRE(reinterpret_cast<uintptr_t>(g->split) % 16 == 0, "split-field storage not 16-byte aligned");
    } else {
        SW_REQUIRE(g->px.lo + g->nx - g->px.hi == 0 || g->mem_x, "x memory slab missing");
        SW_REQUIRE(g->py.lo + g->ny - g->py.hi == 0 || g->mem_y, "y memory slab missing");
        SW_REQUIRE(g->pz.lo + g->nz - g->pz.hi == 0 || g->mem_z, "z memory slab missing");
    }
    return SW_OK;
}

int check_ssg2d(const sw_ssg2d *g)
{
    SW_REQUIRE(g != nullptr, "null descriptor");
    SW_REQUIRE(g->order >= 1 && g->order <= SW_MAX_ORDER, "order %d outside 1..%d", g->order, SW_MAX_ORDER);
    SW_REQUIRE(g->scheme == SW_2D_VS || g->scheme == SW_2D_SV || g->scheme == SW_2D_RSG, "unknown 2-D scheme %d",
               g->scheme);
    const int need = 2 * g->order + 2;
    SW_REQUIRE(g->nx >= need && g->nz >= need, "grid %dx%d too small for order %d", g->nx, g->nz, g->order);
    SW_REQUIRE(g->rho && g->lam && g->mu && g->vx && g->vz && g->sxx && g->szz && g->sxz && g->mem,
               "2-D arrays missing");
    int rc;
    if ((rc = check_axis(g->px, g->nx, "x", true, false)) != SW_OK) return rc;
    if ((rc = check_axis(g->pz, g->nz, "z", true, false)) != SW_OK) return rc;
    SW_REQUIRE(g->px.in && g->px.in_h && g->pz.in && g->pz.in_h, "2-D in-PML flags missing");
    return SW_OK;
}

namespace {

// ------------------------------------------------------------------ sources
__global__ void k_inject3d(sw_ssg3d g, sw_source s, int it)
{
    const float amp = s.amp[it];
    const long p = ((long)s.iz * g.ny + s.iy) * g.nx + s.ix;
    switch (s.kind) {
    case SW_SRC_FORCE_VZ:
    case SW_SRC_FORCE_VX: force_src3d(g, s.kind, amp, p); break;
    case SW_SRC_MOMENT:
    case SW_SRC_EXPLOSIVE: stress_src3d(g, s.kind, s.mt, amp, p); break;
    default: break;
    }
}

__device__ __forceinline__ void inject2d(sw_ssg2d &g, const sw_source &s, int it)
{
    const float amp = s.amp[it];
    const long p = (long)s.iz * g.nx + s.ix;
    if (g.scheme == SW_2D_RSG) {
        // src/Elastic2D_CPML.py:322-323: vz[isx+a, isz+b] += dt*source/(8*rho), a,b in {-1,0}
        for (int a = -1; a < 1; ++a)
            for (int b = -1; b < 1; ++b) {
                const long q = p + a + (long)b * g.nx;
                g.vz[q] += g.dt * amp / (8.f * g.rho[q]);
            }
        return;
    }
    switch (s.kind) {
    case SW_SRC_FORCE_VZ:
    case SW_SRC_FORCE_VX: force_src2d(g, s.kind, amp, p); break;
    case SW_SRC_MOMENT:
    case SW_SRC_EXPLOSIVE: stress_src2d(g, s.kind, s.mt[0], s.mt[2], s.mt[4], amp, p); break;
    default: break;
    }
}

__global__ void k_inject2d(sw_ssg2d g, sw_source s, int it)
{
    inject2d(g, s, it);
}

// Source cells of the shots of a batch (at most SW_MAX_BATCH).
struct ShotCells {
    int ix[SW_MAX_BATCH], iz[SW_MAX_BATCH];
};

// One block per shot: shot b's state at + b * bs, its source at cell b.
__global__ void k_inject2d_batch(sw_ssg2d g, sw_source s, ShotCells sc, int it, long bs)
{
    const int b = blockIdx.x;
    shift_state(g, (long)b * bs);
    s.ix = sc.ix[b];
    s.iz = sc.iz[b];
    inject2d(g, s, it);
}

// Adjoint of a force dt*amp[it]/rho[p] added to a velocity between the
// sweeps: the state's adjoint passes through unchanged, and d misfit /
// d amp[it] is the adjoint of that velocity at p, accumulated here times
// amp[it] (one block per shot, steps in a fixed order).
__global__ void k_adj_force_batch(sw_adj2d q, sw_source s, ShotCells sc, int it, long as, int nx)
{
    const int b = blockIdx.x;
    const float *lam = (s.kind == SW_SRC_FORCE_VZ ? q.vz : q.vx) + (long)b * as;
    q.g_src[b] += s.amp[it] * lam[(long)sc.iz[b] * nx + sc.ix[b]];
}

inline bool force_source(const sw_source *s) { return s->kind == SW_SRC_FORCE_VZ || s->kind == SW_SRC_FORCE_VX; }
inline bool stress_source(const sw_source *s) { return s->kind == SW_SRC_MOMENT || s->kind == SW_SRC_EXPLOSIVE; }
// What of a 2-D step's source injection and receiver gather the sweeps can
// do themselves (Fold2D): not on the rotated grid (its force spreads over
// four cells), only for source cells that own a sweep thread ([N, n-N) on
// both axes), receivers recording velocities only, and a next-step stress
// source only when the gather is folded too (a separate gather after the
// stress sweep would see it).  Unset parts run as separate launches.
Fold2D plan_fold2d(const sw_ssg2d *g, const sw_source *src, const sw_receivers *rec, int nbatch, const int *ix,
                   const int *iz, long data_stride)
{
    Fold2D fo{};
    const int N = g->order;
    bool cells_ok = g->scheme != SW_2D_RSG && src->kind != SW_SRC_NONE;
    for (int b = 0; b < nbatch && cells_ok; ++b) {
        cells_ok = ix[b] >= N && ix[b] < g->nx - N && iz[b] >= N && iz[b] < g->nz - N;
        fo.ix[b] = ix[b];
        fo.iz[b] = iz[b];
    }
    bool rec_ok = rec && rec->nrec > 0;
    for (int c = 0; rec_ok && c < rec->ncomp; ++c) rec_ok = rec->comp[c] == 0 || rec->comp[c] == 2;
    if (rec_ok) {
        fo.record = 1;
        fo.rec = *rec;
        fo.ds = data_stride;
    }
    const bool no_gather_after = !rec || rec->nrec == 0 || rec_ok;
    if (cells_ok && force_source(src)) fo.force = src->kind;
    if (cells_ok && stress_source(src) && no_gather_after) fo.stress = src->kind;
    fo.amp = src->amp;
    fo.m0 = src->mt[0];
    fo.m2 = src->mt[2];
    fo.m4 = src->mt[4];
    return fo;
}

// What of a 3-D step the sweeps can do themselves (Fold3D): a source cell
// that owns a sweep thread ([N, n-N) on every axis), receivers recording
// velocities only, and a next-step stress source only when the gather is
// folded too and no free surface is zeroed between the steps (the reference
// zeroes szz before it injects).
Fold3D plan_fold3d(const sw_ssg3d *g, const sw_source *src, const sw_receivers *rec, int flags)
{
    Fold3D fo{};
    const int N = g->order;
    const bool cell_ok = src->kind != SW_SRC_NONE && src->ix >= N && src->ix < g->nx - N && src->iy >= N &&
                         src->iy < g->ny - N && src->iz >= N && src->iz < g->nz - N;
    bool rec_ok = rec && rec->nrec > 0;
    for (int c = 0; rec_ok && c < rec->ncomp; ++c) rec_ok = rec->comp[c] >= 0 && rec->comp[c] <= 2;
    if (rec_ok) {
        fo.record = 1;
        fo.rec = *rec;
        fo.gather = (int)std::min<long>(64, ((long)rec->nrec * rec->ncomp + 255) / 256);
    }
    const bool no_gather_after = !rec || rec->nrec == 0 || rec_ok;
    if (cell_ok && force_source(src)) fo.force = src->kind;
    if (cell_ok && stress_source(src) && no_gather_after && !(flags & SW_RUN_FREE_SURFACE)) fo.stress = src->kind;
    fo.amp = src->amp;
    for (int t = 0; t < 6; ++t) fo.mt[t] = src->mt[t];
    fo.ix = src->ix;
    fo.iy = src->iy;
    fo.iz = src->iz;
    return fo;
}

// The stress-source fold injects step it+1's source in step it's stress
// sweep: the first step of a call injects its own, the last one nothing.
inline void fold_step(Fold2D &fo, const Fold2D &plan, int it, int it_end)
{
    fo.force = plan.force;
    fo.it_force = it;
    fo.stress = it + 1 < it_end ? plan.stress : 0;
    fo.it_next = it + 1;
    fo.record = plan.record;
    fo.it_rec = it;
}

// Receivers of shot blockIdx.y: state at + y * bs, traces at + y * ds.
__global__ void k_record2d_batch(sw_ssg2d g, sw_receivers r, int it, long bs, long ds)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= r.nrec * r.ncomp) return;
    shift_state(g, (long)blockIdx.y * bs);
    const int comp = t / r.nrec, rr = t - comp * r.nrec;
    const int code = r.comp[comp];
    const float *f = code == 0 ? g.vx : code == 2 ? g.vz : code == 3 ? g.sxx : code == 5 ? g.szz : g.sxz;
    r.data[(long)blockIdx.y * ds + ((long)it * r.ncomp + comp) * r.nrec + rr] = f[r.idx[rr]];
}

// ---------------------------------------------------------------- receivers
struct FieldPtrs {
    const float *f[9];
};

__global__ void k_record(FieldPtrs fp, sw_receivers r, int it)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= r.nrec * r.ncomp) return;
    const int comp = t / r.nrec, rr = t - comp * r.nrec;
    r.data[((long)it * r.ncomp + comp) * r.nrec + rr] = fp.f[r.comp[comp]][r.idx[rr]];
}

int check_source3d(const sw_ssg3d *g, const sw_source *s, int it)
{
    SW_REQUIRE(s != nullptr, "null source");
    if (s->kind == SW_SRC_NONE) return SW_OK;
    SW_REQUIRE(s->kind == SW_SRC_FORCE_VZ || s->kind == SW_SRC_FORCE_VX || s->kind == SW_SRC_MOMENT ||
                   s->kind == SW_SRC_EXPLOSIVE,
               "unknown source kind %d", s->kind);
    SW_REQUIRE(s->amp != nullptr, "source amplitude array missing");
    SW_REQUIRE(it >= 0 && it < s->namp, "step %d outside the source amplitude array (%d steps)", it, s->namp);
    SW_REQUIRE(s->ix >= 0 && s->ix < g->nx && s->iy >= 0 && s->iy < g->ny && s->iz >= 0 && s->iz < g->nz,
               "source cell (%d,%d,%d) outside the %dx%dx%d grid", s->ix, s->iy, s->iz, g->nx, g->ny, g->nz);
    return SW_OK;
}

int check_source2d(const sw_ssg2d *g, const sw_source *s, int it)
{
    SW_REQUIRE(s != nullptr, "null source");
    if (s->kind == SW_SRC_NONE) return SW_OK;
    SW_REQUIRE(s->amp != nullptr, "source amplitude array missing");
    SW_REQUIRE(it >= 0 && it < s->namp, "step %d outside the source amplitude array (%d steps)", it, s->namp);
    const int lo = g->scheme == SW_2D_RSG ? 1 : 0;
    SW_REQUIRE(s->ix >= lo && s->ix < g->nx && s->iz >= lo && s->iz < g->nz,
               "source cell (%d,%d) outside the %dx%d grid", s->ix, s->iz, g->nx, g->nz);
    return SW_OK;
}

int check_receivers(const sw_receivers *r, int it)
{
    SW_REQUIRE(r != nullptr, "null receivers");
    if (r->nrec == 0) return SW_OK;
    SW_REQUIRE(r->nrec > 0 && r->idx && r->data, "receiver arrays missing");
    SW_REQUIRE(r->ncomp >= 1 && r->ncomp <= 9, "ncomp %d outside 1..9", r->ncomp);
    for (int c = 0; c < r->ncomp; ++c) SW_REQUIRE(r->comp[c] >= 0 && r->comp[c] < 9, "bad component code");
    SW_REQUIRE(it >= 0 && it < r->nt, "step %d outside the receiver record (%d steps)", it, r->nt);
    return SW_OK;
}

int record_launch(FieldPtrs fp, const sw_receivers *r, int it, hipStream_t st)
{
    if (r->nrec == 0) return SW_OK;
    const int n = r->nrec * r->ncomp;
    hipLaunchKernelGGL(k_record, dim3((n + 255) / 256), dim3(256), 0, st, fp, *r, it);
    return launch_status("receiver gather");
}

FieldPtrs fields3d(const sw_ssg3d *g)
{
    return FieldPtrs{{g->vx, g->vy, g->vz, g->sxx, g->syy, g->szz, g->sxy, g->sxz, g->syz}};
}

FieldPtrs fields2d(const sw_ssg2d *g)
{
    return FieldPtrs{{g->vx, nullptr, g->vz, g->sxx, nullptr, g->szz, nullptr, g->sxz, nullptr}};
}

__global__ void k_free_surface_init3d(int nx, int ny, int order, float *rho, float *lam)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)nx * ny * order;
    if (t >= n) return;  // planes k < order are the first order*ny*nx elements
    lam[t] = 0.f;
    rho[t] = 0.5f * rho[t];
}

__global__ void k_free_surface_init2d(int nx, int nz, int order, float *rho, float *lam)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nx * nz) return;
    const int i = (int)(t % nx);
    if (i >= order) return;
    lam[t] = 0.f;
    rho[t] = 0.5f * rho[t];
}

__global__ void k_free_surface2d(int nx, int nz, int order, float *sxx)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nz) sxx[(long)j * nx + order] = 0.f;
}

// src/Elastic3D_SSG_Tensor.py:159-163: szz[i, j, order] = 0 (plane k = order)
__global__ void k_free_surface3d(int nx, int ny, int order, float *szz)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long)nx * ny) szz[(long)order * nx * ny + t] = 0.f;
}

int free_surface3d_launch(const sw_ssg3d *g, hipStream_t st)
{
    const long n = (long)g->nx * g->ny;
    hipLaunchKernelGGL(k_free_surface3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g->nx, g->ny, g->order,
                       g->szz);
    return launch_status("3-D free surface");
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int sw_version(void) { return (0 << 16) | (1 << 8) | 0; }

int sw_last_error(char *buf, size_t n)
{
    const size_t len = strlen(g_err);
    if (buf && n) {
        strncpy(buf, g_err, n - 1);
        buf[n - 1] = '\0';
    }
    return (int)len;
}

int sw_device_init(int device)
{
    SW_TRY_HIP(hipSetDevice(device));
    SW_TRY_HIP(hipFree(nullptr));
    float *scratch = nullptr;  // the allocation-free entry point's scratch address on this device
    return mlp_scratch_lookup(&scratch);
}

int sw_ssg3d_velocity(const sw_ssg3d *g, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    SW_REQUIRE(g->cls != SW_CLASS_SPML, "velocity sweep needs a CPML/Tensor descriptor (split-field: sw_ssg3d_velocity_spml)");
    return launch_ssg3d_velocity(g, as_stream(stream));
}

int sw_ssg3d_stress(const sw_ssg3d *g, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    SW_REQUIRE(g->cls != SW_CLASS_SPML, "stress sweep needs a CPML/Tensor descriptor (split-field: sw_ssg3d_stress_spml)");
    return launch_ssg3d_stress(g, as_stream(stream));
}

int sw_ssg3d_velocity_spml(const sw_ssg3d *g, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    SW_REQUIRE(g->cls == SW_CLASS_SPML, "split-field velocity sweep needs a split-field descriptor");
    return launch_ssg3d_velocity_spml(g, as_stream(stream));
}

int sw_ssg3d_stress_spml(const sw_ssg3d *g, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    SW_REQUIRE(g->cls == SW_CLASS_SPML, "split-field stress sweep needs a split-field descriptor");
    return launch_ssg3d_stress_spml(g, as_stream(stream));
}

int sw_set_launch_plan(int32_t mode)
{
    SW_REQUIRE(mode == SW_PLAN_AUTO || mode == SW_PLAN_ONE || mode == SW_PLAN_BOXES, "unknown launch plan mode %d",
               mode);
    return sw::g_plan_mode.exchange(mode);
}

int sw_ssg3d_launch_plan(const sw_ssg3d *g, int32_t *pm, int32_t max)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    SW_REQUIRE(max >= 0 && (max == 0 || pm), "null mask array");
    return ssg3d_plan_masks(g, pm, max);
}

// Shell storage of the split-field class: host-only, from the grid shape and
// the axes' lo/hi (no device pointer is read).
static int check_split_shape(const sw_ssg3d *g)
{
    SW_REQUIRE(g != nullptr, "null descriptor");
    SW_REQUIRE(g->order >= 1 && g->order <= SW_MAX_ORDER, "order %d outside 1..%d", g->order, SW_MAX_ORDER);
    SW_REQUIRE(g->nx >= 2 * g->order + 2 && g->ny >= 2 * g->order + 2 && g->nz >= 2 * g->order + 2,
               "grid %dx%dx%d too small for order %d", g->nx, g->ny, g->nz, g->order);
    const sw_axis *ax[3] = {&g->px, &g->py, &g->pz};
    const int n[3] = {g->nx, g->ny, g->nz};
    for (int d = 0; d < 3; ++d)
        SW_REQUIRE(ax[d]->lo >= 0 && ax[d]->lo <= ax[d]->hi && ax[d]->hi <= n[d],
                   "axis %c: split-free range lo=%d hi=%d outside [0,%d]", "xyz"[d], ax[d]->lo, ax[d]->hi, n[d]);
    return SW_OK;
}

int64_t sw_ssg3d_split_cells(const sw_ssg3d *g)
{
    int rc = check_split_shape(g);
    if (rc) return rc;
    return split_shell(*g).S;
}

int64_t sw_ssg3d_split_map(const sw_ssg3d *g, int32_t sweep, int64_t *cell, int64_t *slot, int64_t max)
{
    int rc = check_split_shape(g);
    if (rc) return rc;
    SW_REQUIRE(sweep == 0 || sweep == 1, "sweep %d: 0 (velocity) or 1 (stress)", sweep);
    SW_REQUIRE(max >= 0 && (max == 0 || (cell && slot)), "null output arrays");
    return split_shell_map(g, sweep, cell, slot, max);
}

int sw_ssg3d_inject(const sw_ssg3d *g, const sw_source *src, int32_t it, void *stream)
{
    int rc = check_source3d(g, src, it);
    if (rc) return rc;
    if (src->kind == SW_SRC_NONE) return SW_OK;
    hipLaunchKernelGGL(k_inject3d, dim3(1), dim3(1), 0, as_stream(stream), *g, *src, it);
    return launch_status("source injection");
}

int sw_ssg3d_record(const sw_ssg3d *g, const sw_receivers *rec, int32_t it, void *stream)
{
    int rc = check_receivers(rec, it);
    if (rc) return rc;
    return record_launch(fields3d(g), rec, it, as_stream(stream));
}

int sw_ssg3d_run(const sw_ssg3d *g, const sw_source *src, const sw_receivers *rec, int32_t it0, int32_t nsteps,
                 void *stream)
{
    return sw_ssg3d_run_ex(g, src, rec, it0, nsteps, 0, stream);
}

int sw_ssg3d_run_ex(const sw_ssg3d *g, const sw_source *src, const sw_receivers *rec, int32_t it0,
                    int32_t nsteps, int32_t flags, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    SW_REQUIRE(nsteps >= 0 && it0 >= 0, "bad step range");
    if (nsteps == 0) return SW_OK;
    if ((rc = check_source3d(g, src, it0)) || (rc = check_source3d(g, src, it0 + nsteps - 1))) return rc;
    if (rec && ((rc = check_receivers(rec, it0)) || (rc = check_receivers(rec, it0 + nsteps - 1)))) return rc;
    SW_REQUIRE((flags & ~SW_RUN_FREE_SURFACE) == 0, "unknown run flags 0x%x", flags);
    hipStream_t st = as_stream(stream);
    const FieldPtrs fp = fields3d(g);
    const int it_end = it0 + nsteps;
    const Fold3D plan = plan_fold3d(g, src, rec, flags);
    Fold3D fo = plan;
    for (int it = it0; it < it_end; ++it) {
        fo.it_force = it;
        fo.stress = it + 1 < it_end ? plan.stress : 0;
        fo.it_next = it + 1;
        fo.it_rec = it;
        if ((flags & SW_RUN_FREE_SURFACE) && (rc = free_surface3d_launch(g, st))) return rc;
        if (stress_source(src) && (it == it0 || !plan.stress))
            hipLaunchKernelGGL(k_inject3d, dim3(1), dim3(1), 0, st, *g, *src, it);
        if ((rc = launch_ssg3d_velocity_fold(g, fo, st))) return rc;
        if (force_source(src) && !plan.force) hipLaunchKernelGGL(k_inject3d, dim3(1), dim3(1), 0, st, *g, *src, it);
        if ((rc = launch_ssg3d_stress_fold(g, fo, st))) return rc;
        if (rec && !plan.record && (rc = record_launch(fp, rec, it, st))) return rc;
    }
    return launch_status("ssg3d run");
}

int sw_ssg2d_velocity(const sw_ssg2d *g, void *stream)
{
    int rc = check_ssg2d(g);
    if (rc) return rc;
    return launch_ssg2d_velocity(g, as_stream(stream));
}

int sw_ssg2d_stress(const sw_ssg2d *g, void *stream)
{
    int rc = check_ssg2d(g);
    if (rc) return rc;
    return launch_ssg2d_stress(g, as_stream(stream));
}

int sw_ssg2d_inject(const sw_ssg2d *g, const sw_source *src, int32_t it, void *stream)
{
    int rc = check_ssg2d(g);
    if (rc) return rc;
    if ((rc = check_source2d(g, src, it))) return rc;
    if (src->kind == SW_SRC_NONE) return SW_OK;
    hipLaunchKernelGGL(k_inject2d, dim3(1), dim3(1), 0, as_stream(stream), *g, *src, it);
    return launch_status("2-D source injection");
}

int sw_ssg2d_record(const sw_ssg2d *g, const sw_receivers *rec, int32_t it, void *stream)
{
    int rc = check_receivers(rec, it);
    if (rc) return rc;
    for (int c = 0; c < rec->ncomp; ++c)
        SW_REQUIRE(fields2d(g).f[rec->comp[c]] != nullptr, "component %d does not exist in 2-D", rec->comp[c]);
    return record_launch(fields2d(g), rec, it, as_stream(stream));
}

int sw_ssg2d_run(const sw_ssg2d *g, const sw_source *src, const sw_receivers *rec, int32_t it0, int32_t nsteps,
                 void *stream)
{
    int rc = check_ssg2d(g);
    if (rc) return rc;
    SW_REQUIRE(nsteps >= 0 && it0 >= 0, "bad step range");
    if (nsteps == 0) return SW_OK;
    if ((rc = check_source2d(g, src, it0)) || (rc = check_source2d(g, src, it0 + nsteps - 1))) return rc;
    if (rec) {
        if ((rc = check_receivers(rec, it0)) || (rc = check_receivers(rec, it0 + nsteps - 1))) return rc;
        for (int c = 0; c < rec->ncomp; ++c)
            SW_REQUIRE(fields2d(g).f[rec->comp[c]] != nullptr, "component %d does not exist in 2-D", rec->comp[c]);
    }
    hipStream_t st = as_stream(stream);
    const FieldPtrs fp = fields2d(g);
    const int ix = src->ix, iz = src->iz, it_end = it0 + nsteps;
    const Fold2D plan = plan_fold2d(g, src, rec, 1, &ix, &iz, 0);
    Fold2D fo = plan;
    for (int it = it0; it < it_end; ++it) {
        fold_step(fo, plan, it, it_end);
        if (stress_source(src) && (it == it0 || !plan.stress))
            hipLaunchKernelGGL(k_inject2d, dim3(1), dim3(1), 0, st, *g, *src, it);
        if ((rc = launch_ssg2d_velocity_batch(g, 1, 0, fo, st))) return rc;
        if (force_source(src) && !plan.force) hipLaunchKernelGGL(k_inject2d, dim3(1), dim3(1), 0, st, *g, *src, it);
        if ((rc = launch_ssg2d_stress_batch(g, 1, 0, fo, st))) return rc;
        if (rec && !plan.record && (rc = record_launch(fp, rec, it, st))) return rc;
    }
    return launch_status("ssg2d run");
}

int sw_ssg2d_free_surface_init(const sw_ssg2d *g, void *stream)
{
    SW_REQUIRE(g && g->rho && g->lam, "free-surface init needs rho and lam");
    SW_REQUIRE(g->nx > 0 && g->nz > 0 && g->order >= 1 && g->order <= g->nx, "bad 2-D shape/order");
    const long n = (long)g->nx * g->nz;
    hipLaunchKernelGGL(k_free_surface_init2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream),
                       g->nx, g->nz, g->order, g->rho, g->lam);
    return launch_status("2-D free-surface init");
}

int sw_ssg2d_free_surface(const sw_ssg2d *g, void *stream)
{
    SW_REQUIRE(g && g->sxx, "free surface needs sxx");
    SW_REQUIRE(g->nx > 0 && g->nz > 0 && g->order >= 0 && g->order < g->nx, "bad 2-D shape/order");
    hipLaunchKernelGGL(k_free_surface2d, dim3((g->nz + 255) / 256), dim3(256), 0, as_stream(stream), g->nx, g->nz,
                       g->order, g->sxx);
    return launch_status("2-D free surface");
}

int sw_ssg3d_free_surface(const sw_ssg3d *g, void *stream)
{
    SW_REQUIRE(g && g->szz, "null argument");
    SW_REQUIRE(g->order >= 1 && g->order < g->nz, "bad order");
    return free_surface3d_launch(g, as_stream(stream));
}

int sw_ssg3d_free_surface_init(const sw_ssg3d *g, float *rho, float *lam, void *stream)
{
    SW_REQUIRE(g && rho && lam, "null argument");
    SW_REQUIRE(g->order >= 1 && g->order <= g->nz, "bad order");
    const long n = (long)g->nx * g->ny * g->order;
    hipLaunchKernelGGL(k_free_surface_init3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream),
                       g->nx, g->ny, g->order, rho, lam);
    return launch_status("3-D free-surface init");
}

static int check_adj2d(const sw_ssg2d *g, const sw_adj2d *q)
{
    int rc = check_ssg2d(g);
    if (rc) return rc;
    SW_REQUIRE(g->scheme == SW_2D_VS, "the 2-D adjoint covers the VS scheme");
    SW_REQUIRE(q && q->vx && q->vz && q->sxx && q->szz && q->sxz && q->mem && q->tmp && q->g_b1 && q->g_b4 &&
                   q->g_lamh && q->g_muh && q->g_muz,
               "adjoint arrays missing");
    return SW_OK;
}

int sw_ssg2d_adjoint_stress(const sw_ssg2d *fwd, const sw_adj2d *adj, void *stream)
{
    int rc = check_adj2d(fwd, adj);
    if (rc) return rc;
    return launch_adj2d_stress(fwd, adj, as_stream(stream));
}

int sw_ssg2d_adjoint_velocity(const sw_ssg2d *fwd, const sw_adj2d *adj, void *stream)
{
    int rc = check_adj2d(fwd, adj);
    if (rc) return rc;
    return launch_adj2d_velocity(fwd, adj, as_stream(stream));
}

int sw_ssg2d_adjoint_inject(const sw_ssg2d *g, const sw_adj2d *adj, const sw_receivers *residual, int32_t it,
                            void *stream)
{
    int rc = check_adj2d(g, adj);
    if (rc) return rc;
    if ((rc = check_receivers(residual, it))) return rc;
    SW_REQUIRE(residual->ncomp <= 16, "too many residual components");
    for (int c = 0; c < residual->ncomp; ++c)
        SW_REQUIRE(fields2d(g).f[residual->comp[c]] != nullptr, "component %d does not exist in 2-D",
                   residual->comp[c]);
    if (residual->nrec == 0) return SW_OK;
    return launch_adj2d_inject(adj, residual, it, as_stream(stream));
}

int sw_ssg2d_model_gradient(const sw_ssg2d *g, const sw_adj2d *adj, const float *vp, const float *vs, float *gvp,
                            float *gvs, float *grho, void *stream)
{
    int rc = check_adj2d(g, adj);
    if (rc) return rc;
    SW_REQUIRE(vp && vs && gvp && gvs && grho, "model gradient arrays missing");
    return launch_adj2d_model_gradient(g, adj, vp, vs, gvp, gvs, grho, as_stream(stream));
}

int sw_ssg2d_adjoint_segment(const sw_ssg2d *g, const sw_source *src, const sw_adj2d *adj,
                             const sw_receivers *residual, float *work, int32_t it0, int32_t n, void *stream)
{
    SW_REQUIRE(g && src, "null descriptor or source");
    const int32_t cell[2] = {src->ix, src->iz};
    return sw_ssg2d_adjoint_segment_batch(g, 1, 15L * g->nx * g->nz, src, cell, adj, 0, residual, 0, work, it0, n,
                                          stream);
}

// ---- shot batches of the 2-D VS path ----
static int check_batch2d(const sw_ssg2d *g, int nbatch, long state_stride, const sw_source *src,
                         const int32_t *cells, ShotCells &sc)
{
    int rc = check_ssg2d(g);
    if (rc) return rc;
    SW_REQUIRE(nbatch >= 1 && nbatch <= SW_MAX_BATCH, "batch of %d shots outside 1..%d", nbatch, SW_MAX_BATCH);
    SW_REQUIRE(src && cells, "null source or shot cells");
    const long S = 15L * g->nx * g->nz;
    SW_REQUIRE(nbatch == 1 || state_stride >= S, "state stride %ld below one state (%ld floats)", state_stride, S);
    const long P = (long)g->nx * g->nz;
    SW_REQUIRE(g->vz == g->vx + P && g->sxx == g->vx + 2 * P && g->szz == g->vx + 3 * P &&
                   g->sxz == g->vx + 4 * P && g->mem == g->vx + 5 * P,
               "the state must be one block of 15 planes (vx vz sxx szz sxz mem) starting at vx");
    SW_REQUIRE(src->kind == SW_SRC_NONE || stress_source(src) || force_source(src),
               "shot batches take explosive, moment-tensor and force sources");
    for (int b = 0; b < nbatch; ++b) {
        const int ix = cells[2 * b], iz = cells[2 * b + 1];
        SW_REQUIRE(ix >= 1 && ix < g->nx - 1 && iz >= 1 && iz < g->nz - 1, "shot %d cell (%d,%d) outside the grid", b,
                   ix, iz);
        sc.ix[b] = ix;
        sc.iz[b] = iz;
    }
    return SW_OK;
}

static int check_batch_receivers(const sw_ssg2d *g, const sw_receivers *r, int it0, int n)
{
    int rc;
    if ((rc = check_receivers(r, it0)) || (rc = check_receivers(r, it0 + n - 1))) return rc;
    SW_REQUIRE(r->ncomp <= 16, "too many components");
    for (int c = 0; c < r->ncomp; ++c)
        SW_REQUIRE(fields2d(g).f[r->comp[c]] != nullptr, "component %d does not exist in 2-D", r->comp[c]);
    return SW_OK;
}

int sw_ssg2d_run_batch(const sw_ssg2d *g, int32_t nbatch, int64_t state_stride, const sw_source *src,
                       const int32_t *cells, const sw_receivers *rec, int64_t data_stride, int32_t it0,
                       int32_t nsteps, void *stream)
{
    ShotCells sc{};
    int rc = check_batch2d(g, nbatch, state_stride, src, cells, sc);
    if (rc) return rc;
    SW_REQUIRE(nsteps >= 0 && it0 >= 0, "bad step range");
    if (nsteps == 0) return SW_OK;
    if (src->kind != SW_SRC_NONE) {
        SW_REQUIRE(src->amp != nullptr && it0 + nsteps <= src->namp, "source amplitudes missing for the steps");
    }
    if (rec && (rc = check_batch_receivers(g, rec, it0, nsteps))) return rc;
    hipStream_t st = as_stream(stream);
    const int it_end = it0 + nsteps;
    const long bs = nbatch == 1 ? 0 : (long)state_stride, ds = nbatch == 1 ? 0 : (long)data_stride;
    const Fold2D plan = plan_fold2d(g, src, rec, nbatch, sc.ix, sc.iz, ds);
    Fold2D fo = plan;
    for (int it = it0; it < it_end; ++it) {
        fold_step(fo, plan, it, it_end);
        if (stress_source(src) && (it == it0 || !plan.stress))
            hipLaunchKernelGGL(k_inject2d_batch, dim3(nbatch), dim3(1), 0, st, *g, *src, sc, it, bs);
        if ((rc = launch_ssg2d_velocity_batch(g, nbatch, bs, fo, st))) return rc;
        if (force_source(src) && !plan.force)
            hipLaunchKernelGGL(k_inject2d_batch, dim3(nbatch), dim3(1), 0, st, *g, *src, sc, it, bs);
        if ((rc = launch_ssg2d_stress_batch(g, nbatch, bs, fo, st))) return rc;
        if (rec && rec->nrec && !plan.record) {
            const int n = rec->nrec * rec->ncomp;
            hipLaunchKernelGGL(k_record2d_batch, dim3((n + 255) / 256, nbatch), dim3(256), 0, st, *g, *rec, it, bs,
                               ds);
        }
    }
    return launch_status("ssg2d batch run");
}

int sw_ssg2d_adjoint_segment_batch(const sw_ssg2d *g, int32_t nbatch, int64_t state_stride, const sw_source *src,
                                   const int32_t *cells, const sw_adj2d *adj, int64_t adj_stride,
                                   const sw_receivers *residual, int64_t data_stride, float *work, int32_t it0,
                                   int32_t n, void *stream)
{
    ShotCells sc{};
    int rc = check_batch2d(g, nbatch, state_stride, src, cells, sc);
    if (rc) return rc;
    if ((rc = check_adj2d(g, adj))) return rc;
    SW_REQUIRE(residual && work, "null residual or work buffer");
    SW_REQUIRE(n >= 0 && it0 >= 0, "bad step range");
    if (n == 0) return SW_OK;
    if (src->kind != SW_SRC_NONE) {
        SW_REQUIRE(src->amp != nullptr && it0 + n <= src->namp, "source amplitudes missing for the steps");
    }
    if ((rc = check_batch_receivers(g, residual, it0, n))) return rc;
    SW_REQUIRE(!force_source(src) || adj->g_src, "force sources need the adj->g_src accumulator");
    const long P = (long)g->nx * g->nz;
    SW_REQUIRE(nbatch == 1 || adj_stride >= 24 * P, "adjoint stride %ld below one adjoint block (%ld floats)",
               (long)adj_stride, 24 * P);
    const long bs = nbatch == 1 ? 0 : (long)state_stride, as = nbatch == 1 ? 0 : (long)adj_stride;
    // Kept per step and shot: the 5 planes of CPML-corrected derivative sums
    // the saved sweeps write, all the model gradient needs of the forward
    // state; the adjoint stages themselves are linear in the adjoint state.
    const long save_n = 5 * P, step_n = (long)nbatch * save_n;
    hipStream_t st = as_stream(stream);
    const Fold2D plan = plan_fold2d(g, src, nullptr, nbatch, sc.ix, sc.iz, 0);
    Fold2D fo = plan;
    for (int t = 0; t < n; ++t) {
        float *sv = work + (size_t)t * step_n;
        const int it = it0 + t;
        fold_step(fo, plan, it, it0 + n);
        if (stress_source(src) && (t == 0 || !plan.stress))
            hipLaunchKernelGGL(k_inject2d_batch, dim3(nbatch), dim3(1), 0, st, *g, *src, sc, it, bs);
        if ((rc = launch_ssg2d_sweeps_saved(g, nbatch, bs, false, sv, save_n, fo, st))) return rc;
        if (force_source(src) && !plan.force)
            hipLaunchKernelGGL(k_inject2d_batch, dim3(nbatch), dim3(1), 0, st, *g, *src, sc, it, bs);
        if ((rc = launch_ssg2d_sweeps_saved(g, nbatch, bs, true, sv, save_n, fo, st))) return rc;
    }
    // Backwards through the steps.  With velocity residuals only, each stage's
    // gather runs fused with the next stage's per-cell pass (two launches per
    // step; the residuals enter q.vx, q.vz, which the fused stress pass does
    // not read); stress residuals take the four separate passes.
    bool fuse = true;
    for (int c = 0; c < residual->ncomp; ++c) fuse = fuse && (residual->comp[c] == 0 || residual->comp[c] == 2);
    float *tv = work + (size_t)n * step_n;  // velocity-stage pass-A outputs, 4 planes per shot
    const long tvs = 4 * P;
    auto inject = [&](int t) {
        return residual->nrec ? launch_adj2d_inject_batch(adj, residual, it0 + t, nbatch, as, data_stride, st) : SW_OK;
    };
    for (int t = n - 1; t >= 0; --t) {
        const float *sv = work + (size_t)t * step_n;
        if (!fuse) {
            if ((rc = inject(t))) return rc;
            if ((rc = launch_adj2d_stress_batch(g, adj, nbatch, 0, as, sv, save_n, st))) return rc;
            if (force_source(src))
                hipLaunchKernelGGL(k_adj_force_batch, dim3(nbatch), dim3(1), 0, st, *adj, *src, sc, it0 + t, as,
                                   g->nx);
            if ((rc = launch_adj2d_velocity_batch(g, adj, nbatch, 0, as, sv, save_n, st))) return rc;
            continue;
        }
        if (t == n - 1) {
            if ((rc = inject(t))) return rc;
            if ((rc = launch_adj2d_fused(g, adj, nbatch, as, 0, sv, save_n, tv, tvs, st))) return rc;
        }
        if ((rc = launch_adj2d_fused(g, adj, nbatch, as, 1, sv, save_n, tv, tvs, st))) return rc;
        if (force_source(src))
            hipLaunchKernelGGL(k_adj_force_batch, dim3(nbatch), dim3(1), 0, st, *adj, *src, sc, it0 + t, as, g->nx);
        if (t > 0) {
            if ((rc = inject(t - 1))) return rc;
            if ((rc = launch_adj2d_fused(g, adj, nbatch, as, 2, sv - step_n, save_n, tv, tvs, st))) return rc;
        } else if ((rc = launch_adj2d_fused(g, adj, nbatch, as, 3, sv, save_n, tv, tvs, st))) {
            return rc;
        }
    }
    return launch_status("ssg2d adjoint segment batch");
}

int sw_ssg2d_model_gradient_batch(const sw_ssg2d *g, const sw_adj2d *adj, int32_t nbatch, int64_t adj_stride,
                                  const float *vp, const float *vs, float *gvp, float *gvs, float *grho, void *stream)
{
    int rc = check_adj2d(g, adj);
    if (rc) return rc;
    SW_REQUIRE(nbatch >= 1 && nbatch <= SW_MAX_BATCH, "batch of %d shots outside 1..%d", nbatch, SW_MAX_BATCH);
    SW_REQUIRE(vp && vs && gvp && gvs && grho, "null model or gradient arrays");
    return launch_adj2d_model_gradient_batch(g, adj, nbatch, nbatch == 1 ? 0 : (long)adj_stride, vp, vs, gvp, gvs,
                                             grho, as_stream(stream));
}

static int check_band(const sw_ssg3d *g, const sw_band *b)
{
    SW_REQUIRE(b != nullptr, "null band");
    SW_REQUIRE(b->ex >= 2 * SW_BAND_DEPTH + 1 && b->ey >= 2 * SW_BAND_DEPTH + 1 && b->ez >= 2 * SW_BAND_DEPTH + 1,
               "interior box %dx%dx%d too small for a band of depth %d", b->ex, b->ey, b->ez, SW_BAND_DEPTH);
    SW_REQUIRE(b->ox >= 0 && b->oy >= 0 && b->oz >= 0 && b->ox + b->ex <= g->nx && b->oy + b->ey <= g->ny &&
                   b->oz + b->ez <= g->nz,
               "interior box outside the grid");
    return SW_OK;
}

int sw_mlp_param_count(void) { return mlp_param_count(); }

int sw_band_columns(const sw_band *b)
{
    SW_REQUIRE(b != nullptr, "null band");
    return learned_band_columns(b);
}

int sw_band_gather(const sw_ssg3d *g, const sw_band *b, float *out, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    if ((rc = check_band(g, b))) return rc;
    SW_REQUIRE(out != nullptr, "null output");
    return launch_band_gather(g, b, out, as_stream(stream));
}

int sw_band_correct(const sw_ssg3d *g, const sw_band *b, const float *params, float *prev, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    if ((rc = check_band(g, b))) return rc;
    SW_REQUIRE(params && prev, "null network or band state");
    return launch_band_correct(g, b, params, prev, as_stream(stream));
}

// A cell in the band of the learned boundary (the kD outermost cells of the
// interior box on each face), whose values the correction reads and writes.
static bool in_band(const sw_band *b, int i, int j, int k)
{
    const int a = i - b->ox, c = j - b->oy, e = k - b->oz;
    if (a < 0 || c < 0 || e < 0 || a >= b->ex || c >= b->ey || e >= b->ez) return false;
    return a < SW_BAND_DEPTH || c < SW_BAND_DEPTH || e < SW_BAND_DEPTH || b->ex - 1 - a < SW_BAND_DEPTH ||
           b->ey - 1 - c < SW_BAND_DEPTH || b->ez - 1 - e < SW_BAND_DEPTH;
}

// Steps with the band correction after every stress sweep.  A force source
// is folded into the velocity sweep as in sw_ssg3d_run_ex; a stress source
// into the previous step's stress sweep unless its cell is a band cell (the
// correction after that sweep must see the cell without the next step's
// injection).  Receivers are recorded after the correction (a receiver in the
// band records the corrected value).
int sw_ssg3d_run_learned(const sw_ssg3d *g, const sw_source *src, const sw_receivers *rec, const sw_band *b,
                         const float *params, float *prev, int32_t it0, int32_t nsteps, void *stream)
{
    int rc = check_ssg3d(g);
    if (rc) return rc;
    if ((rc = check_band(g, b))) return rc;
    SW_REQUIRE(params && prev, "null network or band state");
    SW_REQUIRE(nsteps >= 0 && it0 >= 0, "bad step range");
    if (nsteps == 0) return SW_OK;
    if ((rc = check_source3d(g, src, it0)) || (rc = check_source3d(g, src, it0 + nsteps - 1))) return rc;
    if (rec && ((rc = check_receivers(rec, it0)) || (rc = check_receivers(rec, it0 + nsteps - 1)))) return rc;
    hipStream_t st = as_stream(stream);
    const FieldPtrs fp = fields3d(g);
    const int it_end = it0 + nsteps;
    Fold3D plan = plan_fold3d(g, src, nullptr, 0);
    if (plan.stress && in_band(b, src->ix, src->iy, src->iz)) plan.stress = 0;
    Fold3D fo = plan;
    for (int it = it0; it < it_end; ++it) {
        fo.it_force = it;
        fo.stress = it + 1 < it_end ? plan.stress : 0;
        fo.it_next = it + 1;
        if (stress_source(src) && (it == it0 || !plan.stress))
            hipLaunchKernelGGL(k_inject3d, dim3(1), dim3(1), 0, st, *g, *src, it);
        if ((rc = launch_ssg3d_velocity_fold(g, fo, st))) return rc;
        if (force_source(src) && !plan.force) hipLaunchKernelGGL(k_inject3d, dim3(1), dim3(1), 0, st, *g, *src, it);
        if ((rc = launch_ssg3d_stress_fold(g, fo, st))) return rc;
        if ((rc = launch_band_correct(g, b, params, prev, st))) return rc;
        if (rec && (rc = record_launch(fp, rec, it, st))) return rc;
    }
    return launch_status("ssg3d learned run");
}

int sw_stream_probe(const float *const *in, int32_t nin, float *const *out, int32_t nout, int64_t n,
                    int32_t nblock, void *stream)
{
    SW_REQUIRE(in && out && nin >= 1 && nout >= 1 && n > 0 && n % 4 == 0 && nblock > 0,
               "bad stream probe arguments (n a positive multiple of 4)");
    for (int q = 0; q < nin; ++q)
        SW_REQUIRE(in[q] && reinterpret_cast<uintptr_t>(in[q]) % 16 == 0, "stream probe input %d not 16-byte aligned", q);
    for (int o = 0; o < nout; ++o)
        SW_REQUIRE(out[o] && reinterpret_cast<uintptr_t>(out[o]) % 16 == 0, "stream probe output %d not 16-byte aligned",
                   o);
    return launch_stream_probe(in, nin, out, nout, n, nblock, as_stream(stream));
}

int sw_mfma_probe(float *out, int32_t nblock, int32_t steps, void *stream)
{
    SW_REQUIRE(out != nullptr && nblock > 0 && steps > 0, "bad matrix-core probe arguments");
    return launch_mfma_probe(out, nblock, steps, as_stream(stream));
}

int sw_mlp_forward(const float *params, const float *X, int32_t B, float *H1, float *H2, float *Y, void *stream)
{
    SW_REQUIRE(params && X && H1 && H2 && Y && B > 0, "bad mlp forward arguments");
    return launch_mlp_forward(params, X, B, H1, H2, Y, as_stream(stream));
}

int sw_mlp_loss_backward(const float *params, const float *X, const float *H1, const float *H2, const float *Y,
                         const float *T, int32_t B, float *grad, float *work, float *loss_rows, void *stream)
{
    SW_REQUIRE(params && X && H1 && H2 && Y && T && grad && work && loss_rows && B > 0,
               "bad mlp backward arguments");
    return launch_mlp_loss_backward(params, X, H1, H2, Y, T, B, grad, work, loss_rows, as_stream(stream));
}

int sw_mlp_loss_backward_noalloc(const float *params, const float *X, const float *H1, const float *H2,
                                 const float *Y, const float *T, int32_t B, float *grad, float *loss_rows,
                                 void *stream)
{
    SW_REQUIRE(params && X && H1 && H2 && Y && T && grad && loss_rows && B > 0, "bad mlp backward arguments");
    return launch_mlp_loss_backward_noalloc(params, X, H1, H2, Y, T, B, grad, loss_rows, as_stream(stream));
}

int sw_adam(float *p, const float *g, float *m, float *v, int32_t n, float lr, float beta1, float beta2, float eps,
            int32_t t, void *stream)
{
    SW_REQUIRE(p && g && m && v && n > 0 && t >= 1, "bad adam arguments");
    return launch_adam(p, g, m, v, n, lr, beta1, beta2, eps, t, as_stream(stream));
}

}  // extern "C"
