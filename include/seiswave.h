/*
 * seiswave.h -- C ABI of libseiswave.so, the MI355X (gfx950) elastic
 * staggered-grid time stepper.
 *
 * The library replaces the per-step Taichi kernels of SeisWave_Taichi's
 * ElasticWAVE classes (paths relative to the reference root):
 *   src/Elastic3D_SSG_CPML.py    update_SSG        :144-309
 *   src/Elastic3D_SSG_Tensor.py  update_SSG        :166-345
 *   src/Elastic3D_SSG_SPML.py    update_SSG_SPML   :160-393
 *   src/Elastic2D_CPML.py        update_SSG_VS     :127-200
 *                                update_SSG_SV     :202-274
 *                                update_RSG        :276-353
 *                                AEA_init / AEA    :112-125
 * The reference binds these through Taichi's @ti.kernel launch; here the
 * Python classes in seiswave_taichi_amd bind them through ctypes.
 *
 * Conventions
 *   - Plain pointers and sizes only.  Every pointer in a descriptor is a
 *     device pointer owned by the caller; the library never frees caller
 *     memory and allocates nothing on the step path.
 *   - Device layout is x fastest: a 3-D field (nx,ny,nz) is stored as
 *     [nz][ny][nx], element (i,j,k) at (k*ny + j)*nx + i; a 2-D field
 *     (nx,nz) as [nz][nx], element (i,j) at j*nx + i.
 *   - Work is enqueued on `stream` (a hipStream_t passed as void*; NULL is
 *     the default stream).  No call synchronises the device.
 *   - Every function returns 0 on success and a negative SW_E* code on
 *     failure; sw_last_error() returns the message of the calling thread's
 *     last failure.
 */
#ifndef SEISWAVE_H
#define SEISWAVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW_MAX_ORDER 8

#define SW_OK 0
#define SW_EINVAL (-1) /* invalid argument or shape                      */
#define SW_EHIP (-2)   /* HIP runtime error (launch or device)            */

/* Which reference class's arithmetic a 3-D descriptor follows.  They differ
 * in the term order of the vz sum (src/Elastic3D_SSG_CPML.py:212 sums z,x,y;
 * src/Elastic3D_SSG_Tensor.py:243 sums x,y,z). */
#define SW_CLASS_CPML 0
#define SW_CLASS_TENSOR 1
#define SW_CLASS_SPML 2

/* One axis of PML coefficients as SetADEPML3D leaves them
 * (src/Elastic3D_SSG_CPML.py:364-522), all device arrays of length n.
 * a/b/k drive the CPML memory variables; d drives the split-field PML.
 * CPML: memory variables of this axis exist only for indices [0,lo) and
 * [hi,n): outside that range the coefficients are the identity (a = 0, k = 1)
 * and the reference's memory variable stays exactly 0.
 * Split-field PML: in[i] and in_h[i] are both 0 for every i in [lo,hi) (the
 * split-free interior of this axis; lo = hi = 0 if there is none). */
typedef struct sw_axis {
    const float *a, *b, *k;       /* integer points */
    const float *a_h, *b_h, *k_h; /* half points    */
    const float *d, *d_h;         /* damping profile (split-field PML) */
    const uint8_t *in, *in_h;     /* 1 where the kernel's f32 coordinate test puts the point in this axis's PML */
    int32_t lo, hi;
} sw_axis;

/* 3-D state.  mem_x holds 6 slab arrays [nz][ny][wx] back to back
 * (wx = px.lo + nx - px.hi), in the order
 *   x: sxx_dx sxy_dx sxz_dx dvx_dx dvy_dx dvz_dx
 * mem_y holds 6 arrays [nz][wy][nx]:
 *   y: sxy_dy syy_dy syz_dy dvy_dy dvx_dy dvz_dy
 * mem_z holds 6 arrays [wz][ny][nx]:
 *   z: sxz_dz syz_dz szz_dz dvz_dz dvx_dz dvy_dz
 * split (SPML only, 16-byte aligned) holds the 24 split components
 *   c:  0    1    2    3    4    5    6    7    8    9     10    11
 *       vx_x vx_y vx_z vy_x vy_y vy_z vz_x vz_y vz_z sxx_x sxx_y sxx_z
 *   c:  12    13    14    15    16    17    18    19    20    21    22    23
 *       syy_x syy_y syy_z sxy_x sxy_y sxz_x sxz_z syz_y syz_z szz_x szz_y szz_z
 * of the cells outside the split-free box [px.lo,px.hi) x [py.lo,py.hi) x
 * [pz.lo,pz.hi) only (the reference's full-grid split arrays are exactly 0
 * inside it).  Those S = sw_ssg3d_split_cells(g) cells are numbered (their
 * "slot") through three parts, with w_d = n_d - (hi_d - lo_d),
 * m_d = hi_d - lo_d and the slab index s(i) = i < lo ? i : i - (hi - lo):
 *   planes k outside [pz.lo,pz.hi):          (s(k)*ny + j)*nx + i
 *   then planes inside, rows outside:         (k-pz.lo)*wy*nx + s(j)*nx + i
 *   then those planes, rows inside, columns
 *   outside [px.lo,px.hi):                    (k-pz.lo)*my*wx + (j-py.lo)*wx + s(i)
 * (each part after the previous one's last slot).  The slots come in blocks
 * of 64, block b holding floats [1536b, 1536b + 1536) of split; component c
 * of slot 64b + l sits at float 1536b + off + w*l + (c - c0) of the group
 * (c0..c0+w-1 at offset off) that holds it:
 *   c 0-3 at 0, 4-7 at 256, 9-12 at 576, 13-16 at 832, 17-20 at 1088 (w 4);
 *   c 21-22 at 1344 (w 2); c 8 at 512, c 23 at 1472 (w 1).
 * Storage: 1536 * ceil(S / 64) floats. */
typedef struct sw_ssg3d {
    int32_t nx, ny, nz, order; /* order = accuracy N (half stencil), 1..SW_MAX_ORDER */
    int32_t cls;               /* SW_CLASS_* */
    float dx, dy, dz, dt;
    float c[SW_MAX_ORDER]; /* staggered coefficients, f32 (src/Elastic3D_SSG_CPML.py:311-322) */
    const float *rho, *lam, *mu;
    float *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *sxz, *syz;
    sw_axis px, py, pz;
    float *mem_x, *mem_y, *mem_z;
    float *split;
} sw_ssg3d;

/* Source kinds. */
#define SW_SRC_NONE 0
#define SW_SRC_FORCE_VZ 1  /* vz[p] += dt*amp[it]/rho[p] after the velocity sweep
                              (src/Elastic3D_SSG_CPML.py:245, src/Elastic3D_SSG_SPML.py:294) */
#define SW_SRC_MOMENT 2    /* s_ab[p] += amp[it]*(-M_ab) before the velocity sweep
                              (src/Elastic3D_SSG_Tensor.py:186-193) */
#define SW_SRC_EXPLOSIVE 3 /* sxx,syy,szz[p] += amp[it]*(-1/sqrt(3)) before the velocity sweep
                              (MOMENT with M = I/sqrt(3); no reference counterpart) */
#define SW_SRC_FORCE_VX 4  /* vx[p] += dt*amp[it]/rho[p] after the velocity sweep
                              (2-D SV, src/Elastic2D_CPML.py:245) */

typedef struct sw_source {
    int32_t kind;
    int32_t ix, iy, iz; /* cell (iy unused in 2-D) */
    float mt[6];        /* Mxx Myy Mzz Mxy Mxz Myz */
    const float *amp;   /* device, one amplitude per step index */
    int32_t namp;
} sw_source;

/* Receiver gather: data[it*nrec + r] = field[idx[r]] after the stress sweep
 * (src/Elastic3D_SSG_Tensor.py:344-345).  field: 0 vx 1 vy 2 vz 3 sxx 4 syy
 * 5 szz 6 sxy 7 sxz 8 syz (2-D: 0 vx 2 vz 3 sxx 5 szz 7 sxz). */
typedef struct sw_receivers {
    int32_t nrec;
    int32_t ncomp;          /* number of recorded components (1..9)       */
    int32_t comp[9];        /* field code per component                   */
    const int32_t *idx;     /* device, nrec linear device indices         */
    float *data;            /* device, [nt][ncomp][nrec]                  */
    int32_t nt;
} sw_receivers;

/* ---- library ---- */
int sw_version(void);                        /* (major<<16)|(minor<<8)|patch */
int sw_last_error(char *buf, size_t n);      /* copies the message; returns its length */
int sw_device_init(int device);              /* selects and warms the device; idempotent */

/* ---- 3-D sweeps (one reference loop group each) ---- */
/* Velocity loops of update_SSG: src/Elastic3D_SSG_CPML.py:164-212,
 * src/Elastic3D_SSG_Tensor.py:195-243.  Requires cls CPML or TENSOR. */
int sw_ssg3d_velocity(const sw_ssg3d *g, void *stream);
/* Stress loops of update_SSG: src/Elastic3D_SSG_CPML.py:247-307,
 * src/Elastic3D_SSG_Tensor.py:246-310. */
int sw_ssg3d_stress(const sw_ssg3d *g, void *stream);
/* Split-field velocity / stress loops: src/Elastic3D_SSG_SPML.py:182-259, :296-391. */
int sw_ssg3d_velocity_spml(const sw_ssg3d *g, void *stream);
int sw_ssg3d_stress_spml(const sw_ssg3d *g, void *stream);
/* Shell storage of the split-field class (no counterpart in the reference,
 * which keeps 24 full grids): floats per split component for this grid shape
 * and split-free box; host-only, reads only nx, ny, nz, order and the axes'
 * lo/hi.  Negative: SW_EINVAL. */
int64_t sw_ssg3d_split_cells(const sw_ssg3d *g);
/* The (cell, slot) pairs whose split values a split-field sweep (0 velocity,
 * 1 stress) updates, in its launches' logical block order: cell =
 * (k*ny + j)*nx + i, slot the cell's index in the shell storage (layout
 * above).  Writes at most max pairs and returns their total number;
 * host-only, for callers that read split values and for tests of the
 * layout. */
int64_t sw_ssg3d_split_map(const sw_ssg3d *g, int32_t sweep, int64_t *cell, int64_t *slot, int64_t max);
/* Launch plan of the two 3-D sweeps for this descriptor (no counterpart in
 * the reference, whose loops are one Taichi task each): returns the number of
 * launches per sweep (1, or 5: the CPML-free bulk with the x CPML only, the
 * y slabs, the z slabs; for the split-field class 2: the shell, mask 7, and
 * the split-free interior, mask 0) and writes up to max CPML axis masks
 * (bit 0 x, 1 y, 2 z: the axes whose CPML or split-field code the launch
 * instantiates).  Host only: no device access. */
int sw_ssg3d_launch_plan(const sw_ssg3d *g, int32_t *pm, int32_t max);
/* How the 3-D sweeps choose their launch plan, process-wide (no counterpart
 * in the reference).  SW_PLAN_AUTO (the default): five launches when the
 * CPML-free bulk is >= 75 % of the updated cells and every slab launch has
 * at least as many blocks as the device has compute units, else one.
 * SW_PLAN_ONE: always one launch.  SW_PLAN_BOXES: five launches whenever the
 * bulk box is not empty (for tests of the per-box kernels and for tuning).
 * Returns the previous mode, or SW_EINVAL for another value.  Calls that
 * overlap with running sweeps on other threads race. */
#define SW_PLAN_AUTO 0
#define SW_PLAN_ONE 1
#define SW_PLAN_BOXES 2
int sw_set_launch_plan(int32_t mode);
/* Source injection for step it (the serial statements of the reference kernels). */
int sw_ssg3d_inject(const sw_ssg3d *g, const sw_source *src, int32_t it, void *stream);
/* Receiver gather for step it (src/Elastic3D_SSG_CPML.py:309, Tensor :344-345). */
int sw_ssg3d_record(const sw_ssg3d *g, const sw_receivers *rec, int32_t it, void *stream);
/* nsteps full reference steps it0..it0+nsteps-1 in the order of g->cls:
 *   CPML/SPML: velocity, force, stress, record;  TENSOR: moment, velocity, stress, record.
 * rec may be NULL. */
int sw_ssg3d_run(const sw_ssg3d *g, const sw_source *src, const sw_receivers *rec,
                 int32_t it0, int32_t nsteps, void *stream);
/* sw_ssg3d_run with options (no counterpart in the reference's step methods):
 *   SW_RUN_FREE_SURFACE  sw_ssg3d_free_surface before every step, as the
 *                        reference's 2-D drivers call AEA() before each step
 *                        (Relay_wave.py:55-58). */
#define SW_RUN_FREE_SURFACE 1
int sw_ssg3d_run_ex(const sw_ssg3d *g, const sw_source *src, const sw_receivers *rec,
                    int32_t it0, int32_t nsteps, int32_t flags, void *stream);

/* ---- 2-D (src/Elastic2D_CPML.py) ---- */
#define SW_2D_VS 0  /* update_SSG_VS :127-200 */
#define SW_2D_SV 1  /* update_SSG_SV :202-274 */
#define SW_2D_RSG 2 /* update_RSG    :276-353 */

/* 2-D state.  mem holds 10 full-grid arrays [nz][nx] back to back:
 *   sxx_dx sxx_dz sxz_dx sxz_dz szz_dx szz_dz dvx_dx dvx_dz dvz_dx dvz_dz */
typedef struct sw_ssg2d {
    int32_t nx, nz, order, scheme; /* scheme = SW_2D_* */
    float dx, dz, dt;
    float c[SW_MAX_ORDER];
    float *rho, *lam, *mu;
    float *vx, *vz, *sxx, *szz, *sxz;
    sw_axis px, pz; /* py unused in 2-D */
    float *mem;
} sw_ssg2d;

int sw_ssg2d_velocity(const sw_ssg2d *g, void *stream);
int sw_ssg2d_stress(const sw_ssg2d *g, void *stream);
int sw_ssg2d_inject(const sw_ssg2d *g, const sw_source *src, int32_t it, void *stream);
int sw_ssg2d_record(const sw_ssg2d *g, const sw_receivers *rec, int32_t it, void *stream);
/* nsteps reference steps: velocity, source, stress, record.  For SW_2D_RSG the
 * source is the 2x2 spread of src/Elastic2D_CPML.py:322-323 with amp[it]. */
int sw_ssg2d_run(const sw_ssg2d *g, const sw_source *src, const sw_receivers *rec,
                 int32_t it0, int32_t nsteps, void *stream);
/* Free surface (src/Elastic2D_CPML.py:113-125): lam = 0, rho *= 0.5 on planes
 * i < order; and per step sxx = 0 on plane i = order. */
int sw_ssg2d_free_surface_init(const sw_ssg2d *g, void *stream);
int sw_ssg2d_free_surface(const sw_ssg2d *g, void *stream);
/* 3-D variant (src/Elastic3D_SSG_Tensor.py:151-157): lam = 0, rho *= 0.5 on k < order. */
int sw_ssg3d_free_surface_init(const sw_ssg3d *g, float *rho, float *lam, void *stream);
/* 3-D free surface per step (src/Elastic3D_SSG_Tensor.py:159-163, a ti.func the
 * reference never calls): szz = 0 on plane k = order. */
int sw_ssg3d_free_surface(const sw_ssg3d *g, void *stream);

/* ---- 2-D discrete adjoint (no counterpart in the reference; the gradient
 * path of a seismogram misfit with respect to vp, vs, rho) ----
 * Exact transpose of the SW_2D_VS step above, CPML memory variables included.
 * Per backward step n: sw_ssg2d_adjoint_inject (receiver residuals of step n),
 * sw_ssg2d_adjoint_stress with the forward state after the velocity stage of
 * step n, then sw_ssg2d_adjoint_velocity with the forward state before it.
 * The descriptors passed as `fwd` hold that forward state (and the model). */
typedef struct sw_adj2d {
    float *vx, *vz, *sxx, *szz, *sxz; /* adjoint wavefields [nz][nx]               */
    float *mem;                       /* 10 adjoint memory variables, as sw_ssg2d  */
    float *tmp;                       /* 4 scratch arrays [nz][nx]                 */
    float *g_b1, *g_b4;               /* d misfit / d(dt/rho) at vx, vz cells      */
    float *g_lamh, *g_muh, *g_muz;    /* sxx cells: d/d lam_h and d/d mu_h - 2 d/d lam_h;
                                         sxz cells: d/d mu_z                        */
    float *g_src;                     /* force sources only: g_src[b] += amp[it] * the
                                         adjoint of the forced velocity of shot b at
                                         its source cell, summed over the steps (the
                                         force dt*amp/rho[p] then adds
                                         -dt*g_src[b]/rho[p]^2 to d misfit/d rho[p]);
                                         indexed by the shot, not shifted by
                                         adj_stride; may be NULL for stress sources */
} sw_adj2d;

int sw_ssg2d_adjoint_stress(const sw_ssg2d *fwd, const sw_adj2d *adj, void *stream);
int sw_ssg2d_adjoint_velocity(const sw_ssg2d *fwd, const sw_adj2d *adj, void *stream);
/* adds residual[it] (an sw_receivers whose data holds residuals) at the receivers;
 * receivers sharing a cell add up in no fixed order (f32 atomics) */
int sw_ssg2d_adjoint_inject(const sw_ssg2d *g, const sw_adj2d *adj, const sw_receivers *residual,
                            int32_t it, void *stream);
/* gvp, gvs, grho [nz][nx] from the accumulated coefficient gradients */
int sw_ssg2d_model_gradient(const sw_ssg2d *g, const sw_adj2d *adj, const float *vp, const float *vs,
                            float *gvp, float *gvs, float *grho, void *stream);
/* One checkpoint segment of the adjoint, steps it0 .. it0+n-1, in one call.
 * g holds the forward state at step it0 as one block of 15 planes of nx*nz
 * floats starting at g->vx (vx vz sxx szz sxz, then the 10 memory planes at
 * g->mem); the call advances it to step it0+n (source injection, velocity
 * and stress sweeps of src/Elastic2D_CPML.py:127-200 per step), keeping in
 * work ((n*5 + 4)*nx*nz floats, caller-owned; the last 4 planes are scratch)
 * what the model gradient needs of every step: the CPML-corrected derivative sums the two sweeps multiply
 * (vx and vz cells of the velocity sweep, dvx/dx, dvz/dz and dvx/dz + dvz/dx
 * of the stress sweep).  Then, for it = it0+n-1 down to it0, it does what
 * sw_ssg2d_adjoint_inject(residual, it), sw_ssg2d_adjoint_stress (state after
 * the velocity sweep) and sw_ssg2d_adjoint_velocity (state before it) do,
 * with the same results, reading those sums instead of recomputing them from
 * the forward state.  src: explosive or moment
 * tensor (injected before the velocity sweep), a force SW_SRC_FORCE_VZ/VX
 * (after it, as sw_ssg2d_run does; needs adj->g_src), or none. */
int sw_ssg2d_adjoint_segment(const sw_ssg2d *g, const sw_source *src, const sw_adj2d *adj,
                             const sw_receivers *residual, float *work, int32_t it0, int32_t n, void *stream);

/* ---- shot batches (no counterpart in the reference): nbatch shots of the
 * SW_2D_VS scheme in each launch.  Shot b's state is the 15-plane block at
 * g->vx + b*state_stride (layout as for sw_ssg2d_adjoint_segment), its
 * source is src (any kind of sw_ssg2d_adjoint_segment, shared wavelet) at cell
 * (cells[2b], cells[2b+1]), its traces / residuals are rec->data +
 * b*data_stride, and its adjoint block is every sw_adj2d pointer +
 * b*adj_stride (24 planes: 15 state, 4 scratch, 5 accumulators).  The
 * model, PML and receiver positions are shared. */
#define SW_MAX_BATCH 64
/* nsteps of sw_ssg2d_run for every shot */
int sw_ssg2d_run_batch(const sw_ssg2d *g, int32_t nbatch, int64_t state_stride, const sw_source *src,
                       const int32_t *cells, const sw_receivers *rec, int64_t data_stride, int32_t it0,
                       int32_t nsteps, void *stream);
/* sw_ssg2d_adjoint_segment for every shot; work holds (n*5 + 4)*nbatch*nx*nz floats */
int sw_ssg2d_adjoint_segment_batch(const sw_ssg2d *g, int32_t nbatch, int64_t state_stride, const sw_source *src,
                                   const int32_t *cells, const sw_adj2d *adj, int64_t adj_stride,
                                   const sw_receivers *residual, int64_t data_stride, float *work, int32_t it0,
                                   int32_t n, void *stream);
/* sw_ssg2d_model_gradient of the accumulators summed over the shots */
int sw_ssg2d_model_gradient_batch(const sw_ssg2d *g, const sw_adj2d *adj, int32_t nbatch, int64_t adj_stride,
                                  const float *vp, const float *vs, float *gvp, float *gvs, float *grho,
                                  void *stream);

/* ---- learned thin boundary (no counterpart in the reference: a thin CPML
 * plus a per-boundary-column MLP in place of most of a wide CPML) ----
 * The interior box (cells outside every CPML slab) starts at (ox, oy, oz)
 * and spans (ex, ey, ez) cells.  Its SW_BAND_DEPTH outermost cells along
 * each face form the band; a column is one lateral position of one face
 * (x faces span the full lateral box, y faces exclude the x bands, z faces
 * the x and y bands).  Per column: 72 features x (the 9 fields at the 4
 * band cells after the current step, cur, then cur minus the corrected band
 * of the previous step) and 36 outputs (corrections added to cur).  Each of
 * the 6 faces has its own network of sw_mlp_param_count() floats:
 *   W1[72][64] b1[64] W2[64][64] b2[64] W3[64][36] b3[36] in_scale[72] out_scale[36]
 *   u = x/in_scale, s = RMS of the 72 entries of u,
 *   y = (W3^T tanh(W2^T tanh(W1^T (u/s) + b1) + b2) + b3) * out_scale * s  (0 if s = 0)
 * sw_mlp_forward / sw_mlp_loss_backward take samples already divided by s. */
#define SW_BAND_DEPTH 4
#define SW_BAND_FEATURES 72
#define SW_BAND_OUTPUTS 36
#define SW_MLP_HIDDEN 64
typedef struct sw_band {
    int32_t ox, oy, oz, ex, ey, ez;
} sw_band;

int sw_mlp_param_count(void);
int sw_band_columns(const sw_band *b);
/* out[c][36]: current band values of every column */
int sw_band_gather(const sw_ssg3d *g, const sw_band *b, float *out, void *stream);
/* fused inference: fields += MLP correction on every band column; prev[c][36]
 * holds the corrected band of the previous step and is updated in place */
int sw_band_correct(const sw_ssg3d *g, const sw_band *b, const float *params, float *prev, void *stream);
/* nsteps of sw_ssg3d_run with sw_band_correct after each stress sweep; the
 * receivers are recorded after the correction (a receiver on a band cell
 * records the corrected value).  Equal, bit for bit, to calling per step
 * sw_ssg3d_inject, sw_ssg3d_velocity, sw_ssg3d_stress, sw_band_correct and
 * sw_ssg3d_record. */
int sw_ssg3d_run_learned(const sw_ssg3d *g, const sw_source *src, const sw_receivers *rec, const sw_band *b,
                         const float *params, float *prev, int32_t it0, int32_t nsteps, void *stream);
/* training of one face network on a batch X[B][72] -> targets T[B][36] */
int sw_mlp_forward(const float *params, const float *X, int32_t B, float *H1, float *H2, float *Y, void *stream);
/* work: B*(36+128) floats; grad: the trainable part (all but the scales); loss_rows[B] */
int sw_mlp_loss_backward(const float *params, const float *X, const float *H1, const float *H2, const float *Y,
                         const float *T, int32_t B, float *grad, float *work, float *loss_rows, void *stream);
int sw_adam(float *p, const float *g, float *m, float *v, int32_t n, float lr, float beta1, float beta2, float eps,
            int32_t t, void *stream);
/* sw_mlp_loss_backward without the work argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight training
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The dY/dZ2/dZ1 scratch is a static device array of
 * SW_MLP_NOALLOC_MAX_ROWS * 164 floats (10.7 MB) that each device holds until
 * the process ends; its address is looked up by sw_device_init or by the
 * first call on a device.  B above SW_MLP_NOALLOC_MAX_ROWS is refused
 * (SW_EINVAL).  Calls on one device share that scratch: calls that overlap
 * on one GPU (concurrent streams) race; calls in one stream order do not.
 * Nothing in this call allocates. */
#define SW_MLP_NOALLOC_MAX_ROWS 16384
int sw_mlp_loss_backward_noalloc(const float *params, const float *X, const float *H1, const float *H2,
                                 const float *Y, const float *T, int32_t B, float *grad, float *loss_rows,
                                 void *stream);

/* ---- measurement (no counterpart in the reference) ----
 * Streaming probe: out[o][e] = (o+1) * sum_q in[q][e] for e < n, every element
 * read or written once with 16-byte accesses, nblock blocks of 256 threads
 * in a grid-stride loop.  Instantiated mixes (nin read / nout written):
 * 10/3 (the velocity sweep's arrays), 11/6 (the stress sweep's), 16/12 and
 * 20/21 (the two sweeps at a split-field PML shell cell), 1/1 (a copy).  n a
 * multiple of 4, arrays 16-byte aligned and device resident. */
int sw_stream_probe(const float *const *in, int32_t nin, float *const *out, int32_t nout, int64_t n,
                    int32_t nblock, void *stream);
/* Matrix-core probe: nblock blocks of 256 threads, each wave issuing `steps`
 * rounds of 4 independent f32 16x16x4 MFMA products (2048 flop each); out
 * receives one float per thread (nblock * 256 floats, device resident). */
int sw_mfma_probe(float *out, int32_t nblock, int32_t steps, void *stream);

#ifdef __cplusplus
}
#endif
#endif
