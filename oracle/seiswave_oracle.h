/*
 * seiswave_oracle.h -- CPU oracle for the elastic staggered-grid time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Taichi kernels of SeisWave_Taichi (paths relative to the reference root):
 *   src/Differential3D.py, src/Differential2D.py, src/BaseFun.py,
 *   src/Elastic3D_SSG_CPML.py, src/Elastic3D_SSG_Tensor.py,
 *   src/Elastic3D_SSG_SPML.py, src/Elastic2D_CPML.py.
 * It is used by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg
 * of bench.py, as the checker; the product (seiswave_taichi_amd) never links,
 * imports or calls it.
 *
 * The file is compiled twice: with -DOR_REAL=float (liboracle_f32.so, the
 * reference's f32 arithmetic) and -DOR_REAL=double (liboracle_f64.so, the
 * fp64 oracle).  Arrays use the reference layout: (nx,ny,nz) C order, z
 * contiguous, index (i*ny + j)*nz + k; 2-D arrays (nx,nz), index i*nz + j.
 *
 * Arithmetic order follows the reference source term by term (sum taps, then
 * divide by the spacing; multiply by dt then divide by rho; averages summed
 * left to right then scaled).  Build with -ffp-contract=off so that no
 * multiply-add is fused.
 */
#ifndef SEISWAVE_ORACLE_H
#define SEISWAVE_ORACLE_H

#include <stdint.h>

#ifndef OR_REAL
#define OR_REAL float
#endif
typedef OR_REAL real;

#define OR_MAX_ORDER 8

/* One axis of PML coefficients, as SetADEPML2D/3D leaves them
 * (src/Elastic3D_SSG_CPML.py:364-522).  Arrays of length n. */
typedef struct {
    real *d, *d_h;       /* damping d(x) at integer / half points (pml_x, pml_x_half)   */
    real *alpha, *alpha_h;
    real *k, *k_h;
    real *b, *b_h;
    real *a, *a_h;
} or_axis_pml;

typedef struct {
    int32_t nx, ny, nz, s;              /* s = accuracy (half stencil length)      */
    real dx, dy, dz, dt;
    real c[OR_MAX_ORDER];
    real xmin_pml, xmax_pml, ymin_pml, ymax_pml, zmin_pml, zmax_pml;
    or_axis_pml px, py, pz;
    const real *rho, *lam, *mu;
    real *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *sxz, *syz;
    /* CPML memory variables (full grid, as in the reference):
     * 0 sxx_dx 1 sxz_dx 2 sxy_dx 3 syy_dy 4 sxy_dy 5 syz_dy 6 szz_dz 7 sxz_dz 8 syz_dz
     * 9 dvx_dx 10 dvy_dx 11 dvz_dx 12 dvx_dy 13 dvy_dy 14 dvz_dy 15 dvx_dz 16 dvy_dz 17 dvz_dz */
    real *mem[18];
    /* SPML split fields (full grid):
     * 0-2 vx_x vx_y vx_z, 3-5 vy_*, 6-8 vz_*, 9-11 sxx_*, 12-14 syy_*, 15-16 sxy_x sxy_y,
     * 17-18 sxz_x sxz_z, 19-20 syz_y syz_z, 21-23 szz_* */
    real *split[24];
    int32_t isx, isy, isz;              /* source cell                                */
    real mt[6];                         /* Mxx Myy Mzz Mxy Mxz Myz (Tensor class)      */
    int32_t nrec;                       /* receivers: cell indices (i,j,k) triplets    */
    const int32_t *rec;                 /* length 3*nrec                               */
    real *data;                         /* (nt, nrec) in the oracle precision          */
} or_state3d;

typedef struct {
    int32_t nx, nz, s;
    real dx, dz, dt;
    real c[OR_MAX_ORDER];
    real xmin_pml, xmax_pml, zmin_pml, zmax_pml;
    or_axis_pml px, pz;
    real *rho, *lam, *mu;               /* rho, lam writable: AEA_init mutates them   */
    real *vx, *vz, *sxx, *szz, *sxz;
    /* 0 sxx_dx 1 sxx_dz 2 sxz_dx 3 sxz_dz 4 szz_dx 5 szz_dz 6 dvx_dx 7 dvx_dz 8 dvz_dx 9 dvz_dz */
    real *mem[10];
    int32_t isx, isz, rsx, rsz;
    real *data;                         /* (nt,) in the field dtype                    */
} or_state2d;

/* src/Elastic3D_SSG_CPML.py:311-322 -- staggered Taylor coefficients in f64. */
int or_diff_coff(int32_t order, double *c_out);

/* src/BaseFun.py:4-11 and :13-19 (f32 in the f32 build). */
real or_ricker(int32_t nt, real dt, real freq, real scale);
real or_ricker2(int32_t nt, real dt, real t0, real freq, real scale);

/* src/Elastic3D_SSG_CPML.py:364-522 for one axis.  xmin is 0; xmax_coord is
 * d*(n-1-2s).  Writes the ten arrays and returns min_pml / max_pml. */
void or_pml_axis(int32_t n, int32_t s, double d, double dt, double thick, double vp_max,
                 double rcoef, double alpha_max, double kmax, double theta,
                 int32_t use_lo, int32_t use_hi, or_axis_pml *out,
                 double *min_pml, double *max_pml);

/* 3-D step bodies. */
void or3d_cpml_velocity(or_state3d *st, int32_t tensor_order); /* :164-212 / Tensor :195-243 */
void or3d_cpml_stress(or_state3d *st);                           /* :247-307 / Tensor :246-310 */
void or3d_spml_velocity(or_state3d *st);                         /* SPML :182-259 */
void or3d_spml_stress(or_state3d *st);                           /* SPML :296-391 */

/* Full reference steps (update_SSG / update_SSG_SPML) with the wavelet
 * amplitude for this step supplied by the caller. */
void or3d_cpml_step(or_state3d *st, int32_t it, real amp);    /* src/Elastic3D_SSG_CPML.py:144-309 */
void or3d_tensor_step(or_state3d *st, int32_t it, real amp);  /* src/Elastic3D_SSG_Tensor.py:166-345 */
void or3d_spml_step(or_state3d *st, int32_t it, real amp);    /* src/Elastic3D_SSG_SPML.py:160-393 */

/* src/Elastic3D_SSG_Tensor.py:151-157 (rho_w, lam_w alias st->rho, st->lam). */
void or3d_aea_init(or_state3d *st, real *rho_w, real *lam_w);

/* 2-D (src/Elastic2D_CPML.py). */
void or2d_vs_step(or_state2d *st, int32_t it, real amp);      /* :127-200 */
void or2d_sv_step(or_state2d *st, int32_t it, real amp);      /* :202-274 */
void or2d_rsg_step(or_state2d *st, real source);              /* :276-353 */
void or2d_aea_init(or_state2d *st);                           /* :113-119 */
void or2d_aea(or_state2d *st);                                /* :121-125 */

/* Sets the OpenMP thread count (n > 0) and returns the count in effect. */
int or_set_threads(int32_t n);

#endif
