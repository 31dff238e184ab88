/*
 * seiswave_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see the header).
 *
 * Each function restates one reference kernel; the reference file:line is
 * given at the function.  OpenMP parallelises the outermost index of every
 * loop, one parallel region per reference loop (a Taichi kernel runs its
 * top-level loops as ordered tasks with a barrier between them).
 */
#include "seiswave_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>

#define I3(i, j, k) (((size_t)(i) * (size_t)st->ny + (size_t)(j)) * (size_t)st->nz + (size_t)(k))
#define I2(i, j) ((size_t)(i) * (size_t)st->nz + (size_t)(j))

/* pi**2 is evaluated by Python in f64 before it reaches the kernel
 * (src/BaseFun.py:7,15); the kernel sees it as a constant of the field type. */
#define OR_PI2 9.869604401089358

/* Grids smaller than this run their loops on one thread (thread start-up and
 * barrier costs dominate below it). */
#define OR_PAR_MIN 200000L

/* ---------------------------------------------------------------- stencils
 * src/Differential3D.py:7-44 / src/Differential2D.py:4-28.  Forward ("fm"):
 * sum_id c[id]*(f[i+id+1]-f[i-id]); backward ("bm"): c[id]*(f[i+id]-f[i-id-1]).
 * Accumulator starts at 0 and adds in id order. */
static inline real d_fm(const real *f, size_t idx, size_t stride, const real *c, int n)
{
    real v = (real)0;
    for (int id = 0; id < n; ++id)
        v += c[id] * (f[idx + (size_t)(id + 1) * stride] - f[idx - (size_t)id * stride]);
    return v;
}

static inline real d_bm(const real *f, size_t idx, size_t stride, const real *c, int n)
{
    real v = (real)0;
    for (int id = 0; id < n; ++id)
        v += c[id] * (f[idx + (size_t)id * stride] - f[idx - (size_t)(id + 1) * stride]);
    return v;
}

/* Rotated stencils, src/Differential2D.py:32-56 (2-D, index i*nz + j). */
static inline real drx_fm(const or_state2d *st, const real *f, int i, int j)
{
    real v = (real)0;
    for (int id = 0; id < st->s; ++id)
        v += st->c[id] * (f[I2(i + id + 1, j + id + 1)] - f[I2(i - id, j - id)]);
    return v;
}
static inline real drz_fm(const or_state2d *st, const real *f, int i, int j)
{
    real v = (real)0;
    for (int id = 0; id < st->s; ++id)
        v += st->c[id] * (f[I2(i - id, j + id + 1)] - f[I2(i + id + 1, j - id)]);
    return v;
}
static inline real drx_bm(const or_state2d *st, const real *f, int i, int j)
{
    real v = (real)0;
    for (int id = 0; id < st->s; ++id)
        v += st->c[id] * (f[I2(i + id, j + id)] - f[I2(i - id - 1, j - id - 1)]);
    return v;
}
static inline real drz_bm(const or_state2d *st, const real *f, int i, int j)
{
    real v = (real)0;
    for (int id = 0; id < st->s; ++id)
        v += st->c[id] * (f[I2(i - id - 1, j + id)] - f[I2(i + id, j - id - 1)]);
    return v;
}

/* ------------------------------------------------------------ coefficients
 * src/Elastic3D_SSG_CPML.py:311-322: solve A c = e0 with A[i][j] = (2j+1)^(2i+1)
 * (Gaussian elimination with partial pivoting, f64). */
int or_diff_coff(int32_t order, double *c_out)
{
    if (order < 1 || order > OR_MAX_ORDER) return -1;
    double A[OR_MAX_ORDER][OR_MAX_ORDER + 1];
    for (int i = 0; i < order; ++i) {
        for (int j = 0; j < order; ++j) A[i][j] = pow((double)(2 * j + 1), (double)(2 * i + 1));
        A[i][order] = (i == 0) ? 1.0 : 0.0;
    }
    for (int col = 0; col < order; ++col) {
        int piv = col;
        for (int r = col + 1; r < order; ++r)
            if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
        if (piv != col)
            for (int q = 0; q <= order; ++q) {
                double t = A[col][q]; A[col][q] = A[piv][q]; A[piv][q] = t;
            }
        for (int r = col + 1; r < order; ++r) {
            double f = A[r][col] / A[col][col];
            for (int q = col; q <= order; ++q) A[r][q] -= f * A[col][q];
        }
    }
    for (int r = order - 1; r >= 0; --r) {
        double acc = A[r][order];
        for (int q = r + 1; q < order; ++q) acc -= A[r][q] * c_out[q];
        c_out[r] = acc / A[r][r];
    }
    return 0;
}

/* ---------------------------------------------------------------- wavelets */
#define OR_EXP(x) (sizeof(real) == sizeof(float) ? (real)expf((float)(x)) : (real)exp((double)(x)))

/* src/BaseFun.py:4-11: t0 = 1.5/f, a = pi^2 f^2, -scale*2*a*tau*exp(-a*tau^2). */
real or_ricker(int32_t nt, real dt, real freq, real scale)
{
    real t0 = (real)1.5 / freq;
    real a = (real)OR_PI2 * (freq * freq);
    real t = (real)nt * dt;
    real tau = t - t0;
    real e = OR_EXP((-a) * (tau * tau));
    return (((-scale) * (real)2) * a) * tau * e;
}

/* src/BaseFun.py:13-19: scale*(1 - 2*a*tau^2)*exp(-a*tau^2). */
real or_ricker2(int32_t nt, real dt, real t0, real freq, real scale)
{
    real a = (real)OR_PI2 * (freq * freq);
    real t = (real)nt * dt;
    real tau = t - t0;
    real tt = tau * tau;
    real e = OR_EXP((-a) * tt);
    return (scale * ((real)1 - ((real)2 * a) * tt)) * e;
}

/* --------------------------------------------------------------- PML setup
 * src/Elastic3D_SSG_CPML.py:324-522 (one axis; identical in the 2-D class,
 * src/Elastic2D_CPML.py:392-528).  Profiles are computed in f64 and stored in
 * the field type; b and a are then computed in f64 from the stored values. */
static void pml_profile(int32_t n, int32_t s, double d, double half, double min_pml, double max_pml,
                        double thick, double d0, double alpha_max, double kmax,
                        int32_t use_lo, int32_t use_hi, real *D, real *AL, real *K)
{
    for (int i = 0; i < n; ++i) {
        double x = (double)(i - s) * d + half;
        double dd, al, kk;
        if (x <= min_pml && use_lo) {
            double r = (min_pml - x) / thick;
            dd = d0 * (r * r);
            al = alpha_max * (1.0 - r);
            kk = 1.0 + (kmax - 1.0) * (r * r);
        } else if (x >= max_pml && use_hi) {
            double r = (x - max_pml) / thick;
            dd = d0 * (r * r);
            al = alpha_max * (1.0 - r);
            kk = 1.0 + (kmax - 1.0) * (r * r);
        } else {
            dd = 0.0; al = 0.0; kk = 1.0;
        }
        D[i] = (real)dd;
        AL[i] = (real)al;
        K[i] = (real)kk;
        if (AL[i] < 0) AL[i] = 0;
    }
}

void or_pml_axis(int32_t n, int32_t s, double d, double dt, double thick, double vp_max,
                 double rcoef, double alpha_max, double kmax, double theta,
                 int32_t use_lo, int32_t use_hi, or_axis_pml *out,
                 double *min_pml, double *max_pml)
{
    double xmax = d * (double)(n - 1 - 2 * s);
    double mn = 0.0 + thick;
    double mx = xmax - thick;
    double d0 = -3.0 * vp_max * log(rcoef) / (2.0 * thick);
    pml_profile(n, s, d, 0.0, mn, mx, thick, d0, alpha_max, kmax, use_lo, use_hi,
                out->d, out->alpha, out->k);
    pml_profile(n, s, d, d / 2.0, mn, mx, thick, d0, alpha_max, kmax, use_lo, use_hi,
                out->d_h, out->alpha_h, out->k_h);
    for (int i = 0; i < n; ++i) {
        double D = out->d[i], K = out->k[i], A = out->alpha[i];
        out->b[i] = (real)((1.0 - (1.0 - theta) * dt * (D / K + A)) / (1.0 + theta * dt * (D / K + A)));
        out->a[i] = (real)0;
        if (fabs(D) > 1e-6) out->a[i] = (real)((-dt) * D / (K * K) / (1.0 + theta * dt * (D / K + A)));
        D = out->d_h[i]; K = out->k_h[i]; A = out->alpha_h[i];
        out->b_h[i] = (real)((1.0 - (1.0 - theta) * dt * (D / K + A)) / (1.0 + theta * dt * (D / K + A)));
        out->a_h[i] = (real)0;
        if (fabs(D) > 1e-6) out->a_h[i] = (real)((-dt) * D / (K * K) / (1.0 + theta * dt * (D / K + A)));
    }
    *min_pml = mn;
    *max_pml = mx;
}

/* ------------------------------------------------------------- 3-D helpers */
static inline int in_pml3(const or_state3d *st, real x, real y, real z)
{
    return x <= st->xmin_pml || x >= st->xmax_pml || y <= st->ymin_pml || y >= st->ymax_pml ||
           z <= st->zmin_pml || z >= st->zmax_pml;
}

/* ------------------------------------------------------ 3-D CPML velocity
 * src/Elastic3D_SSG_CPML.py:164-212 (tensor_order=0) and
 * src/Elastic3D_SSG_Tensor.py:195-243 (tensor_order=1).  The two differ only
 * in the order of the three terms of the vz sum (:212 vs Tensor :243). */
void or3d_cpml_velocity(or_state3d *st, int32_t tensor_order)
{
    const int s = st->s, nx = st->nx, ny = st->ny, nz = st->nz;
    const size_t SX = (size_t)ny * nz, SY = (size_t)nz;
    const real dx = st->dx, dy = st->dy, dz = st->dz, dt = st->dt;
    const real hdx = dx / 2, hdy = dy / 2, hdz = dz / 2;
    const real *c = st->c;
    real **m = st->mem;

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s + 1; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx + hdx, y = (real)(j - s) * dy, z = (real)(k - s) * dz;
                real dsxx = d_fm(st->sxx, p, SX, c, s) / dx;
                real dsxyy = d_bm(st->sxy, p, SY, c, s) / dy;
                real dsxzz = d_bm(st->sxz, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    m[0][p] = st->px.b_h[i] * m[0][p] + st->px.a_h[i] * dsxx;
                    m[4][p] = st->py.b[j] * m[4][p] + st->py.a[j] * dsxyy;
                    m[7][p] = st->pz.b[k] * m[7][p] + st->pz.a[k] * dsxzz;
                    dsxx = dsxx / st->px.k_h[i] + m[0][p];
                    dsxyy = dsxyy / st->py.k[j] + m[4][p];
                    dsxzz = dsxzz / st->pz.k[k] + m[7][p];
                }
                st->vx[p] += (dsxx + dsxzz + dsxyy) * dt / st->rho[p];
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s; j < ny - s - 1; ++j)
            for (int k = s + 1; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy + hdy, z = (real)(k - s) * dz;
                real rho_half_y = (real)0.5 * (st->rho[p] + st->rho[p + SY]);
                real dsxy = d_bm(st->sxy, p, SX, c, s) / dx;
                real dsyy = d_fm(st->syy, p, SY, c, s) / dy;
                real dsyz = d_bm(st->syz, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    m[2][p] = st->px.b[i] * m[2][p] + st->px.a[i] * dsxy;
                    m[3][p] = st->py.b_h[j] * m[3][p] + st->py.a_h[j] * dsyy;
                    m[8][p] = st->pz.b[k] * m[8][p] + st->pz.a[k] * dsyz;
                    dsxy = dsxy / st->px.k[i] + m[2][p];
                    dsyy = dsyy / st->py.k_h[j] + m[3][p];
                    dsyz = dsyz / st->pz.k[k] + m[8][p];
                }
                st->vy[p] += (dsxy + dsyy + dsyz) * dt / rho_half_y;
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s; k < nz - s - 1; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy, z = (real)(k - s) * dz + hdz;
                real rho_avg = (real)0.25 * (st->rho[p] + st->rho[p + SY] + st->rho[p + SX + SY] + st->rho[p + SX]);
                real dsxz = d_bm(st->sxz, p, SX, c, s) / dx;
                real dsyz = d_bm(st->syz, p, SY, c, s) / dy;
                real dszz = d_fm(st->szz, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    m[1][p] = st->px.b[i] * m[1][p] + st->px.a[i] * dsxz;
                    m[5][p] = st->py.b[j] * m[5][p] + st->py.a[j] * dsyz;
                    m[6][p] = st->pz.b_h[k] * m[6][p] + st->pz.a_h[k] * dszz;
                    dsxz = dsxz / st->px.k[i] + m[1][p];
                    dsyz = dsyz / st->py.k[j] + m[5][p];
                    dszz = dszz / st->pz.k_h[k] + m[6][p];
                }
                if (tensor_order)
                    st->vz[p] += (dsxz + dsyz + dszz) * dt / rho_avg;
                else
                    st->vz[p] += (dszz + dsxz + dsyz) * dt / rho_avg;
            }
}

/* -------------------------------------------------------- 3-D CPML stress
 * src/Elastic3D_SSG_CPML.py:247-307 (identical arithmetic in
 * src/Elastic3D_SSG_Tensor.py:246-310). */
void or3d_cpml_stress(or_state3d *st)
{
    const int s = st->s, nx = st->nx, ny = st->ny, nz = st->nz;
    const size_t SX = (size_t)ny * nz, SY = (size_t)nz;
    const real dx = st->dx, dy = st->dy, dz = st->dz, dt = st->dt;
    const real hdx = dx / 2, hdy = dy / 2, hdz = dz / 2;
    const real *c = st->c;
    real **m = st->mem;

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s + 1; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy, z = (real)(k - s) * dz;
                real lam_h = (real)0.5 * (st->lam[p] + st->lam[p + SX]);
                real mu_h = (real)0.5 * (st->mu[p] + st->mu[p + SX]);
                real lpm = lam_h + (real)2 * mu_h;
                real dvxx = d_bm(st->vx, p, SX, c, s) / dx;
                real dvyy = d_bm(st->vy, p, SY, c, s) / dy;
                real dvzz = d_bm(st->vz, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    m[9][p] = st->px.b[i] * m[9][p] + st->px.a[i] * dvxx;
                    m[13][p] = st->py.b[j] * m[13][p] + st->py.a[j] * dvyy;
                    m[17][p] = st->pz.b[k] * m[17][p] + st->pz.a[k] * dvzz;
                    dvxx = dvxx / st->px.k[i] + m[9][p];
                    dvyy = dvyy / st->py.k[j] + m[13][p];
                    dvzz = dvzz / st->pz.k[k] + m[17][p];
                }
                st->sxx[p] += (lpm * dvxx + lam_h * (dvzz + dvyy)) * dt;
                st->syy[p] += (lpm * dvyy + lam_h * (dvxx + dvzz)) * dt;
                st->szz[p] += (lpm * dvzz + lam_h * (dvxx + dvyy)) * dt;
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx + hdx, y = (real)(j - s) * dy + hdy, z = (real)(k - s) * dz;
                real mu_a = (real)0.25 * (st->mu[p + SY] + st->mu[p] + st->mu[p] + st->mu[p + SX]);
                real dvyx = d_fm(st->vy, p, SX, c, s) / dx;
                real dvxy = d_fm(st->vx, p, SY, c, s) / dy;
                if (in_pml3(st, x, y, z)) {
                    m[10][p] = st->px.b_h[i] * m[10][p] + st->px.a_h[i] * dvyx;
                    m[12][p] = st->py.b_h[j] * m[12][p] + st->py.a_h[j] * dvxy;
                    dvyx = dvyx / st->px.k_h[i] + m[10][p];
                    dvxy = dvxy / st->py.k_h[j] + m[12][p];
                }
                st->sxy[p] += mu_a * (dvxy + dvyx) * dt;
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s; j < ny - s; ++j)
            for (int k = s; k < nz - s - 1; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx + hdx, y = (real)(j - s) * dy, z = (real)(k - s) * dz + hdz;
                real mu_a = (real)0.25 * (st->mu[p] + st->mu[p + SX] + st->mu[p] + st->mu[p + SY]);
                real dvzx = d_fm(st->vz, p, SX, c, s) / dx;
                real dvxz = d_fm(st->vx, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    m[11][p] = st->px.b_h[i] * m[11][p] + st->px.a_h[i] * dvzx;
                    m[15][p] = st->pz.b_h[k] * m[15][p] + st->pz.a_h[k] * dvxz;
                    dvzx = dvzx / st->px.k_h[i] + m[11][p];
                    dvxz = dvxz / st->pz.k_h[k] + m[15][p];
                }
                st->sxz[p] += mu_a * (dvxz + dvzx) * dt;
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s; ++i)
        for (int j = s; j < ny - s - 1; ++j)
            for (int k = s; k < nz - s - 1; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy + hdy, z = (real)(k - s) * dz + hdz;
                real mu_a = (real)0.25 * (st->mu[p + SY] + st->mu[p] + st->mu[p + 1] + st->mu[p]);
                real dvyz = d_fm(st->vy, p, 1, c, s) / dz;
                real dvzy = d_fm(st->vz, p, SY, c, s) / dy;
                if (in_pml3(st, x, y, z)) {
                    m[14][p] = st->py.b_h[j] * m[14][p] + st->py.a_h[j] * dvzy;
                    m[16][p] = st->pz.b_h[k] * m[16][p] + st->pz.a_h[k] * dvyz;
                    dvzy = dvzy / st->py.k_h[j] + m[14][p];
                    dvyz = dvyz / st->pz.k_h[k] + m[16][p];
                }
                st->syz[p] += mu_a * (dvyz + dvzy) * dt;
            }
}

/* ------------------------------------------------------ 3-D SPML velocity
 * src/Elastic3D_SSG_SPML.py:182-259.  Inside the PML region each component is
 * split per derivative direction and advanced with a Crank-Nicolson step on d
 * only; the total is the sum of the splits. */
void or3d_spml_velocity(or_state3d *st)
{
    const int s = st->s, nx = st->nx, ny = st->ny, nz = st->nz;
    const size_t SX = (size_t)ny * nz, SY = (size_t)nz;
    const real dx = st->dx, dy = st->dy, dz = st->dz, dt = st->dt;
    const real hdx = dx / 2, hdy = dy / 2, hdz = dz / 2, hdt = (real)0.5 * dt;
    const real *c = st->c;
    real **q = st->split;

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s + 1; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx + hdx, y = (real)(j - s) * dy, z = (real)(k - s) * dz;
                real dsxxdx = d_fm(st->sxx, p, SX, c, s) / dx;
                real dsxydy = d_bm(st->sxy, p, SY, c, s) / dy;
                real dsxzdz = d_bm(st->sxz, p, 1, c, s) / dz;
                real rho = st->rho[p];
                real rho_over = (real)1 / rho;
                if (in_pml3(st, x, y, z)) {
                    real pxn = (real)1 - hdt * st->px.d_h[i], pyn = (real)1 - hdt * st->py.d[j], pzn = (real)1 - hdt * st->pz.d[k];
                    real pxd = (real)1 + hdt * st->px.d_h[i], pyd = (real)1 + hdt * st->py.d[j], pzd = (real)1 + hdt * st->pz.d[k];
                    q[0][p] = (pxn * q[0][p] + dsxxdx * dt * rho_over) / pxd;
                    q[1][p] = (pyn * q[1][p] + dsxydy * dt * rho_over) / pyd;
                    q[2][p] = (pzn * q[2][p] + dsxzdz * dt * rho_over) / pzd;
                    st->vx[p] = q[0][p] + q[1][p] + q[2][p];
                } else {
                    st->vx[p] += (dsxxdx + dsxydy + dsxzdz) * dt * rho_over;
                }
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s; j < ny - s - 1; ++j)
            for (int k = s + 1; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy + hdy, z = (real)(k - s) * dz;
                real rho_half_y = (real)0.5 * (st->rho[p] + st->rho[p + SY]);
                real rho_over = (real)1 / rho_half_y;
                real dsxydx = d_bm(st->sxy, p, SX, c, s) / dx;
                real dsyydy = d_fm(st->syy, p, SY, c, s) / dy;
                real dsyzdz = d_bm(st->syz, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    real pxn = (real)1 - hdt * st->px.d[i], pyn = (real)1 - hdt * st->py.d_h[j], pzn = (real)1 - hdt * st->pz.d[k];
                    real pxd = (real)1 + hdt * st->px.d[i], pyd = (real)1 + hdt * st->py.d_h[j], pzd = (real)1 + hdt * st->pz.d[k];
                    q[3][p] = (pxn * q[3][p] + dt * dsxydx * rho_over) / pxd;
                    q[4][p] = (pyn * q[4][p] + dt * dsyydy * rho_over) / pyd;
                    q[5][p] = (pzn * q[5][p] + dt * dsyzdz * rho_over) / pzd;
                    st->vy[p] = q[3][p] + q[4][p] + q[5][p];
                } else {
                    st->vy[p] += (dsxydx + dsyydy + dsyzdz) * dt * rho_over;
                }
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s; k < nz - s - 1; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy, z = (real)(k - s) * dz + hdz;
                real rho_avg = (real)0.25 * (st->rho[p] + st->rho[p + SY] + st->rho[p + SX + SY] + st->rho[p + SX]);
                real rho_over = (real)1 / rho_avg;
                real dsxzdx = d_bm(st->sxz, p, SX, c, s) / dx;
                real dsyzdy = d_bm(st->syz, p, SY, c, s) / dy;
                real dszzdz = d_fm(st->szz, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    real pxn = (real)1 - hdt * st->px.d[i], pyn = (real)1 - hdt * st->py.d[j], pzn = (real)1 - hdt * st->pz.d_h[k];
                    real pxd = (real)1 + hdt * st->px.d[i], pyd = (real)1 + hdt * st->py.d[j], pzd = (real)1 + hdt * st->pz.d_h[k];
                    q[6][p] = (pxn * q[6][p] + dt * dsxzdx * rho_over) / pxd;
                    q[7][p] = (pyn * q[7][p] + dt * dsyzdy * rho_over) / pyd;
                    q[8][p] = (pzn * q[8][p] + dt * dszzdz * rho_over) / pzd;
                    st->vz[p] = q[6][p] + q[7][p] + q[8][p];
                } else {
                    st->vz[p] += (dsxzdx + dsyzdy + dszzdz) * dt * rho_over;
                }
            }
}

/* -------------------------------------------------------- 3-D SPML stress
 * src/Elastic3D_SSG_SPML.py:296-391 (point lambda, mu for the normal stresses). */
void or3d_spml_stress(or_state3d *st)
{
    const int s = st->s, nx = st->nx, ny = st->ny, nz = st->nz;
    const size_t SX = (size_t)ny * nz, SY = (size_t)nz;
    const real dx = st->dx, dy = st->dy, dz = st->dz, dt = st->dt;
    const real hdx = dx / 2, hdy = dy / 2, hdz = dz / 2, hdt = (real)0.5 * dt;
    const real *c = st->c;
    real **q = st->split;

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s + 1; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy, z = (real)(k - s) * dz;
                real lam = st->lam[p], mu = st->mu[p];
                real l2m = lam + (real)2 * mu;
                real dvxdx = d_bm(st->vx, p, SX, c, s) / dx;
                real dvydy = d_bm(st->vy, p, SY, c, s) / dy;
                real dvzdz = d_bm(st->vz, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    real pxn = (real)1 - hdt * st->px.d[i], pyn = (real)1 - hdt * st->py.d[j], pzn = (real)1 - hdt * st->pz.d[k];
                    real pxd = (real)1 + hdt * st->px.d[i], pyd = (real)1 + hdt * st->py.d[j], pzd = (real)1 + hdt * st->pz.d[k];
                    q[9][p] = (pxn * q[9][p] + dt * l2m * dvxdx) / pxd;
                    q[10][p] = (pyn * q[10][p] + dt * lam * dvydy) / pyd;
                    q[11][p] = (pzn * q[11][p] + dt * lam * dvzdz) / pzd;
                    q[12][p] = (pxn * q[12][p] + dt * lam * dvxdx) / pxd;
                    q[13][p] = (pyn * q[13][p] + dt * l2m * dvydy) / pyd;
                    q[14][p] = (pzn * q[14][p] + dt * lam * dvzdz) / pzd;
                    q[21][p] = (pxn * q[21][p] + dt * lam * dvxdx) / pxd;
                    q[22][p] = (pyn * q[22][p] + dt * lam * dvydy) / pyd;
                    q[23][p] = (pzn * q[23][p] + dt * l2m * dvzdz) / pzd;
                    st->sxx[p] = q[9][p] + q[10][p] + q[11][p];
                    st->syy[p] = q[12][p] + q[13][p] + q[14][p];
                    st->szz[p] = q[21][p] + q[22][p] + q[23][p];
                } else {
                    st->sxx[p] += (l2m * dvxdx + lam * (dvzdz + dvydy)) * dt;
                    st->syy[p] += (l2m * dvydy + lam * (dvxdx + dvzdz)) * dt;
                    st->szz[p] += (l2m * dvzdz + lam * (dvxdx + dvydy)) * dt;
                }
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < ny - s; ++j)
            for (int k = s; k < nz - s; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx + hdx, y = (real)(j - s) * dy + hdy, z = (real)(k - s) * dz;
                real mu_a = (real)0.25 * (st->mu[p + SY] + st->mu[p] + st->mu[p] + st->mu[p + SX]);
                real dvydx = d_fm(st->vy, p, SX, c, s) / dx;
                real dvxdy = d_fm(st->vx, p, SY, c, s) / dy;
                if (in_pml3(st, x, y, z)) {
                    real pxn = (real)1 - hdt * st->px.d_h[i], pyn = (real)1 - hdt * st->py.d_h[j];
                    real pxd = (real)1 + hdt * st->px.d_h[i], pyd = (real)1 + hdt * st->py.d_h[j];
                    q[15][p] = (pxn * q[15][p] + dt * mu_a * dvydx) / pxd;
                    q[16][p] = (pyn * q[16][p] + dt * mu_a * dvxdy) / pyd;
                    st->sxy[p] = q[15][p] + q[16][p];
                } else {
                    st->sxy[p] += mu_a * (dvxdy + dvydx) * dt;
                }
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s; j < ny - s; ++j)
            for (int k = s; k < nz - s - 1; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx + hdx, y = (real)(j - s) * dy, z = (real)(k - s) * dz + hdz;
                real mu_a = (real)0.25 * (st->mu[p] + st->mu[p + SX] + st->mu[p] + st->mu[p + SY]);
                real dvzdx = d_fm(st->vz, p, SX, c, s) / dx;
                real dvxdz = d_fm(st->vx, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    real pxn = (real)1 - hdt * st->px.d_h[i], pzn = (real)1 - hdt * st->pz.d_h[k];
                    real pxd = (real)1 + hdt * st->px.d_h[i], pzd = (real)1 + hdt * st->pz.d_h[k];
                    q[17][p] = (pxn * q[17][p] + dt * mu_a * dvzdx) / pxd;
                    q[18][p] = (pzn * q[18][p] + dt * mu_a * dvxdz) / pzd;
                    st->sxz[p] = q[17][p] + q[18][p];
                } else {
                    st->sxz[p] += mu_a * (dvxdz + dvzdx) * dt;
                }
            }

#pragma omp parallel for schedule(static) if ((long)nx * ny * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s; ++i)
        for (int j = s; j < ny - s - 1; ++j)
            for (int k = s; k < nz - s - 1; ++k) {
                size_t p = I3(i, j, k);
                real x = (real)(i - s) * dx, y = (real)(j - s) * dy + hdy, z = (real)(k - s) * dz + hdz;
                real mu_a = (real)0.25 * (st->mu[p + SY] + st->mu[p] + st->mu[p + 1] + st->mu[p]);
                real dvzdy = d_fm(st->vz, p, SY, c, s) / dy;
                real dvydz = d_fm(st->vy, p, 1, c, s) / dz;
                if (in_pml3(st, x, y, z)) {
                    real pyn = (real)1 - hdt * st->py.d_h[j], pzn = (real)1 - hdt * st->pz.d_h[k];
                    real pyd = (real)1 + hdt * st->py.d_h[j], pzd = (real)1 + hdt * st->pz.d_h[k];
                    q[19][p] = (pyn * q[19][p] + dt * mu_a * dvzdy) / pyd;
                    q[20][p] = (pzn * q[20][p] + dt * mu_a * dvydz) / pzd;
                    st->syz[p] = q[19][p] + q[20][p];
                } else {
                    st->syz[p] += mu_a * (dvydz + dvzdy) * dt;
                }
            }
}

/* ------------------------------------------------------------ 3-D steps */
static void record_vx(or_state3d *st, int32_t it)
{
    for (int r = 0; r < st->nrec; ++r) {
        const int32_t *q = st->rec + 3 * r;
        st->data[(size_t)it * st->nrec + r] = st->vx[I3(q[0], q[1], q[2])];
    }
}

/* src/Elastic3D_SSG_CPML.py:144-309.  The six Dirichlet loops (:213-243) are
 * omitted: they only write zeros to cells no loop ever updates. */
void or3d_cpml_step(or_state3d *st, int32_t it, real amp)
{
    or3d_cpml_velocity(st, 0);
    size_t p = I3(st->isx + 5, st->isy + 5, st->isz + 5);
    st->vz[p] += st->dt * amp / st->rho[p];
    or3d_cpml_stress(st);
    record_vx(st, it);
}

/* src/Elastic3D_SSG_Tensor.py:166-345: moment-tensor injection first. */
void or3d_tensor_step(or_state3d *st, int32_t it, real amp)
{
    size_t p = I3(st->isx, st->isy, st->isz);
    st->sxx[p] += amp * (-st->mt[0]);
    st->syy[p] += amp * (-st->mt[1]);
    st->szz[p] += amp * (-st->mt[2]);
    st->sxy[p] += amp * (-st->mt[3]);
    st->sxz[p] += amp * (-st->mt[4]);
    st->syz[p] += amp * (-st->mt[5]);
    or3d_cpml_velocity(st, 1);
    or3d_cpml_stress(st);
    record_vx(st, it);
}

/* src/Elastic3D_SSG_SPML.py:160-393. */
void or3d_spml_step(or_state3d *st, int32_t it, real amp)
{
    or3d_spml_velocity(st);
    size_t p = I3(st->isx, st->isy, st->isz);
    st->vz[p] += st->dt * amp / st->rho[p];
    or3d_spml_stress(st);
    record_vx(st, it);
}

/* ------------------------------------------------------------------- 2-D */
static inline int in_pml2(const or_state2d *st, real x, real z)
{
    return x <= st->xmin_pml || x >= st->xmax_pml || z <= st->zmin_pml || z >= st->zmax_pml;
}

/* src/Elastic2D_CPML.py:127-200 (update_SSG_VS). */
void or2d_vs_step(or_state2d *st, int32_t it, real amp)
{
    const int s = st->s, nx = st->nx, nz = st->nz;
    const size_t SX = (size_t)nz;
    const real dx = st->dx, dz = st->dz, dt = st->dt, hdx = dx / 2, hdz = dz / 2;
    const real *c = st->c;
    real **m = st->mem;

#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < nz - s; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx, z = (real)(j - s) * dz;
            real dsxxdx = d_bm(st->sxx, p, SX, c, s) / dx;
            real dsxzdz = d_bm(st->sxz, p, 1, c, s) / dz;
            if (in_pml2(st, x, z)) {
                m[0][p] = st->px.b[i] * m[0][p] + st->px.a[i] * dsxxdx;
                m[3][p] = st->pz.b[j] * m[3][p] + st->pz.a[j] * dsxzdz;
                dsxxdx = dsxxdx / st->px.k[i] + m[0][p];
                dsxzdz = dsxzdz / st->pz.k[j] + m[3][p];
            }
            st->vx[p] += (dsxxdx + dsxzdz) * dt / st->rho[p];
        }
#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s; j < nz - s - 1; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx + hdx, z = (real)(j - s) * dz + hdz;
            real rho_avg = (real)0.25 * (st->rho[p] + st->rho[p + 1] + st->rho[p + SX + 1] + st->rho[p + SX]);
            real dsxzdx = d_fm(st->sxz, p, SX, c, s) / dx;
            real dszzdz = d_fm(st->szz, p, 1, c, s) / dz;
            if (in_pml2(st, x, z)) {
                m[2][p] = st->px.b_h[i] * m[2][p] + st->px.a_h[i] * dsxzdx;
                m[5][p] = st->pz.b_h[j] * m[5][p] + st->pz.a_h[j] * dszzdz;
                dsxzdx = dsxzdx / st->px.k_h[i] + m[2][p];
                dszzdz = dszzdz / st->pz.k_h[j] + m[5][p];
            }
            st->vz[p] += (dszzdz + dsxzdx) * dt / rho_avg;
        }
    {
        size_t p = I2(st->isx, st->isz);
        st->vz[p] += dt * amp / st->rho[p];
    }
#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s + 1; j < nz - s; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx + hdx, z = (real)(j - s) * dz;
            real lam_h = (real)0.5 * (st->lam[p] + st->lam[p + SX]);
            real mu_h = (real)0.5 * (st->mu[p] + st->mu[p + SX]);
            real lpm = lam_h + (real)2 * mu_h;
            real dvxdx = d_fm(st->vx, p, SX, c, s) / dx;
            real dvzdz = d_bm(st->vz, p, 1, c, s) / dz;
            if (in_pml2(st, x, z)) {
                m[6][p] = st->px.b_h[i] * m[6][p] + st->px.a_h[i] * dvxdx;
                m[9][p] = st->pz.b[j] * m[9][p] + st->pz.a[j] * dvzdz;
                dvxdx = dvxdx / st->px.k_h[i] + m[6][p];
                dvzdz = dvzdz / st->pz.k[j] + m[9][p];
            }
            st->sxx[p] += (lpm * dvxdx + lam_h * dvzdz) * dt;
            st->szz[p] += (lpm * dvzdz + lam_h * dvxdx) * dt;
        }
#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s; j < nz - s - 1; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx, z = (real)(j - s) * dz + hdz;
            real mu_half_z = (real)0.5 * (st->mu[p + 1] + st->mu[p]);
            real dvzdx = d_bm(st->vz, p, SX, c, s) / dx;
            real dvxdz = d_fm(st->vx, p, 1, c, s) / dz;
            if (in_pml2(st, x, z)) {
                m[8][p] = st->px.b[i] * m[8][p] + st->px.a[i] * dvzdx;
                m[7][p] = st->pz.b_h[j] * m[7][p] + st->pz.a_h[j] * dvxdz;
                dvzdx = dvzdx / st->px.k[i] + m[8][p];
                dvxdz = dvxdz / st->pz.k_h[j] + m[7][p];
            }
            st->sxz[p] += mu_half_z * (dvxdz + dvzdx) * dt;
        }
    st->data[it] = st->vz[I2(st->rsx, st->rsz)];
}

/* src/Elastic2D_CPML.py:202-274 (update_SSG_SV). */
void or2d_sv_step(or_state2d *st, int32_t it, real amp)
{
    const int s = st->s, nx = st->nx, nz = st->nz;
    const size_t SX = (size_t)nz;
    const real dx = st->dx, dz = st->dz, dt = st->dt, hdx = dx / 2, hdz = dz / 2;
    const real *c = st->c;
    real **m = st->mem;

#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s + 1; j < nz - s; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx + hdx, z = (real)(j - s) * dz;
            real dsxxdx = d_fm(st->sxx, p, SX, c, s) / dx;
            real dsxzdz = d_bm(st->sxz, p, 1, c, s) / dz;
            real rho_half_x = (st->rho[p] + st->rho[p + SX]) / (real)2;
            if (in_pml2(st, x, z)) {
                m[0][p] = st->px.b_h[i] * m[0][p] + st->px.a_h[i] * dsxxdx;
                m[3][p] = st->pz.b[j] * m[3][p] + st->pz.a[j] * dsxzdz;
                dsxxdx = dsxxdx / st->px.k_h[i] + m[0][p];
                dsxzdz = dsxzdz / st->pz.k[j] + m[3][p];
            }
            st->vx[p] += (dsxxdx + dsxzdz) * dt / rho_half_x;
        }
#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s; j < nz - s - 1; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx, z = (real)(j - s) * dz + hdz;
            real rho_half_z = (real)0.5 * (st->rho[p] + st->rho[p + 1]);
            real dsxzdx = d_bm(st->sxz, p, SX, c, s) / dx;
            real dszzdz = d_fm(st->szz, p, 1, c, s) / dz;
            if (in_pml2(st, x, z)) {
                m[2][p] = st->px.b[i] * m[2][p] + st->px.a[i] * dsxzdx;
                m[5][p] = st->pz.b_h[j] * m[5][p] + st->pz.a_h[j] * dszzdz;
                dsxzdx = dsxzdx / st->px.k[i] + m[2][p];
                dszzdz = dszzdz / st->pz.k_h[j] + m[5][p];
            }
            st->vz[p] += (dszzdz + dsxzdx) * dt / rho_half_z;
        }
    {
        size_t p = I2(st->isx, st->isz);
        st->vx[p] += dt * amp / st->rho[p];
    }
#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < nz - s; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx, z = (real)(j - s) * dz;
            real lam = st->lam[p];
            real lpm = st->lam[p] + (real)2 * st->mu[p];
            real dvxdx = d_bm(st->vx, p, SX, c, s) / dx;
            real dvzdz = d_bm(st->vz, p, 1, c, s) / dz;
            if (in_pml2(st, x, z)) {
                m[6][p] = st->px.b[i] * m[6][p] + st->px.a[i] * dvxdx;
                m[9][p] = st->pz.b[j] * m[9][p] + st->pz.a[j] * dvzdz;
                dvxdx = dvxdx / st->px.k[i] + m[6][p];
                dvzdz = dvzdz / st->pz.k[j] + m[9][p];
            }
            st->sxx[p] += (lpm * dvxdx + lam * dvzdz) * dt;
            st->szz[p] += (lpm * dvzdz + lam * dvxdx) * dt;
        }
#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < nz - s; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx + hdx, z = (real)(j - s) * dz + hdz;
            real mu_a = (real)0.25 * (st->mu[p + 1] + st->mu[p] + st->mu[p + SX + 1] + st->mu[p + SX]);
            real dvzdx = d_fm(st->vz, p, SX, c, s) / dx;
            real dvxdz = d_fm(st->vx, p, 1, c, s) / dz;
            if (in_pml2(st, x, z)) {
                m[8][p] = st->px.b_h[i] * m[8][p] + st->px.a_h[i] * dvzdx;
                m[7][p] = st->pz.b_h[j] * m[7][p] + st->pz.a_h[j] * dvxdz;
                dvzdx = dvzdx / st->px.k_h[i] + m[8][p];
                dvxdz = dvxdz / st->pz.k_h[j] + m[7][p];
            }
            st->sxz[p] += mu_a * (dvxdz + dvzdx) * dt;
        }
    st->data[it] = st->vx[I2(st->rsx, st->rsz)];
}

/* src/Elastic2D_CPML.py:276-353 (update_RSG). */
void or2d_rsg_step(or_state2d *st, real source)
{
    const int s = st->s, nx = st->nx, nz = st->nz;
    const real dx = st->dx, dz = st->dz, dt = st->dt, hdx = dx / 2, hdz = dz / 2;
    const real dx2 = (real)2 * dx, dz2 = (real)2 * dz;
    real **m = st->mem;

#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s + 1; i < nx - s; ++i)
        for (int j = s + 1; j < nz - s; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx, z = (real)(j - s) * dz;
            real sxxrx = drx_bm(st, st->sxx, i, j), sxxrz = drz_bm(st, st->sxx, i, j);
            real szzrx = drx_bm(st, st->szz, i, j), szzrz = drz_bm(st, st->szz, i, j);
            real sxzrx = drx_bm(st, st->sxz, i, j), sxzrz = drz_bm(st, st->sxz, i, j);
            real dsxxdx = (sxxrz - sxxrx) / dx2;
            real dszzdz = (szzrz + szzrx) / dz2;
            real dsxzdx = (sxzrz - sxzrx) / dx2;
            real dsxzdz = (sxzrz + sxzrx) / dz2;
            real rho = st->rho[p];
            if (in_pml2(st, x, z)) {
                m[0][p] = st->px.b[i] * m[0][p] + st->px.a[i] * dsxxdx;
                m[5][p] = st->pz.b[j] * m[5][p] + st->pz.a[j] * dszzdz;
                m[2][p] = st->px.b[i] * m[2][p] + st->px.a[i] * dsxzdx;
                m[3][p] = st->pz.b[j] * m[3][p] + st->pz.a[j] * dsxzdz;
                dsxxdx = dsxxdx / st->px.k[i] + m[0][p];
                dszzdz = dszzdz / st->pz.k[j] + m[5][p];
                dsxzdx = dsxzdx / st->px.k[i] + m[2][p];
                dsxzdz = dsxzdz / st->pz.k[j] + m[3][p];
            }
            st->vx[p] += (dsxxdx + dsxzdz) * dt / rho;
            st->vz[p] += (dszzdz + dsxzdx) * dt / rho;
        }
    for (int a = -1; a < 1; ++a)
        for (int b = -1; b < 1; ++b) {
            size_t p = I2(st->isx + a, st->isz + b);
            st->vz[p] += dt * source / ((real)8 * st->rho[p]);
        }
#pragma omp parallel for schedule(static) if ((long)nx * nz > OR_PAR_MIN)
    for (int i = s; i < nx - s - 1; ++i)
        for (int j = s; j < nz - s - 1; ++j) {
            size_t p = I2(i, j);
            real x = (real)(i - s) * dx + hdx, z = (real)(j - s) * dz + hdz;
            real lam = st->lam[p], mu = st->mu[p];
            real lpm = lam + (real)2 * mu;
            real vxrx = drx_fm(st, st->vx, i, j), vxrz = drz_fm(st, st->vx, i, j);
            real vzrx = drx_fm(st, st->vz, i, j), vzrz = drz_fm(st, st->vz, i, j);
            real dvxdx = (vxrz - vxrx) / dx2;
            real dvxdz = (vxrz + vxrx) / dz2;
            real dvzdx = (vzrz - vzrx) / dx2;
            real dvzdz = (vzrz + vzrx) / dz2;
            if (in_pml2(st, x, z)) {
                m[6][p] = st->px.b_h[i] * m[6][p] + st->px.a_h[i] * dvxdx;
                m[7][p] = st->pz.b_h[j] * m[7][p] + st->pz.a_h[j] * dvxdz;
                m[8][p] = st->px.b_h[i] * m[8][p] + st->px.a_h[i] * dvzdx;
                m[9][p] = st->pz.b_h[j] * m[9][p] + st->pz.a_h[j] * dvzdz;
                dvxdx = dvxdx / st->px.k_h[i] + m[6][p];
                dvxdz = dvxdz / st->pz.k_h[j] + m[7][p];
                dvzdx = dvzdx / st->px.k_h[i] + m[8][p];
                dvzdz = dvzdz / st->pz.k_h[j] + m[9][p];
            }
            st->sxx[p] += (lpm * dvxdx + lam * dvzdz) * dt;
            st->szz[p] += (lpm * dvzdz + lam * dvxdx) * dt;
            st->sxz[p] += (mu * (dvxdz + dvzdx)) * dt;
        }
}

/* src/Elastic2D_CPML.py:113-119: lambda = 0 and rho halved on planes i < s. */
void or2d_aea_init(or_state2d *st)
{
    for (int i = 0; i < st->s; ++i)
        for (int j = 0; j < st->nz; ++j) {
            size_t p = I2(i, j);
            st->lam[p] = (real)0;
            st->rho[p] = (real)0.5 * st->rho[p];
        }
}

/* src/Elastic2D_CPML.py:121-125: sxx = 0 on plane i = s. */
void or2d_aea(or_state2d *st)
{
    for (int j = 0; j < st->nz; ++j) st->sxx[I2(st->s, j)] = (real)0;
}

/* src/Elastic3D_SSG_Tensor.py:151-157: lambda = 0 and rho halved on planes
 * k < s (the companion per-step AEA at :159-163 is never called). */
void or3d_aea_init(or_state3d *st, real *rho_w, real *lam_w)
{
    for (int i = 0; i < st->nx; ++i)
        for (int j = 0; j < st->ny; ++j)
            for (int k = 0; k < st->s; ++k) {
                size_t p = I3(i, j, k);
                lam_w[p] = (real)0;
                rho_w[p] = (real)0.5 * rho_w[p];
            }
}

/* Thread count of the OpenMP loops (the CPU baseline states what it used). */
#ifdef _OPENMP
#include <omp.h>
#endif
int or_set_threads(int32_t n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
