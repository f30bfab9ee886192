/*
 * les_oracle.c -- CPU FP64 restatement of the LES per-timestep hot path.
 *
 * TEST INFRASTRUCTURE: this is the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each routine cites the reference routine it restates (file:line in the
 * KhaledGhannam/LES sources).  Evaluation order of floating-point
 * expressions follows the Fortran text; the FP32 round trips of
 * lagrange_Sdep.f90 are reproduced as (double)(float)x.
 *
 * The reference calls FFTW 2.1.5 (not vendored); the transforms here are a
 * project-written mixed-radix Stockham FFT with FFTW-2 rfftwnd in-place
 * semantics (unnormalised, forward sign -1, halfcomplex-style c2r that
 * ignores Im of the kx=0 and kx=nx/2 entries), with pad columns defined = 0
 * after every inverse transform.
 */
#include "les_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BOGUS (-1234567890.0)
static const double PI = 3.1415926535897932384626433;
static const double VONK = 0.41;   /* param.f90:98 */
static const double TADV1 = 1.5;   /* param.f90:151 */
static const double TADV2 = -0.5;  /* 1 - tadv1 */

/* Planted mistakes, for the tests that prove the closed-form pins catch them
   (tests/test_oracle_mutations.py builds the oracle with -DOR_MUTANT=n into a
   temporary directory).  The oracle's own build leaves OR_MUTANT at 0.
     1  lagrange_Sdep: the 2-Delta / 4-Delta factors tf1^2 = 4 and tf2^2 = 16 swapped
     2  lagrange_Sdep: Beta = Cs2_2d / Cs2_4d (inverted)
     3  step: AB2 weights tadv1 / tadv2 swapped
     4  step: mean pressure forcing applied to v instead of u
     5  press_stag_array: sign of the divtz boundary rows flipped
     6  lagrange_scalar_Sdep: the factors 4 and 16 swapped
     7  lagrange_scalar_Sdep: Beta inverted
     8  step: no Euler first step when initu is false
     9  step: Coriolis term of RHSy with the wrong sign
    10  step: sponge of RHSx, RHSy taken at the w-node below (sponge(jz)) instead
        of the average of sponge(jz) and sponge(jz+1)
    11  interpolag_Ssim: the w_lag clamp of interpolag_Sdep applied
    12  model 1: the wall level's mixing length taken at z = dz instead of dz/2 */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif

typedef struct { double re, im; } cpx;

/* ------------------------------------------------------------------ */
/* 1-D complex FFT: Stockham autosort, radices 4,2,3,5                 */
/* ------------------------------------------------------------------ */
typedef struct { int n, ns, rad[48]; cpx *tw; } fplan;

static void fplan_init(fplan *p, int n) {
  p->n = n; p->ns = 0;
  int m = n;
  while (m % 4 == 0) { p->rad[p->ns++] = 4; m /= 4; }
  while (m % 2 == 0) { p->rad[p->ns++] = 2; m /= 2; }
  while (m % 3 == 0) { p->rad[p->ns++] = 3; m /= 3; }
  while (m % 5 == 0) { p->rad[p->ns++] = 5; m /= 5; }
  if (m != 1) { fprintf(stderr, "les_oracle: FFT length %d not 2^a3^b5^c\n", n); abort(); }
  p->tw = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (int k = 0; k < n; k++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    p->tw[k].re = (double)cosl(a);
    p->tw[k].im = (double)sinl(a);
  }
}
static void fplan_free(fplan *p) { free(p->tw); p->tw = NULL; }

/* x: data (length n), t: scratch (length n).  sign = -1 forward, +1 backward. */
static void fft_exec(const fplan *p, cpx *x, cpx *t, int sign) {
  const int n = p->n;
  cpx *src = x, *dst = t;
  int Ns = 1;
  const double s3 = 0.86602540378443864676372317075294; /* sqrt(3)/2 */
  const double c51 = 0.30901699437494742410229341718282, s51 = 0.95105651629515357211643933337938;
  const double c52 = -0.80901699437494742410229341718282, s52 = 0.58778525229247312916870595463907;
  for (int s = 0; s < p->ns; s++) {
    const int R = p->rad[s], nR = n / R, tws = n / (Ns * R);
    for (int j = 0; j < nR; j++) {
      const int k = j % Ns;
      cpx v[5] = { { 0, 0 } };
      for (int r = 0; r < R; r++) {
        cpx a = src[j + r * nR];
        if (r && k) {
          int idx = (r * k * tws) % n;
          double wr = p->tw[idx].re, wi = sign < 0 ? p->tw[idx].im : -p->tw[idx].im;
          cpx b = { a.re * wr - a.im * wi, a.re * wi + a.im * wr };
          a = b;
        }
        v[r] = a;
      }
      cpx y[5] = { { 0, 0 } };
      if (R == 2) {
        y[0].re = v[0].re + v[1].re; y[0].im = v[0].im + v[1].im;
        y[1].re = v[0].re - v[1].re; y[1].im = v[0].im - v[1].im;
      } else if (R == 4) {
        cpx a0 = { v[0].re + v[2].re, v[0].im + v[2].im };
        cpx a1 = { v[0].re - v[2].re, v[0].im - v[2].im };
        cpx a2 = { v[1].re + v[3].re, v[1].im + v[3].im };
        cpx d  = { v[1].re - v[3].re, v[1].im - v[3].im };
        /* a3 = d * W, W = exp(sign*i*pi/2) = sign*i */
        cpx a3 = sign < 0 ? (cpx){ d.im, -d.re } : (cpx){ -d.im, d.re };
        y[0].re = a0.re + a2.re; y[0].im = a0.im + a2.im;
        y[1].re = a1.re + a3.re; y[1].im = a1.im + a3.im;
        y[2].re = a0.re - a2.re; y[2].im = a0.im - a2.im;
        y[3].re = a1.re - a3.re; y[3].im = a1.im - a3.im;
      } else if (R == 3) {
        cpx sm = { v[1].re + v[2].re, v[1].im + v[2].im };
        cpx df = { v[1].re - v[2].re, v[1].im - v[2].im };
        cpx t1 = { v[0].re - 0.5 * sm.re, v[0].im - 0.5 * sm.im };
        double f = sign < 0 ? -s3 : s3;          /* t2 = i*f*df */
        cpx t2 = { -f * df.im, f * df.re };
        y[0].re = v[0].re + sm.re; y[0].im = v[0].im + sm.im;
        y[1].re = t1.re + t2.re; y[1].im = t1.im + t2.im;
        y[2].re = t1.re - t2.re; y[2].im = t1.im - t2.im;
      } else { /* R == 5 */
        double sg = sign < 0 ? -1.0 : 1.0;
        cpx a1 = { v[1].re + v[4].re, v[1].im + v[4].im }, b1 = { v[1].re - v[4].re, v[1].im - v[4].im };
        cpx a2 = { v[2].re + v[3].re, v[2].im + v[3].im }, b2 = { v[2].re - v[3].re, v[2].im - v[3].im };
        y[0].re = v[0].re + a1.re + a2.re; y[0].im = v[0].im + a1.im + a2.im;
        cpx c1 = { v[0].re + c51 * a1.re + c52 * a2.re, v[0].im + c51 * a1.im + c52 * a2.im };
        cpx c2 = { v[0].re + c52 * a1.re + c51 * a2.re, v[0].im + c52 * a1.im + c51 * a2.im };
        cpx d1 = { sg * (s51 * b1.re + s52 * b2.re), sg * (s51 * b1.im + s52 * b2.im) };
        cpx d2 = { sg * (s52 * b1.re - s51 * b2.re), sg * (s52 * b1.im - s51 * b2.im) };
        /* y1 = c1 + i d1, y4 = c1 - i d1, y2 = c2 + i d2, y3 = c2 - i d2 */
        y[1].re = c1.re - d1.im; y[1].im = c1.im + d1.re;
        y[4].re = c1.re + d1.im; y[4].im = c1.im - d1.re;
        y[2].re = c2.re - d2.im; y[2].im = c2.im + d2.re;
        y[3].re = c2.re + d2.im; y[3].im = c2.im - d2.re;
      }
      const int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) dst[base + r * Ns] = y[r];
    }
    cpx *tmp = src; src = dst; dst = tmp;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(cpx) * (size_t)n);
}

void or_cfft1(int n, double *re_im, int sign) {
  fplan p; fplan_init(&p, n);
  cpx *t = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  fft_exec(&p, (cpx *)re_im, t, sign);
  free(t); fplan_free(&p);
}

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
typedef struct { int nx, ny, lh, ld; fplan px, py; } plan2;

struct or_ctx {
  or_params P;
  int nx, ny, nz, lh, ld, nx2, ny2, lh_big, ld_big;
  double dx, dy, dz, delta;
  double *kx, *ky, *k2, *Gt, *Gtt;
  plan2 sm, bg;
  int flm_init, fqn_init;   /* SAVE flags, lagrange_Sdep.f90:67-68 */
  int jt, jt_total;         /* mirrors param jt/jt_total for routine calls */
  int sgs_model;            /* param.f90 'model' of the whole step: 1..6 (5 default; 7 = 5 + scalar model 7) */
  or_scalar_params S;       /* scalar switches and constants (zero: no scalars) */
  double *wt_s, *wq_s;      /* (nx,ny) kinematic surface fluxes */
  int glm_init, gqn_init;   /* SAVE flags, lagrange_scalar_Sdep.f90:65-66 */
  or_forcing F;             /* Coriolis and sponge switches (or_set_forcing) */
  double coriol, cug, cvg;  /* coriol, coriol*ug, coriol*vg */
  double *sponge;           /* topbc sponge(1..nz) at [1..nz]; [0] unused */
};

static void plan2_init(plan2 *p, int nx, int ny) {
  p->nx = nx; p->ny = ny; p->lh = nx / 2 + 1; p->ld = 2 * p->lh;
  fplan_init(&p->px, nx); fplan_init(&p->py, ny);
}

/* fft.f90:54-81 init_wavenumber, incl. the ky(lh,:)=0 quirk */
static void init_wavenumber(or_ctx *c) {
  const int lh = c->lh, ny = c->ny;
  for (int jy = 0; jy < ny; jy++)
    for (int jx = 0; jx < lh; jx++) {
      double kx = (jx < lh - 1) ? (double)jx : 0.0;
      int m = ((jy + ny / 2) % ny) - ny / 2;
      double ky = (double)m;
      if (jx == lh - 1) { kx = 0.0; ky = 0.0; }
      if (jy == ny / 2) { kx = 0.0; ky = 0.0; }
      kx = 2.0 * PI / c->P.Lx * kx;
      ky = 2.0 * PI / c->P.Ly * ky;
      c->kx[jx + lh * jy] = kx;
      c->ky[jx + lh * jy] = ky;
      c->k2[jx + lh * jy] = kx * kx + ky * ky;
    }
}

or_ctx *or_create(const or_params *p) {
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->P = *p;
  if (c->P.cs_count <= 0) c->P.cs_count = 5;
  c->nx = p->nx; c->ny = p->ny; c->nz = p->nz;
  c->lh = p->nx / 2 + 1; c->ld = 2 * c->lh;
  c->nx2 = 3 * p->nx / 2; c->ny2 = 3 * p->ny / 2;
  c->lh_big = c->nx2 / 2 + 1; c->ld_big = 2 * c->lh_big;
  c->dx = p->Lx / p->nx; c->dy = p->Ly / p->ny;
  c->dz = p->Lz / (p->nz - 1);                 /* param.f90:64 */
  c->delta = sqrt(c->dx * c->dy);              /* sgs_stag.f90:43, filter_size=1 */
  size_t n2 = (size_t)c->lh * c->ny;
  c->kx = (double *)malloc(n2 * sizeof(double));
  c->ky = (double *)malloc(n2 * sizeof(double));
  c->k2 = (double *)malloc(n2 * sizeof(double));
  c->Gt = (double *)malloc(n2 * sizeof(double));
  c->Gtt = (double *)malloc(n2 * sizeof(double));
  plan2_init(&c->sm, c->nx, c->ny);
  plan2_init(&c->bg, c->nx2, c->ny2);
  init_wavenumber(c);
  or_test_filter_init(c, 2.0, c->Gt);    /* main.f90:166 */
  or_test_filter_init(c, 4.0, c->Gtt);   /* main.f90:169 */
  c->sponge = (double *)calloc((size_t)c->nz + 1, sizeof(double));
  c->sgs_model = 5;
  return c;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  free(c->kx); free(c->ky); free(c->k2); free(c->Gt); free(c->Gtt);
  free(c->wt_s); free(c->wq_s); free(c->sponge);
  fplan_free(&c->sm.px); fplan_free(&c->sm.py); fplan_free(&c->bg.px); fplan_free(&c->bg.py);
  free(c);
}

void or_get_tables(const or_ctx *c, double *kx, double *ky, double *k2, double *G, double *GG) {
  size_t n2 = (size_t)c->lh * c->ny * sizeof(double);
  if (kx) memcpy(kx, c->kx, n2);
  if (ky) memcpy(ky, c->ky, n2);
  if (k2) memcpy(k2, c->k2, n2);
  if (G) memcpy(G, c->Gt, n2);
  if (GG) memcpy(GG, c->Gtt, n2);
}

/* ------------------------------------------------------------------ */
/* 2-D real transforms, FFTW-2 rfftwnd in-place semantics              */
/* ------------------------------------------------------------------ */
/* The rows of a plane, then its columns, are independent 1-D transforms:
   with OpenMP they are shared among the threads (each with its own work
   vectors); every transform is computed the same way whatever the thread
   count, so the results do not depend on it. */
static void rfft2_fwd(const plan2 *P, double *a) {
  const int nx = P->nx, ny = P->ny, lh = P->lh, ld = P->ld;
  const size_t nmax = (size_t)(nx > ny ? nx : ny);
  cpx *out = (cpx *)malloc(sizeof(cpx) * (size_t)lh * ny);
#pragma omp parallel
  {
    cpx *z = (cpx *)malloc(sizeof(cpx) * nmax);
    cpx *t = (cpx *)malloc(sizeof(cpx) * nmax);
#pragma omp for schedule(static)
    for (int y = 0; y < ny; y += 2) {          /* two real rows per complex FFT */
      for (int x = 0; x < nx; x++) { z[x].re = a[x + (size_t)ld * y]; z[x].im = a[x + (size_t)ld * (y + 1)]; }
      fft_exec(&P->px, z, t, -1);
      for (int k = 0; k < lh; k++) {
        cpx zk = z[k % nx], zn = z[(nx - k) % nx];
        /* A = (zk + conj zn)/2, B = (zk - conj zn)/(2i) */
        cpx A = { 0.5 * (zk.re + zn.re), 0.5 * (zk.im - zn.im) };
        double dr = zk.re - zn.re, di = zk.im + zn.im;
        cpx B = { 0.5 * di, -0.5 * dr };
        out[k + (size_t)lh * y] = A;
        out[k + (size_t)lh * (y + 1)] = B;
      }
    }
#pragma omp for schedule(static)
    for (int k = 0; k < lh; k++) {
      for (int y = 0; y < ny; y++) z[y] = out[k + (size_t)lh * y];
      fft_exec(&P->py, z, t, -1);
      for (int y = 0; y < ny; y++) out[k + (size_t)lh * y] = z[y];
    }
    free(z); free(t);
  }
  memcpy(a, out, sizeof(cpx) * (size_t)lh * ny);
  free(out);
}

static void rfft2_inv(const plan2 *P, double *a) {
  const int nx = P->nx, ny = P->ny, lh = P->lh, ld = P->ld;
  const size_t nmax = (size_t)(nx > ny ? nx : ny);
  cpx *X = (cpx *)malloc(sizeof(cpx) * (size_t)lh * ny);
  memcpy(X, a, sizeof(cpx) * (size_t)lh * ny);
#pragma omp parallel
  {
    cpx *z = (cpx *)malloc(sizeof(cpx) * nmax);
    cpx *t = (cpx *)malloc(sizeof(cpx) * nmax);
#pragma omp for schedule(static)
    for (int k = 0; k < lh; k++) {
      for (int y = 0; y < ny; y++) z[y] = X[k + (size_t)lh * y];
      fft_exec(&P->py, z, t, +1);
      for (int y = 0; y < ny; y++) X[k + (size_t)lh * y] = z[y];
    }
#pragma omp for schedule(static)
    for (int y = 0; y < ny; y += 2) {
      cpx *A = X + (size_t)lh * y, *B = X + (size_t)lh * (y + 1);
      A[0].im = 0.0; B[0].im = 0.0; A[nx / 2].im = 0.0; B[nx / 2].im = 0.0;
      for (int k = 0; k <= nx / 2; k++) {
        z[k].re = A[k].re - B[k].im; z[k].im = A[k].im + B[k].re;          /* A + iB */
        if (k > 0 && k < nx / 2) {
          z[nx - k].re = A[k].re + B[k].im; z[nx - k].im = B[k].re - A[k].im; /* conj A + i conj B */
        }
      }
      fft_exec(&P->px, z, t, +1);
      for (int x = 0; x < nx; x++) { a[x + (size_t)ld * y] = z[x].re; a[x + (size_t)ld * (y + 1)] = z[x].im; }
      for (int x = nx; x < ld; x++) { a[x + (size_t)ld * y] = 0.0; a[x + (size_t)ld * (y + 1)] = 0.0; }
    }
    free(z); free(t);
  }
  free(X);
}

void or_rfft2_fwd(or_ctx *c, double *plane, int big) { rfft2_fwd(big ? &c->bg : &c->sm, plane); }
void or_rfft2_inv(or_ctx *c, double *plane, int big) { rfft2_inv(big ? &c->bg : &c->sm, plane); }

/* helpers on complex (lh,ny) planes stored interleaved in double arrays */
#define CRE(a, i) ((a)[2 * (size_t)(i)])
#define CIM(a, i) ((a)[2 * (size_t)(i) + 1])

static void kill_oddballs(const or_ctx *c, double *pl) {
  const int lh = c->lh, ny = c->ny;
  for (int y = 0; y < ny; y++) { CRE(pl, lh - 1 + lh * y) = 0.0; CIM(pl, lh - 1 + lh * y) = 0.0; }
  for (int x = 0; x < lh; x++) { CRE(pl, x + lh * (ny / 2)) = 0.0; CIM(pl, x + lh * (ny / 2)) = 0.0; }
}
/* out = (i*k) * in, as Fortran's eye*k*f (eye*k first) */
static void mul_ik(const or_ctx *c, const double *k, const double *in, double *out) {
  const size_t n = (size_t)c->lh * c->ny;
  for (size_t i = 0; i < n; i++) {
    double ar = CRE(in, i), ai = CIM(in, i), kk = k[i];
    double wr = 0.0 * kk, wi = 1.0 * kk;           /* eye*k */
    CRE(out, i) = wr * ar - wi * ai;
    CIM(out, i) = wr * ai + wi * ar;
  }
}

/* ------------------------------------------------------------------ */
/* filt_da.f90:3-42, ddx.f90:1-31, ddy.f90:1-31, ddxy.f90:1-29          */
/* ------------------------------------------------------------------ */
void or_filt_da(or_ctx *c, double *f, double *dfdx, double *dfdy, int nlev) {
  const size_t pl = (size_t)c->ld * c->ny;
  const double cst = 1.0 / (c->nx * c->ny);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nlev; k++) {
    double *F = f + pl * k, *DX = dfdx + pl * k, *DY = dfdy + pl * k;
    for (size_t i = 0; i < pl; i++) F[i] = cst * F[i];
    rfft2_fwd(&c->sm, F);
    kill_oddballs(c, F);
    mul_ik(c, c->ky, F, DY);
    mul_ik(c, c->kx, F, DX);
    rfft2_inv(&c->sm, F);
    rfft2_inv(&c->sm, DX);
    rfft2_inv(&c->sm, DY);
  }
}

static void ddxy_one(or_ctx *c, double *out, const double *f, const double *k, int nlev) {
  const size_t pl = (size_t)c->ld * c->ny;
  const double cst = 1.0 / (c->nx * c->ny);
#pragma omp parallel for schedule(dynamic)
  for (int l = 0; l < nlev; l++) {
    double *O = out + pl * l; const double *F = f + pl * l;
    for (size_t i = 0; i < pl; i++) O[i] = cst * F[i];
    rfft2_fwd(&c->sm, O);
    kill_oddballs(c, O);
    mul_ik(c, k, O, O);
    rfft2_inv(&c->sm, O);
  }
}
void or_ddx(or_ctx *c, double *dfdx, const double *f, int nlev) { ddxy_one(c, dfdx, f, c->kx, nlev); }
void or_ddy(or_ctx *c, double *dfdy, const double *f, int nlev) { ddxy_one(c, dfdy, f, c->ky, nlev); }

void or_ddxy(or_ctx *c, double *dfdx, double *dfdy, const double *f, int nlev) {
  const size_t pl = (size_t)c->ld * c->ny;
  const double cst = 1.0 / (c->nx * c->ny);
#pragma omp parallel for schedule(dynamic)
  for (int l = 0; l < nlev; l++) {
    double *DX = dfdx + pl * l, *DY = dfdy + pl * l; const double *F = f + pl * l;
    for (size_t i = 0; i < pl; i++) DX[i] = cst * F[i];
    rfft2_fwd(&c->sm, DX);
    kill_oddballs(c, DX);
    mul_ik(c, c->ky, DX, DY);
    mul_ik(c, c->kx, DX, DX);
    rfft2_inv(&c->sm, DX);
    rfft2_inv(&c->sm, DY);
  }
}

/* test_filtermodule.f90:32-63 (ifilter=1 sharp cutoff) */
void or_test_filter_init(or_ctx *c, double alpha, double *G) {
  const int lh = c->lh, ny = c->ny;
  const double delta = alpha * sqrt(c->dx * c->dy);
  const double q = PI / delta, kc2 = q * q;
  for (int i = 0; i < lh * ny; i++) {
    G[i] = 1.0 / (c->nx * c->ny);
    if (c->k2[i] >= kc2) G[i] = 0.0;
  }
  for (int y = 0; y < ny; y++) G[lh - 1 + lh * y] = 0.0;
  for (int x = 0; x < lh; x++) G[x + lh * (ny / 2)] = 0.0;
}

/* test_filtermodule.f90:13-30 */
void or_test_filter(or_ctx *c, double *plane, const double *G) {
  rfft2_fwd(&c->sm, plane);
  const size_t n = (size_t)c->lh * c->ny;
  for (size_t i = 0; i < n; i++) { CRE(plane, i) *= G[i]; CIM(plane, i) *= G[i]; }
  rfft2_inv(&c->sm, plane);
}

/* padd.f90:1-26 (complex planes) */
void or_padd(or_ctx *c, double *big, const double *u) {
  const int lh = c->lh, lhb = c->lh_big, nx = c->nx, ny = c->ny, ny2 = c->ny2;
  memset(big, 0, sizeof(double) * 2 * (size_t)lhb * ny2);
  for (int y = 0; y < ny / 2; y++)
    for (int x = 0; x < nx / 2; x++) {
      CRE(big, x + lhb * y) = CRE(u, x + lh * y); CIM(big, x + lhb * y) = CIM(u, x + lh * y);
    }
  for (int y = 0; y < ny / 2 - 1; y++)
    for (int x = 0; x < nx / 2; x++) {
      int yb = y + ny2 - ny / 2 + 1, ys = y + ny / 2 + 1;
      CRE(big, x + lhb * yb) = CRE(u, x + lh * ys); CIM(big, x + lhb * yb) = CIM(u, x + lh * ys);
    }
}

/* unpadd.f90:1-22 */
void or_unpadd(or_ctx *c, double *cc, const double *big) {
  const int lh = c->lh, lhb = c->lh_big, nx = c->nx, ny = c->ny, ny2 = c->ny2;
  for (int y = 0; y < ny / 2; y++)
    for (int x = 0; x < nx / 2; x++) {
      CRE(cc, x + lh * y) = CRE(big, x + lhb * y); CIM(cc, x + lh * y) = CIM(big, x + lhb * y);
    }
  for (int y = 0; y < ny; y++) { CRE(cc, nx / 2 + lh * y) = 0.0; CIM(cc, nx / 2 + lh * y) = 0.0; }
  for (int x = 0; x < lh; x++) { CRE(cc, x + lh * (ny / 2)) = 0.0; CIM(cc, x + lh * (ny / 2)) = 0.0; }
  for (int y = 0; y < ny / 2 - 1; y++)
    for (int x = 0; x < nx / 2; x++) {
      int yb = y + ny2 - ny / 2 + 1, ys = y + ny / 2 + 1;
      CRE(cc, x + lh * ys) = CRE(big, x + lhb * yb); CIM(cc, x + lh * ys) = CIM(big, x + lhb * yb);
    }
}

/* dealias1.f90:1-31 */
void or_dealias1(or_ctx *c, const double *u, double *u_big, int nlev) {
  const size_t pl = (size_t)c->ld * c->ny, plb = (size_t)c->ld_big * c->ny2;
  const double cst = 1.0 / (c->nx * c->ny);
#pragma omp parallel
  {
    double *tmp = (double *)malloc(sizeof(double) * pl);
#pragma omp for schedule(dynamic)
    for (int k = 0; k < nlev; k++) {
      for (size_t i = 0; i < pl; i++) tmp[i] = cst * u[pl * k + i];
      rfft2_fwd(&c->sm, tmp);
      or_padd(c, u_big + plb * k, tmp);
      rfft2_inv(&c->bg, u_big + plb * k);
    }
    free(tmp);
  }
}

/* dealias2.f90:1-33 */
void or_dealias2(or_ctx *c, double *u, double *u_big, int nlev) {
  const size_t pl = (size_t)c->ld * c->ny, plb = (size_t)c->ld_big * c->ny2;
  const double cst = 1.0 / (c->nx2 * c->ny2);
  for (size_t i = 0; i < plb * nlev; i++) u_big[i] = cst * u_big[i];
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nlev; k++) {
    rfft2_fwd(&c->bg, u_big + plb * k);
    or_unpadd(c, u + pl * k, u_big + plb * k);
    rfft2_inv(&c->sm, u + pl * k);
  }
}

/* ------------------------------------------------------------------ */
/* ddz_uv.f90:5-57, ddz_w.f90:6-51 (non-MPI)                            */
/* ------------------------------------------------------------------ */
void or_ddz_uv(or_ctx *c, double *dfdz, const double *f) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double cst = 1.0 / c->dz;
  for (int k = 1; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = x + (size_t)ld * y;
        dfdz[pl * k + i] = cst * (f[pl * k + i] - f[pl * (k - 1) + i]);
      }
  for (size_t i = 0; i < pl; i++) dfdz[pl * (nz - 1) + i] = 0.0;
}

void or_ddz_w(or_ctx *c, double *dfdz, const double *f) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double cst = 1.0 / c->dz;
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = x + (size_t)ld * y;
        dfdz[pl * k + i] = cst * (f[pl * (k + 1) + i] - f[pl * k + i]);
      }
  for (size_t i = 0; i < pl; i++) dfdz[pl * (nz - 1) + i] = 0.0;
}

/* ------------------------------------------------------------------ */
/* convec.f90:1-248 (lbc_mom='wall', non-MPI)                           */
/* ------------------------------------------------------------------ */
void or_convec(or_ctx *c, const double *u1, const double *u2, const double *u3,
               const double *du1d2, const double *du1d3, const double *du2d1,
               const double *du2d3, const double *du3d1, const double *du3d2,
               double *cx, double *cy, double *cz) {
  const int nz = c->nz;
  const size_t pl = (size_t)c->ld * c->ny, plb = (size_t)c->ld_big * c->ny2;
  double *b_u1 = (double *)calloc(plb * nz, sizeof(double));
  double *b_u2 = (double *)calloc(plb * nz, sizeof(double));
  double *b_u3 = (double *)calloc(plb * nz, sizeof(double));
  double *b_v1 = (double *)calloc(plb * nz, sizeof(double));
  double *b_v2 = (double *)calloc(plb * nz, sizeof(double));
  double *b_v3 = (double *)calloc(plb * nz, sizeof(double));
  double cst = 1.0 / (c->nx * c->ny);
  /* convec.f90:45-65 */
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nz; k++) {
    double *X = cx + pl * k, *Y = cy + pl * k, *Z = cz + pl * k;
    for (size_t i = 0; i < pl; i++) { X[i] = cst * u1[pl * k + i]; Y[i] = cst * u2[pl * k + i]; Z[i] = cst * u3[pl * k + i]; }
    rfft2_fwd(&c->sm, X); rfft2_fwd(&c->sm, Y); rfft2_fwd(&c->sm, Z);
    or_padd(c, b_u1 + plb * k, X); or_padd(c, b_u2 + plb * k, Y); or_padd(c, b_u3 + plb * k, Z);
    rfft2_inv(&c->bg, b_u1 + plb * k); rfft2_inv(&c->bg, b_u2 + plb * k); rfft2_inv(&c->bg, b_u3 + plb * k);
  }
  /* convec.f90:67-114 vorticity */
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nz; k++) {
    double *X = cx + pl * k, *Y = cy + pl * k, *Z = cz + pl * k;
    for (size_t i = 0; i < pl; i++) {
      size_t a = pl * k + i;
      if (k == 0) {
        X[i] = cst * (0.5 * (du3d2[a] + du3d2[pl + i]) - du2d3[a]);
        Y[i] = cst * (du1d3[a] - 0.5 * (du3d1[a] + du3d1[pl + i]));
      } else {
        X[i] = cst * (du3d2[a] - du2d3[a]);
        Y[i] = cst * (du1d3[a] - du3d1[a]);
      }
      Z[i] = cst * (du2d1[a] - du1d2[a]);
    }
    rfft2_fwd(&c->sm, X); rfft2_fwd(&c->sm, Y); rfft2_fwd(&c->sm, Z);
    or_padd(c, b_v1 + plb * k, X); or_padd(c, b_v2 + plb * k, Y); or_padd(c, b_v3 + plb * k, Z);
    rfft2_inv(&c->bg, b_v1 + plb * k); rfft2_inv(&c->bg, b_v2 + plb * k); rfft2_inv(&c->bg, b_v3 + plb * k);
  }
  cst = 1.0 / (c->nx2 * c->ny2);
  /* CX: convec.f90:121-152 */
#pragma omp parallel
  {
    double *cc = (double *)malloc(sizeof(double) * plb);
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz - 1; k++) {
    for (size_t i = 0; i < plb; i++) {
      if (k == 0)
        cc[i] = cst * (b_u2[i] * (-b_v3[i]) + 0.5 * b_u3[plb + i] * (b_v2[plb + i]));
      else
        cc[i] = cst * (b_u2[plb * k + i] * (-b_v3[plb * k + i]) +
                       0.5 * (b_u3[plb * (k + 1) + i] * (b_v2[plb * (k + 1) + i]) +
                              b_u3[plb * k + i] * (b_v2[plb * k + i])));
    }
    rfft2_fwd(&c->bg, cc);
    or_unpadd(c, cx + pl * k, cc);
    rfft2_inv(&c->sm, cx + pl * k);
  }
    free(cc);
  }
  /* CY: convec.f90:157-188 */
#pragma omp parallel
  {
    double *cc = (double *)malloc(sizeof(double) * plb);
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz - 1; k++) {
    for (size_t i = 0; i < plb; i++) {
      if (k == 0)
        cc[i] = cst * (b_u1[i] * (b_v3[i]) + 0.5 * b_u3[plb + i] * (-b_v1[plb + i]));
      else
        cc[i] = cst * (b_u1[plb * k + i] * (b_v3[plb * k + i]) +
                       0.5 * (b_u3[plb * (k + 1) + i] * (-b_v1[plb * (k + 1) + i]) +
                              b_u3[plb * k + i] * (-b_v1[plb * k + i])));
    }
    rfft2_fwd(&c->bg, cc);
    or_unpadd(c, cy + pl * k, cc);
    rfft2_inv(&c->sm, cy + pl * k);
  }
    free(cc);
  }
  /* CZ: convec.f90:192-233 */
#pragma omp parallel
  {
    double *cc = (double *)malloc(sizeof(double) * plb);
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz - 1; k++) {
    for (size_t i = 0; i < plb; i++) {
      if (k == 0)
        cc[i] = 0.0;
      else
        cc[i] = cst * 0.5 * ((b_u1[plb * k + i] + b_u1[plb * (k - 1) + i]) * (-b_v2[plb * k + i]) +
                             (b_u2[plb * k + i] + b_u2[plb * (k - 1) + i]) * (b_v1[plb * k + i]));
    }
    rfft2_fwd(&c->bg, cc);
    or_unpadd(c, cz + pl * k, cc);
    rfft2_inv(&c->sm, cz + pl * k);
  }
    free(cc);
  }
  for (size_t i = 0; i < pl; i++) { cx[pl * (nz - 1) + i] = BOGUS; cy[pl * (nz - 1) + i] = BOGUS; cz[pl * (nz - 1) + i] = BOGUS; }
  free(b_u1); free(b_u2); free(b_u3); free(b_v1); free(b_v2); free(b_v3);
}

/* ------------------------------------------------------------------ */
/* wallstress.f90:5-69, 'wall', WALL_HOMOG_FLAG=.false.                 */
/* ------------------------------------------------------------------ */
void or_wallstress(or_ctx *c, const double *u, const double *v, double *dudz,
                   double *dvdz, double *txz, double *tyz, const double *psi_m,
                   const double *psi_m0, const double *phi_m, double *ustar_avg) {
  const int nx = c->nx, ny = c->ny, ld = c->ld;
  const size_t pl = (size_t)ld * ny;
  double *u1 = (double *)malloc(sizeof(double) * pl), *v1 = (double *)malloc(sizeof(double) * pl);
  memcpy(u1, u, sizeof(double) * pl); memcpy(v1, v, sizeof(double) * pl);
  or_test_filter(c, u1, c->Gtt);
  or_test_filter(c, v1, c->Gtt);
  const double dz = c->dz;
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      size_t i = x + (size_t)ld * y, j = x + (size_t)nx * y;
      double denom = log(0.5 * dz / c->P.z0) - psi_m[j] + psi_m0[j];
      double u_avg = sqrt(u1[i] * u1[i] + v1[i] * v1[i]);
      double us = u_avg * VONK / denom;
      ustar_avg[j] = us;
      double cst = -(us * us) / u_avg;
      txz[i] = cst * u1[i];
      tyz[i] = cst * v1[i];
      double du = us / (0.5 * dz * VONK) * u1[i] / u_avg * phi_m[j];
      double dv = us / (0.5 * dz * VONK) * v1[i] / u_avg * phi_m[j];
      dudz[i] = (u[i] == 0.0) ? 0.0 : du;
      dvdz[i] = (v[i] == 0.0) ? 0.0 : dv;
    }
  free(u1); free(v1);
}

/* ------------------------------------------------------------------ */
/* sgs_stag.f90:371-498 calc_Sij ('wall', non-MPI)                      */
/* ------------------------------------------------------------------ */
void or_calc_sij(or_ctx *c, const double *const g[9], double *const S[6]) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  for (int s = 0; s < 6; s++) memset(S[s], 0, sizeof(double) * pl * nz);
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      size_t i = x + (size_t)ld * y;
      double ux = g[0][i], uy = g[1][i], uz = g[2][i];
      double vx = g[3][i], vy = g[4][i], vz = g[5][i];
      double wx = 0.5 * (g[6][i] + g[6][pl + i]);
      double wy = 0.5 * (g[7][i] + g[7][pl + i]);
      double wz = g[8][i];
      S[0][i] = ux; S[1][i] = 0.5 * (uy + vx); S[2][i] = 0.5 * (uz + wx);
      S[3][i] = vy; S[4][i] = 0.5 * (vz + wy); S[5][i] = wz;
    }
  for (int k = 1; k < nz; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y, m = i - pl;
        double ux = 0.5 * (g[0][i] + g[0][m]);
        double uy = 0.5 * (g[1][i] + g[1][m]);
        double uz = g[2][i];
        double vx = 0.5 * (g[3][i] + g[3][m]);
        double vy = 0.5 * (g[4][i] + g[4][m]);
        double vz = g[5][i];
        double wx = g[6][i];
        double wy = g[7][i];
        double wz = 0.5 * (g[8][i] + g[8][m]);
        S[0][i] = ux; S[1][i] = 0.5 * (uy + vx); S[2][i] = 0.5 * (uz + wx);
        S[3][i] = vy; S[4][i] = 0.5 * (vz + wy); S[5][i] = wz;
      }
}

/* ------------------------------------------------------------------ */
/* interpolag_Sdep.f90:2-380 (non-MPI, read_inflow_file=.false.)        */
/* ------------------------------------------------------------------ */
/* the same interpolation serves interpolag_scalar_Sdep.f90:30-316 on
   G_LM..G_NN with u_lag_s..w_lag_s, and interpolag_Ssim.f90 on F_LM, F_MM
   (nf = 2), which lacks the clamp of w_lag to [-1, 1] at the two lowest
   levels (interpolag_Sdep.f90:264-269) */
static void interpolag_core(or_ctx *c, double *const F[4], double *ul, double *vl, double *wl, int nf, int clamp) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz;
  const double cs = (double)c->P.cs_count, dt = c->P.dt;
  double *FF[4];
  for (int q = 0; q < nf; q++) { FF[q] = (double *)malloc(sizeof(double) * n3); memcpy(FF[q], F[q], sizeof(double) * n3); }
  for (size_t i = 0; i < n3; i++) { ul[i] = ul[i] / cs; vl[i] = vl[i] / cs; }
  for (int k = nz - 1; k >= 1; k--)
    for (size_t i = 0; i < pl; i++) {
      ul[pl * k + i] = 0.5 * (ul[pl * k + i] + ul[pl * (k - 1) + i]);
      vl[pl * k + i] = 0.5 * (vl[pl * k + i] + vl[pl * (k - 1) + i]);
    }
  for (size_t i = 0; i < n3; i++) wl[i] = wl[i] / cs;
  for (size_t i = 0; i < pl; i++) wl[i] = 0.25 * wl[pl + i];
  for (size_t i = 0; i < n3; i++) {
    ul[i] = -ul[i] * dt * cs / c->dx;
    vl[i] = -vl[i] * dt * cs / c->dy;
    wl[i] = -wl[i] * dt * cs / c->dz;
  }
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      size_t i = x + (size_t)ld * y;
      wl[pl + i] = wl[pl + i] + fmin(wl[pl + i], 0.0);
      wl[i] = wl[i] + fmax(wl[i], 0.0);
    }
  for (size_t i = 0; clamp && i < pl; i++) {
    wl[pl + i] = fmin(1.0, wl[pl + i]);
    wl[i] = fmin(1.0, wl[pl + i]);
    wl[pl + i] = fmax(-1.0, wl[pl + i]);
    wl[i] = fmax(-1.0, wl[pl + i]);
  }
  for (int k = 0; k < nz; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        int ax = signbit(ul[i]) ? -1 : 1, ay = signbit(vl[i]) ? -1 : 1, az = signbit(wl[i]) ? -1 : 1;
        int xa = (x + ax + nx) % nx, ya = (y + ay + ny) % ny;
        int za = k + az; if (za < 0) za = 0; if (za > nz - 1) za = nz - 1;
        double cx = fabs(ul[i]), cy = fabs(vl[i]), cz = fabs(wl[i]);
        double fx = 1.0 - cx, fy = 1.0 - cy, fz = 1.0 - cz;
        size_t i000 = pl * k + x + (size_t)ld * y, i00z = pl * za + x + (size_t)ld * y;
        size_t i0y0 = pl * k + x + (size_t)ld * ya, i0yz = pl * za + x + (size_t)ld * ya;
        size_t ix00 = pl * k + xa + (size_t)ld * y, ix0z = pl * za + xa + (size_t)ld * y;
        size_t ixy0 = pl * k + xa + (size_t)ld * ya, ixyz = pl * za + xa + (size_t)ld * ya;
        for (int q = 0; q < nf; q++) {
          const double *G = FF[q];
          F[q][i] = fx * fy * (G[i000] * fz + G[i00z] * cz)
                  + fx * cy * (G[i0y0] * fz + G[i0yz] * cz)
                  + cx * fy * (G[ix00] * fz + G[ix0z] * cz)
                  + cx * cy * (G[ixy0] * fz + G[ixyz] * cz);
        }
      }
  memset(ul, 0, sizeof(double) * n3); memset(vl, 0, sizeof(double) * n3); memset(wl, 0, sizeof(double) * n3);
  for (int q = 0; q < nf; q++) free(FF[q]);
}

void or_interpolag_sdep(or_ctx *c, or_sgs_state *s) {
  double *F[4] = { s->F_LM, s->F_MM, s->F_QN, s->F_NN };
  interpolag_core(c, F, s->u_lag, s->v_lag, s->w_lag, 4, 1);
}

void or_interpolag_ssim(or_ctx *c, or_sgs_state *s) {
  double *F[4] = { s->F_LM, s->F_MM, NULL, NULL };
  interpolag_core(c, F, s->u_lag, s->v_lag, s->w_lag, 2, OR_MUTANT == 11);
}

void or_interpolag_scalar_sdep(or_ctx *c, or_scalar_sgs *g) {
  double *F[4] = { g->G_LM, g->G_MM, g->G_QN, g->G_NN };
  interpolag_core(c, F, g->u_lag_s, g->v_lag_s, g->w_lag_s, 4, 1);
}

/* ------------------------------------------------------------------ */
/* lagrange_Sdep.f90:4-478 (non-MPI, no buildings, no inflow)           */
/* ------------------------------------------------------------------ */
static inline double fq(double x) { return (double)(float)x; }   /* Fortran real(x) */

/* Plane sums (Fortran sum() over a level, whose order the compiler chooses):
   compensated with an error-free TwoSum per addition, so the result is the
   exact sum rounded once up to a term of order n eps^2, independent of the
   order.  The plane means of theta (~283 K) that calcbeta subtracts and
   scales by g z_i/u*^2 (scalars_module.f90:199-220) would otherwise carry the
   rounding of a sequential sum. */
typedef struct { double s, c; } csum_t;
static inline void csum_add(csum_t *a, double x) {
  const double t = a->s + x, bb = t - a->s;
  a->c += (a->s - (t - bb)) + (x - bb);
  a->s = t;
}
static inline double csum_val(const csum_t *a) { return a->s + a->c; }
static const float F1EM24 = 1e-24f;                               /* 1E-24 literal */

void or_lagrange_sdep(or_ctx *c, const double *const S[6], const double *u,
                      const double *v, const double *w, or_sgs_state *s) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double delta = c->delta, tf1 = 2.0, tf2 = 4.0, tf1_2 = tf1 * tf1, tf2_2 = tf2 * tf2;
  const double opftdelta = 1.5 * delta, powcoeff = -1.0 / 8.0;
  const double e24 = (double)F1EM24;
  or_interpolag_sdep(c, s);
  const double lagran_dt = c->P.dt * (double)c->P.cs_count;
  const double cst = 2.0 * (delta * delta);
  const double expo = log(tf1) / (log(tf2) - log(tf1));
  enum { UB, VB, WB, UH, VH, WH, L11, L12, L13, L22, L23, L33, Q11, Q12, Q13, Q22, Q23, Q33,
         SB11, SB12, SB13, SB22, SB23, SB33, SH11, SH12, SH13, SH22, SH23, SH33,
         SSB11, SSB12, SSB13, SSB22, SSB23, SSB33, SSH11, SSH12, SSH13, SSH22, SSH23, SSH33,
         SMAG, NT };
  const int jt = c->jt;
  const int init_lm = c->P.inilag && !c->flm_init && (jt == c->P.cs_count || jt == 1);
  const int init_qn = c->P.inilag && !c->fqn_init && (jt == c->P.cs_count || jt == 1);
#pragma omp parallel
  {
    double *T[NT];
    for (int t = 0; t < NT; t++) T[t] = (double *)calloc(pl, sizeof(double));
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz; k++) {
    const size_t o = pl * k;
    for (size_t i = 0; i < pl; i++) {
      double ub, vb, wb;
      if (k == 0) { ub = u[i]; vb = v[i]; wb = 0.25 * w[pl + i]; }
      else { ub = 0.5 * (u[o + i] + u[o - pl + i]); vb = 0.5 * (v[o + i] + v[o - pl + i]); wb = w[o + i]; }
      T[UB][i] = ub; T[VB][i] = vb; T[WB][i] = wb;
      T[UH][i] = ub; T[VH][i] = vb; T[WH][i] = wb;
      T[L11][i] = ub * ub; T[L12][i] = ub * vb; T[L13][i] = ub * wb;
      T[L23][i] = vb * wb; T[L22][i] = vb * vb; T[L33][i] = wb * wb;
      T[Q11][i] = ub * ub; T[Q12][i] = ub * vb; T[Q13][i] = ub * wb;
      T[Q22][i] = vb * vb; T[Q23][i] = vb * wb; T[Q33][i] = wb * wb;
    }
    or_test_filter(c, T[UB], c->Gt); or_test_filter(c, T[VB], c->Gt); or_test_filter(c, T[WB], c->Gt);
    const int Lp[6][3] = { { L11, UB, UB }, { L12, UB, VB }, { L13, UB, WB }, { L22, VB, VB }, { L23, VB, WB }, { L33, WB, WB } };
    for (int q = 0; q < 6; q++) {
      or_test_filter(c, T[Lp[q][0]], c->Gt);
      for (size_t i = 0; i < pl; i++) T[Lp[q][0]][i] = T[Lp[q][0]][i] - T[Lp[q][1]][i] * T[Lp[q][2]][i];
    }
    or_test_filter(c, T[UH], c->Gtt); or_test_filter(c, T[VH], c->Gtt); or_test_filter(c, T[WH], c->Gtt);
    const int Qp[6][3] = { { Q11, UH, UH }, { Q12, UH, VH }, { Q13, UH, WH }, { Q22, VH, VH }, { Q23, VH, WH }, { Q33, WH, WH } };
    for (int q = 0; q < 6; q++) {
      or_test_filter(c, T[Qp[q][0]], c->Gtt);
      for (size_t i = 0; i < pl; i++) T[Qp[q][0]][i] = T[Qp[q][0]][i] - T[Qp[q][1]][i] * T[Qp[q][2]][i];
    }
    memcpy(s->L11t + o, T[L11], sizeof(double) * pl);
    memcpy(s->L22t + o, T[L22], sizeof(double) * pl);
    memcpy(s->L33t + o, T[L33], sizeof(double) * pl);
    memcpy(s->Q11t + o, T[Q11], sizeof(double) * pl);
    memcpy(s->Q22t + o, T[Q22], sizeof(double) * pl);
    memcpy(s->Q33t + o, T[Q33], sizeof(double) * pl);
    for (size_t i = 0; i < pl; i++) {
      double s11 = S[0][o + i], s12 = S[1][o + i], s13 = S[2][o + i];
      double s22 = S[3][o + i], s23 = S[4][o + i], s33 = S[5][o + i];
      double Sm = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
      T[SMAG][i] = Sm;
      T[SB11][i] = s11; T[SB12][i] = s12; T[SB13][i] = s13; T[SB22][i] = s22; T[SB23][i] = s23; T[SB33][i] = s33;
      T[SH11][i] = s11; T[SH12][i] = s12; T[SH13][i] = s13; T[SH22][i] = s22; T[SH23][i] = s23; T[SH33][i] = s33;
      T[SSB11][i] = Sm * s11; T[SSB12][i] = Sm * s12; T[SSB13][i] = Sm * s13;
      T[SSB22][i] = Sm * s22; T[SSB23][i] = Sm * s23; T[SSB33][i] = Sm * s33;
      T[SSH11][i] = T[SSB11][i]; T[SSH12][i] = T[SSB12][i]; T[SSH13][i] = T[SSB13][i];
      T[SSH22][i] = T[SSB22][i]; T[SSH23][i] = T[SSB23][i]; T[SSH33][i] = T[SSB33][i];
    }
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SB11 + q], c->Gt);
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SH11 + q], c->Gtt);
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SSB11 + q], c->Gt);
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SSH11 + q], c->Gtt);
    double *FLM = s->F_LM + o, *FMM = s->F_MM + o, *FQN = s->F_QN + o, *FNN = s->F_NN + o;
    double sum2d = 0.0, sum4d = 0.0, sumcs = 0.0;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < ld; x++) {
        size_t i = x + (size_t)ld * y;
        double Sb = sqrt(2.0 * (T[SB11][i] * T[SB11][i] + T[SB22][i] * T[SB22][i] + T[SB33][i] * T[SB33][i] +
                                2.0 * (T[SB12][i] * T[SB12][i] + T[SB13][i] * T[SB13][i] + T[SB23][i] * T[SB23][i])));
        double Sh = sqrt(2.0 * (T[SH11][i] * T[SH11][i] + T[SH22][i] * T[SH22][i] + T[SH33][i] * T[SH33][i] +
                                2.0 * (T[SH12][i] * T[SH12][i] + T[SH13][i] * T[SH13][i] + T[SH23][i] * T[SH23][i])));
        double M[6], N[6], Lv[6], Qv[6];
        for (int q = 0; q < 6; q++) {
          M[q] = cst * (T[SSB11 + q][i] - (OR_MUTANT == 1 ? tf2_2 : tf1_2) * Sb * T[SB11 + q][i]);
          N[q] = cst * (T[SSH11 + q][i] - (OR_MUTANT == 1 ? tf1_2 : tf2_2) * Sh * T[SH11 + q][i]);
        }
        Lv[0] = T[L11][i]; Lv[1] = T[L12][i]; Lv[2] = T[L13][i]; Lv[3] = T[L22][i]; Lv[4] = T[L23][i]; Lv[5] = T[L33][i];
        Qv[0] = T[Q11][i]; Qv[1] = T[Q12][i]; Qv[2] = T[Q13][i]; Qv[3] = T[Q22][i]; Qv[4] = T[Q23][i]; Qv[5] = T[Q33][i];
        /* index map q: 0=11 1=12 2=13 3=22 4=23 5=33 */
        double LM = Lv[0] * M[0] + Lv[3] * M[3] + Lv[5] * M[5] + 2.0 * (Lv[1] * M[1] + Lv[2] * M[2] + Lv[4] * M[4]);
        double MM = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
        double QN = Qv[0] * N[0] + Qv[3] * N[3] + Qv[5] * N[5] + 2.0 * (Qv[1] * N[1] + Qv[2] * N[2] + Qv[4] * N[4]);
        double NN = N[0] * N[0] + N[3] * N[3] + N[5] * N[5] + 2.0 * (N[1] * N[1] + N[2] * N[2] + N[4] * N[4]);
        if (init_lm) {
          FMM[i] = MM; FLM[i] = 0.03 * MM;
          if (x >= ld - 2) { FMM[i] = 1.0; FLM[i] = 1.0; }
        }
        double Tn = fq(fmax((double)(float)(FLM[i] * FMM[i]), e24));
        s->TnLMMM[o + i] = Tn;
        Tn = opftdelta * pow(Tn, powcoeff);
        Tn = fmax(e24, fq(Tn));
        double dumfac = lagran_dt / Tn;
        double epsi = dumfac / (1.0 + dumfac);
        FLM[i] = (epsi * LM + (1.0 - epsi) * FLM[i]);
        FMM[i] = (epsi * MM + (1.0 - epsi) * FMM[i]);
        FLM[i] = fmax(e24, fq(FLM[i]));
        double cs2d = FLM[i] / FMM[i];
        if (x >= ld - 2) cs2d = e24;
        cs2d = fmax(e24, fq(cs2d));
        if (init_qn) {
          FNN[i] = NN; FQN[i] = 0.03 * NN;
          if (x >= ld - 2) { FNN[i] = 1.0; FQN[i] = 1.0; }
        }
        Tn = fmax(fq(FQN[i] * FNN[i]), e24);
        s->TnQNNN[o + i] = Tn;
        Tn = opftdelta * pow(Tn, powcoeff);
        Tn = fmax(e24, fq(Tn));
        dumfac = lagran_dt / Tn;
        epsi = dumfac / (1.0 + dumfac);
        FQN[i] = (epsi * QN + (1.0 - epsi) * FQN[i]);
        FNN[i] = (epsi * NN + (1.0 - epsi) * FNN[i]);
        FQN[i] = fmax(e24, fq(FQN[i]));
        double cs4d = FQN[i] / FNN[i];
        if (x >= ld - 2) cs4d = e24;
        cs4d = fmax(e24, fq(cs4d));
        double beta = OR_MUTANT == 2 ? pow(cs2d / cs4d, expo) : pow(cs4d / cs2d, expo);
        if (k == nz - 1) beta = 1.0;
        s->Beta[o + i] = beta;
        double bclip = fmax(fq(beta), fq(1.0 / (tf1 * tf2)));
        double cso = cs2d / bclip;
        if (x >= ld - 2) cso = e24;
        s->Cs_opt2[o + i] = fmax(e24, fq(cso));
        if (x < nx) { sum2d += cs2d; sum4d += cs4d; sumcs += s->Cs_opt2[o + i]; }
      }
    s->Beta_avg[k] = sum2d / sumcs;
    s->Betaclip_avg[k] = sum4d / sum2d;
  }
    for (int t = 0; t < NT; t++) free(T[t]);
  }
  if (init_lm) c->flm_init = 1;
  if (init_qn) c->fqn_init = 1;
}

/* ------------------------------------------------------------------ */
/* the 2-Delta filtered quantities of one level shared by lagrange_Ssim  */
/* (lagrange_Ssim.f90:142-215) and std_dynamic (std_dynamic.f90:30-122):  */
/* L_ij = bar(u_i u_j) - bar(u_i) bar(u_j), bar(S_ij), bar(|S| S_ij).    */
/* Both routines put u, v on the first uv-node and 0.25 w(2) there, and  */
/* w-node averages above; the products are formed before filtering.       */
/* ------------------------------------------------------------------ */
enum { LV_UB, LV_VB, LV_WB, LV_L11, LV_L12, LV_L13, LV_L22, LV_L23, LV_L33,
       LV_SB11, LV_SB12, LV_SB13, LV_SB22, LV_SB23, LV_SB33,
       LV_SS11, LV_SS12, LV_SS13, LV_SS22, LV_SS23, LV_SS33, LV_N };

static void level_2delta(or_ctx *c, const double *const S[6], const double *u, const double *v,
                         const double *w, int k, double *T[LV_N]) {
  const size_t pl = (size_t)c->ld * c->ny, o = pl * k;
  for (size_t i = 0; i < pl; i++) {
    double ub, vb, wb;
    if (k == 0) { ub = u[i]; vb = v[i]; wb = 0.25 * w[pl + i]; }
    else { ub = 0.5 * (u[o + i] + u[o - pl + i]); vb = 0.5 * (v[o + i] + v[o - pl + i]); wb = w[o + i]; }
    T[LV_UB][i] = ub; T[LV_VB][i] = vb; T[LV_WB][i] = wb;
    T[LV_L11][i] = ub * ub; T[LV_L12][i] = ub * vb; T[LV_L13][i] = ub * wb;
    T[LV_L22][i] = vb * vb; T[LV_L23][i] = vb * wb; T[LV_L33][i] = wb * wb;
  }
  for (int q = 0; q < 3; q++) or_test_filter(c, T[LV_UB + q], c->Gt);
  const int Lp[6][3] = { { LV_L11, LV_UB, LV_UB }, { LV_L12, LV_UB, LV_VB }, { LV_L13, LV_UB, LV_WB },
                         { LV_L22, LV_VB, LV_VB }, { LV_L23, LV_VB, LV_WB }, { LV_L33, LV_WB, LV_WB } };
  for (int q = 0; q < 6; q++) {
    or_test_filter(c, T[Lp[q][0]], c->Gt);
    for (size_t i = 0; i < pl; i++) T[Lp[q][0]][i] = T[Lp[q][0]][i] - T[Lp[q][1]][i] * T[Lp[q][2]][i];
  }
  for (size_t i = 0; i < pl; i++) {
    const double s11 = S[0][o + i], s12 = S[1][o + i], s13 = S[2][o + i];
    const double s22 = S[3][o + i], s23 = S[4][o + i], s33 = S[5][o + i];
    const double Sm = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
    for (int q = 0; q < 6; q++) { T[LV_SB11 + q][i] = S[q][o + i]; T[LV_SS11 + q][i] = Sm * S[q][o + i]; }
  }
  for (int q = 0; q < 6; q++) or_test_filter(c, T[LV_SB11 + q], c->Gt);
  for (int q = 0; q < 6; q++) or_test_filter(c, T[LV_SS11 + q], c->Gt);
}

/* L:M and M:M at plane point i with M_ij = cst (bar(|S|S_ij) - 4 |S_bar| bar(S_ij))
   (beta = 1: lagrange_Ssim.f90:221-231, std_dynamic.f90:124-132) */
static void level_lm_mm(double *const T[LV_N], size_t i, double cst, double *LM, double *MM) {
  const double *sb[6] = { T[LV_SB11], T[LV_SB12], T[LV_SB13], T[LV_SB22], T[LV_SB23], T[LV_SB33] };
  const double Sb = sqrt(2.0 * (sb[0][i] * sb[0][i] + sb[3][i] * sb[3][i] + sb[5][i] * sb[5][i] +
                                2.0 * (sb[1][i] * sb[1][i] + sb[2][i] * sb[2][i] + sb[4][i] * sb[4][i])));
  double M[6], L[6];
  for (int q = 0; q < 6; q++) {
    M[q] = cst * (T[LV_SS11 + q][i] - 4.0 * Sb * sb[q][i]);
    L[q] = T[LV_L11 + q][i];
  }
  /* q: 0=11 1=12 2=13 3=22 4=23 5=33 */
  *LM = L[0] * M[0] + L[3] * M[3] + L[5] * M[5] + 2.0 * (L[1] * M[1] + L[2] * M[2] + L[4] * M[4]);
  *MM = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
}

/* lagrange_Ssim.f90:4-300 (non-MPI, no buildings, no inflow, Beta = 1) */
void or_lagrange_ssim(or_ctx *c, const double *const S[6], const double *u,
                      const double *v, const double *w, or_sgs_state *s) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double delta = c->delta, eps = 1.e-32;
  const double opftdelta = 1.5 * delta, powcoeff = -1.0 / 8.0;       /* :71-75 */
  or_interpolag_ssim(c, s);                                          /* :123 */
  const double lagran_dt = c->P.dt * (double)c->P.cs_count;
  const double cst = 2.0 * (delta * delta);
  const int jt = c->jt;
  const int init = c->P.inilag && !c->flm_init && (jt == c->P.cs_count || jt == 1);   /* :235-244 */
#pragma omp parallel
  {
    double *T[LV_N];
    for (int t = 0; t < LV_N; t++) T[t] = (double *)calloc(pl, sizeof(double));
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz; k++) {
    const size_t o = pl * k;
    level_2delta(c, S, u, v, w, k, T);
    double *FLM = s->F_LM + o, *FMM = s->F_MM + o;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < ld; x++) {
        const size_t i = x + (size_t)ld * y;
        double LM, MM;
        level_lm_mm(T, i, cst, &LM, &MM);
        if (init) {
          FMM[i] = MM; FLM[i] = 0.025 * MM;
          if (x >= ld - 2) { FMM[i] = 1.0; FLM[i] = 1.0; }
        }
        double Tn = fmax(FLM[i] * FMM[i], eps);                      /* :256-260 */
        Tn = opftdelta * pow(Tn, powcoeff);
        const double dumfac = lagran_dt / Tn;
        const double epsi = dumfac / (1.0 + dumfac);
        FLM[i] = (epsi * LM + (1.0 - epsi) * FLM[i]);                /* :263-265 */
        FMM[i] = (epsi * MM + (1.0 - epsi) * FMM[i]);
        FLM[i] = fmax(eps, FLM[i]);
        double cs = FLM[i] / FMM[i];                                 /* :267-269 */
        if (x >= ld - 2) cs = 0.0;
        s->Cs_opt2[o + i] = fmax(eps, cs);
        s->Beta[o + i] = 1.0;                                        /* :117 */
      }
    s->Beta_avg[k] = (double)nx * ny;                                /* :270-271, sums of Beta = 1 */
    s->Betaclip_avg[k] = (double)nx * ny;
  }
    for (int t = 0; t < LV_N; t++) free(T[t]);
  }
  if (init) c->flm_init = 1;
}

/* std_dynamic.f90:4-140: one coefficient per level from plane sums of L:M
   and M:M over the (ld, ny) arrays, Delta = (dx dy dz)^(1/3) */
void or_std_dynamic(or_ctx *c, const double *const S[6], const double *u, const double *v,
                    const double *w, double *Cs_opt2) {
  const int ld = c->ld, ny = c->ny, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double delta = pow(c->dx * c->dy * c->dz, 1.0 / 3.0);        /* :28 */
  const double cst = 2.0 * (delta * delta);                          /* :124 */
#pragma omp parallel
  {
    double *T[LV_N];
    for (int t = 0; t < LV_N; t++) T[t] = (double *)calloc(pl, sizeof(double));
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz; k++) {
    level_2delta(c, S, u, v, w, k, T);
    csum_t num = { 0.0, 0.0 }, den = { 0.0, 0.0 };
    for (size_t i = 0; i < pl; i++) {
      double LM, MM;
      level_lm_mm(T, i, cst, &LM, &MM);
      csum_add(&num, LM);
      csum_add(&den, MM);
    }
    Cs_opt2[k] = fmax(0.0, csum_val(&num) / csum_val(&den));         /* :135-138 */
  }
    for (int t = 0; t < LV_N; t++) free(T[t]);
  }
}

/* ------------------------------------------------------------------ */
/* sgs_stag.f90:3-366 (model=5, sgs=.true., molec=.false., 'wall')      */
/* ------------------------------------------------------------------ */
static void sgs_stag_impl(or_ctx *c, const double *const g[9], const double *u,
                          const double *v, const double *w, double *theta,
                          double *const tau[6], or_sgs_state *s, or_scalar_sgs *ss, int model);

void or_sgs_stag(or_ctx *c, const double *const g[9], const double *u,
                 const double *v, const double *w, double *const tau[6],
                 or_sgs_state *s) {
  sgs_stag_impl(c, g, u, v, w, NULL, tau, s, NULL, 5);
}

/* model=3: the plane-averaged scale-dependent model (sgs_stag.f90:126-137) */
void or_sgs_stag_model3(or_ctx *c, const double *const g[9], const double *u,
                        const double *v, const double *w, double *const tau[6],
                        or_sgs_state *s) {
  sgs_stag_impl(c, g, u, v, w, NULL, tau, s, NULL, 3);
}

/* model=7: also u_lag_s accumulation (sgs_stag.f90:64-66), the scalar
   Lagrangian model after lagrange_Sdep (:150-153) and Pr_t=Pr while
   jt_total<500 (:171) */
void or_sgs_stag_scalar(or_ctx *c, const double *const g[9], const double *u,
                        const double *v, const double *w, double *theta,
                        double *const tau[6], or_sgs_state *s, or_scalar_sgs *ss) {
  sgs_stag_impl(c, g, u, v, w, theta, tau, s, ss, 7);
}

void or_sgs_stag_model(or_ctx *c, const double *const g[9], const double *u,
                       const double *v, const double *w, double *const tau[6],
                       or_sgs_state *s, int model) {
  sgs_stag_impl(c, g, u, v, w, NULL, tau, s, NULL, model);
}

static void sgs_stag_impl(or_ctx *c, const double *const g[9], const double *u,
                          const double *v, const double *w, double *theta,
                          double *const tau[6], or_sgs_state *s, or_scalar_sgs *ss, int model) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz;
  double *S[6];
  for (int q = 0; q < 6; q++) S[q] = (double *)calloc(n3, sizeof(double));
  or_calc_sij(c, g, S);
  double *txx = tau[0], *txy = tau[1], *txz = tau[2], *tyy = tau[3], *tyz = tau[4], *tzz = tau[5];
  double *txzp = (double *)malloc(sizeof(double) * pl), *tyzp = (double *)malloc(sizeof(double) * pl);
  memcpy(txzp, txz, sizeof(double) * pl); memcpy(tyzp, tyz, sizeof(double) * pl);
  if (model == 4 || model == 5 || model == 7)                   /* sgs_stag.f90:59-69 */
    for (size_t i = 0; i < n3; i++) { s->u_lag[i] += u[i]; s->v_lag[i] += v[i]; s->w_lag[i] += w[i]; }
  if (ss && model == 7)
    for (size_t i = 0; i < n3; i++) { ss->u_lag_s[i] += u[i]; ss->v_lag_s[i] += v[i]; ss->w_lag_s[i] += w[i]; }
  const int jt = c->jt;
  double *lz = (double *)malloc(sizeof(double) * nz);
  for (int k = 0; k < nz; k++) lz[k] = c->delta;                  /* :110 dynamic models: l = delta */
  if (model == 1) {                                               /* :73-104, lbc_mom = 'wall' */
    const double Co = 0.16, vonk = 0.41;                          /* param.f90:98,123; nnn = 2 */
    for (size_t i = 0; i < n3; i++) s->Cs_opt2[i] = Co * Co;
    for (int k = 0; k < nz; k++) {
      const double zz = k == 0 ? (OR_MUTANT == 12 ? 1.0 : 1.0 - 0.5) * c->dz : (double)k * c->dz;
      const double a = vonk * zz;
      lz[k] = pow(Co * Co * (1.0 / (a * a)) + 1.0 / (c->delta * c->delta), -1.0 / 2.0);
    }
  } else if (jt == 1 && c->P.inilag) {
    for (size_t i = 0; i < n3; i++) s->Cs_opt2[i] = 0.03;
  } else if ((jt >= 1 || c->P.initu) && (jt % c->P.cs_count == 0)) {
    if (model == 3) {
      /* sgs_stag.f90:132 passes (ziko,u,v,w,S11,...,S33) to the 7-argument
         routine: its S11..S33 are u, v, w, S11, S12, S13 */
      const double *Sd[6] = { u, v, w, S[0], S[1], S[2] };
      double *ziko = (double *)calloc(nz, sizeof(double));
      or_scaledep_dynamic(c, Sd, u, v, w, ziko, NULL);
      for (int k = 0; k < nz; k++)
        for (size_t i = 0; i < pl; i++) s->Cs_opt2[pl * k + i] = ziko[k];
      free(ziko);
    } else if (model == 2) {                                      /* :123-126 */
      double *ziko = (double *)calloc(nz, sizeof(double));
      or_std_dynamic(c, (const double *const *)S, u, v, w, ziko);
      for (int k = 0; k < nz; k++)
        for (size_t i = 0; i < pl; i++) s->Cs_opt2[pl * k + i] = ziko[k];
      free(ziko);
    } else if (model == 4) {                                      /* :137-138 */
      or_lagrange_ssim(c, (const double *const *)S, u, v, w, s);
    } else if (model == 6) {                                      /* :143-149 */
      double *ziko = (double *)calloc(nz, sizeof(double));
      or_new_scaledep_dynamic(c, (const double *const *)S, u, v, w, ziko);
      for (int k = 0; k < nz; k++)
        for (size_t i = 0; i < pl; i++) s->Cs_opt2[pl * k + i] = ziko[k];
      free(ziko);
      /* Cs2_Pr(nz-1) of the routine lands in the first nz-1 elements of
         sgs_stag's (ld, ny, nz) Cs2_Pr (sequence association) */
      if (ss) or_scalar_scaledep_dynamic(c, (const double *const *)S, u, v, w, theta, s->Cs_opt2, ss->psi_m,
                                         ss->phi_h, ss->Cs2_Pr, ss->Pr_t);
    } else {
      or_lagrange_sdep(c, (const double *const *)S, u, v, w, s);
      if (ss) or_lagrange_scalar_sdep(c, (const double *const *)S, u, v, w, theta, s, ss);
    }
  }
  if ((model == 5 || model == 6 || model == 7) && c->jt_total < 500) {   /* :165-174 */
    const double cs0 = (double)(0.16f * 0.16f);
    for (size_t i = 0; i < n3; i++) s->Cs_opt2[i] = cs0;
    if (ss) for (size_t i = 0; i < n3; i++) ss->Pr_t[i] = c->S.Pr;
  }
  for (int k = 0; k < nz; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        double Sm = sqrt(2.0 * (S[0][i] * S[0][i] + S[3][i] * S[3][i] + S[5][i] * S[5][i] +
                                2.0 * (S[1][i] * S[1][i] + S[2][i] * S[2][i] + S[4][i] * S[4][i])));
        s->Nu_t[i] = Sm * s->Cs_opt2[i] * (lz[k] * lz[k]);
      }
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      size_t i = x + (size_t)ld * y;
      double cst = s->Nu_t[i];
      txx[i] = -2.0 * (cst + 0.0) * S[0][i];
      txy[i] = -2.0 * (cst + 0.0) * S[1][i];
      tyy[i] = -2.0 * (cst + 0.0) * S[3][i];
      tzz[i] = -2.0 * (cst + 0.0) * S[5][i];
    }
  for (int k = 1; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y, p = i + pl;
        double cst = 0.5 * (s->Nu_t[i] + s->Nu_t[p]);
        txx[i] = -2.0 * (cst + 0.0) * 0.5 * (S[0][i] + S[0][p]);
        txy[i] = -2.0 * (cst + 0.0) * 0.5 * (S[1][i] + S[1][p]);
        tyy[i] = -2.0 * (cst + 0.0) * 0.5 * (S[3][i] + S[3][p]);
        tzz[i] = -2.0 * (cst + 0.0) * 0.5 * (S[5][i] + S[5][p]);
      }
  for (int k = 1; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        double cst = s->Nu_t[i];
        txz[i] = -2.0 * (cst + 0.0) * S[2][i];
        tyz[i] = -2.0 * (cst + 0.0) * S[4][i];
      }
  memcpy(txz, txzp, sizeof(double) * pl); memcpy(tyz, tyzp, sizeof(double) * pl);
  const size_t top = pl * (nz - 1);
  for (size_t i = 0; i < pl; i++) {
    txx[top + i] = BOGUS; txy[top + i] = BOGUS; tyy[top + i] = BOGUS; tzz[top + i] = BOGUS;
    txz[top + i] = 0.0; tyz[top + i] = 0.0;
  }
  free(txzp); free(tyzp); free(lz);
  for (int q = 0; q < 6; q++) free(S[q]);
}

/* ------------------------------------------------------------------ */
/* divstress_uv.f90:1-63, divstress_w.f90:4-86 (non-MPI)               */
/* ------------------------------------------------------------------ */
void or_divstress_uv(or_ctx *c, double *divt, const double *tx, const double *ty, const double *tz) {
  const int nz = c->nz, ld = c->ld, ny = c->ny;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz;
  double *a = (double *)calloc(n3, sizeof(double)), *b = (double *)calloc(n3, sizeof(double));
  double *d = (double *)calloc(n3, sizeof(double));
  or_ddx(c, a, tx, nz);
  or_ddy(c, b, ty, nz);
  or_ddz_w(c, d, tz);
  for (size_t i = 0; i < pl * (nz - 1); i++) divt[i] = a[i] + b[i] + d[i];
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++) { divt[pl * k + ld - 2 + (size_t)ld * y] = 0.0; divt[pl * k + ld - 1 + (size_t)ld * y] = 0.0; }
  for (size_t i = 0; i < pl; i++) divt[pl * (nz - 1) + i] = BOGUS;
  free(a); free(b); free(d);
}

void or_divstress_w(or_ctx *c, double *divt, const double *tx, const double *ty, const double *tz) {
  const int nz = c->nz, ld = c->ld, ny = c->ny, nx = c->nx;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz;
  double *a = (double *)calloc(n3, sizeof(double)), *b = (double *)calloc(n3, sizeof(double));
  double *d = (double *)calloc(n3, sizeof(double));
  or_ddx(c, a, tx, nz);
  or_ddy(c, b, ty, nz);
  or_ddz_uv(c, d, tz);
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) { size_t i = x + (size_t)ld * y; divt[i] = a[i] + b[i]; }
  for (int k = 1; k < nz; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) { size_t i = pl * k + x + (size_t)ld * y; divt[i] = a[i] + b[i] + d[i]; }
  for (int k = 0; k < nz; k++)
    for (int y = 0; y < ny; y++) { divt[pl * k + ld - 2 + (size_t)ld * y] = 0.0; divt[pl * k + ld - 1 + (size_t)ld * y] = 0.0; }
  free(a); free(b); free(d);
}

/* ------------------------------------------------------------------ */
/* tridag_array.f90:1-109 (non-MPI)                                    */
/* ------------------------------------------------------------------ */
int or_tridag_array(or_ctx *c, const double *a, const double *b, const double *cc, const double *r, double *u) {
  const int lh = c->lh, ny = c->ny, n = c->nz + 1;
  const size_t pl = (size_t)lh * ny;
  double *bet = (double *)malloc(sizeof(double) * pl);
  double *gam = (double *)calloc(pl * n, sizeof(double));
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < lh - 1; x++)
      if (b[x + (size_t)lh * y] == 0.0) { free(bet); free(gam); return 1; }
  for (size_t i = 0; i < pl; i++) {
    bet[i] = b[i];
    CRE(u, i) = CRE(r, i) / bet[i];
    CIM(u, i) = CIM(r, i) / bet[i];
  }
  for (int j = 1; j < n; j++)
    for (int y = 0; y < ny; y++) {
      if (y == ny / 2) continue;
      for (int x = 0; x < lh - 1; x++) {
        if (x == 0 && y == 0) continue;
        size_t m = x + (size_t)lh * y, i = pl * j + m, ip = i - pl;
        gam[i] = cc[ip] / bet[m];
        bet[m] = b[i] - a[i] * gam[i];
        if (bet[m] == 0.0) { free(bet); free(gam); return 2; }
        CRE(u, i) = (CRE(r, i) - a[i] * CRE(u, ip)) / bet[m];
        CIM(u, i) = (CIM(r, i) - a[i] * CIM(u, ip)) / bet[m];
      }
    }
  for (int j = n - 2; j >= 0; j--)
    for (int y = 0; y < ny; y++) {
      if (y == ny / 2) continue;
      for (int x = 0; x < lh - 1; x++) {
        if (x == 0 && y == 0) continue;
        size_t m = x + (size_t)lh * y, i = pl * j + m, in = i + pl;
        CRE(u, i) = CRE(u, i) - gam[in] * CRE(u, in);
        CIM(u, i) = CIM(u, i) - gam[in] * CIM(u, in);
      }
    }
  free(bet); free(gam);
  return 0;
}

/* ------------------------------------------------------------------ */
/* press_stag_array.f90:1-294 (non-MPI, lbc=1)                          */
/* ------------------------------------------------------------------ */
void or_press_stag_array(or_ctx *c, const double *u, const double *v, const double *w,
                         const double *divtz, double *p, double *dpdx, double *dpdy) {
  const int nz = c->nz, ld = c->ld, ny = c->ny, lh = c->lh;
  const size_t pl = (size_t)ld * ny, cpl = (size_t)lh * ny;
  const double dz = c->dz, dt = c->P.dt;
  const double cst = 1.0 / (c->nx * c->ny);
  double *Hx = (double *)calloc(pl * nz, sizeof(double));
  double *Hy = (double *)calloc(pl * nz, sizeof(double));
  double *Hz = (double *)calloc(pl * nz, sizeof(double));
  double *topw = (double *)malloc(sizeof(double) * pl), *botw = (double *)malloc(sizeof(double) * pl);
  const int n = nz + 1;
  double *a = (double *)calloc(cpl * n, sizeof(double)), *b = (double *)calloc(cpl * n, sizeof(double));
  double *cc = (double *)calloc(cpl * n, sizeof(double)), *R = (double *)calloc(2 * cpl * n, sizeof(double));
  /* p_hat(:,:,0)=0 */
  for (size_t i = 0; i < pl; i++) p[i] = 0.0;
  /* jz = 1..nz-1 -> array level k = jz-1 */
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nz - 1; k++) {
    for (size_t i = 0; i < pl; i++) {
      Hx[pl * k + i] = cst / TADV1 * (u[pl * k + i] / dt);
      Hy[pl * k + i] = cst / TADV1 * (v[pl * k + i] / dt);
      Hz[pl * k + i] = cst / TADV1 * (w[pl * k + i] / dt);
    }
    rfft2_fwd(&c->sm, Hx + pl * k); rfft2_fwd(&c->sm, Hy + pl * k); rfft2_fwd(&c->sm, Hz + pl * k);
  }
  /* H_x(:,:,nz) = H_y(:,:,nz) = BOGUS (complex), H_z(:,:,nz) = 0 */
  for (size_t i = 0; i < cpl; i++) {
    CRE(Hx + pl * (nz - 1), i) = BOGUS; CIM(Hx + pl * (nz - 1), i) = 0.0;
    CRE(Hy + pl * (nz - 1), i) = BOGUS; CIM(Hy + pl * (nz - 1), i) = 0.0;
    CRE(Hz + pl * (nz - 1), i) = 0.0; CIM(Hz + pl * (nz - 1), i) = 0.0;
  }
  const double wsgn = OR_MUTANT == 5 ? -1.0 : 1.0;
  for (size_t i = 0; i < pl; i++) { botw[i] = wsgn * cst * divtz[i]; topw[i] = wsgn * cst * divtz[pl * (nz - 1) + i]; }
  rfft2_fwd(&c->sm, botw); rfft2_fwd(&c->sm, topw);
  for (int k = 0; k < nz - 1; k++) { kill_oddballs(c, Hx + pl * k); kill_oddballs(c, Hy + pl * k); kill_oddballs(c, Hz + pl * k); }
  kill_oddballs(c, topw); kill_oddballs(c, botw);
  /* rows: RHS_col index j=1..nz+1 -> array j-1 */
  for (size_t m = 0; m < cpl; m++) {
    a[m] = BOGUS; b[m] = -1.0; cc[m] = 1.0;
    CRE(R, m) = -dz * CRE(botw, m); CIM(R, m) = -dz * CIM(botw, m);
    size_t t = cpl * nz + m;
    a[t] = -1.0; b[t] = 1.0; cc[t] = BOGUS;
    CRE(R, t) = -CRE(topw, m) * dz; CIM(R, t) = -CIM(topw, m) * dz;
  }
  for (int jz = 2; jz <= nz; jz++)
    for (int y = 0; y < ny; y++) {
      if (y == ny / 2) continue;
      for (int x = 0; x < lh - 1; x++) {
        if (x == 0 && y == 0) continue;
        size_t m = x + (size_t)lh * y, t = cpl * (jz - 1) + m;
        double kx = c->kx[m], ky = c->ky[m];
        a[t] = 1.0 / (dz * dz);
        b[t] = -(kx * kx + ky * ky + 2.0 / (dz * dz));
        cc[t] = 1.0 / (dz * dz);
        const double *hxm = Hx + pl * (jz - 2), *hym = Hy + pl * (jz - 2);
        const double *hzm = Hz + pl * (jz - 2), *hz = Hz + pl * (jz - 1);
        /* eye*(kx*Hx + ky*Hy) + (Hz(jz) - Hz(jz-1))/dz */
        double sr = kx * CRE(hxm, m) + ky * CRE(hym, m), si = kx * CIM(hxm, m) + ky * CIM(hym, m);
        double er = 0.0 * sr - 1.0 * si, ei = 0.0 * si + 1.0 * sr;
        CRE(R, t) = er + (CRE(hz, m) - CRE(hzm, m)) / dz;
        CIM(R, t) = ei + (CIM(hz, m) - CIM(hzm, m)) / dz;
      }
    }
  /* p_hat levels 0..nz = u(1..n): p array stores complex planes interleaved */
  double *U = (double *)calloc(2 * cpl * n, sizeof(double));
  /* tridag writes only solved modes for j>=2; other modes keep p_hat's prior
     content, which is zeroed below (Nyquist) or set by the zero-mode chain */
  int err = or_tridag_array(c, a, b, cc, R, U);
  if (err) fprintf(stderr, "or_press_stag_array: tridag failed (%d)\n", err);
  for (int j = 0; j < n; j++) {
    double *P = p + pl * j;
    for (size_t m = 0; m < cpl; m++) { CRE(P, m) = CRE(U + 2 * cpl * j, m); CIM(P, m) = CIM(U + 2 * cpl * j, m); }
  }
  /* zero mode: press_stag_array.f90:238-247 */
  CRE(p, 0) = 0.0; CIM(p, 0) = 0.0;
  CRE(p + pl, 0) = CRE(p, 0) - dz * CRE(botw, 0);
  CIM(p + pl, 0) = CIM(p, 0) - dz * CIM(botw, 0);
  for (int jz = 2; jz <= nz; jz++) {
    const double *hz = Hz + pl * (jz - 1);
    CRE(p + pl * jz, 0) = CRE(p + pl * (jz - 1), 0) + CRE(hz, 0) * dz;
    CIM(p + pl * jz, 0) = CIM(p + pl * (jz - 1), 0) + CIM(hz, 0) * dz;
  }
  for (int j = 0; j < n; j++) kill_oddballs(c, p + pl * j);
  rfft2_inv(&c->sm, p);
#pragma omp parallel for schedule(dynamic)
  for (int k = 1; k <= nz - 1; k++) {
    double *P = p + pl * k;
    mul_ik(c, c->kx, P, dpdx + pl * (k - 1));
    mul_ik(c, c->ky, P, dpdy + pl * (k - 1));
    rfft2_inv(&c->sm, dpdx + pl * (k - 1));
    rfft2_inv(&c->sm, dpdy + pl * (k - 1));
    rfft2_inv(&c->sm, P);
  }
  for (size_t m = 0; m < cpl; m++) {
    CRE(dpdx + pl * (nz - 1), m) = BOGUS; CIM(dpdx + pl * (nz - 1), m) = 0.0;
    CRE(dpdy + pl * (nz - 1), m) = BOGUS; CIM(dpdy + pl * (nz - 1), m) = 0.0;
    CRE(p + pl * nz, m) = BOGUS; CIM(p + pl * nz, m) = 0.0;
  }
  free(Hx); free(Hy); free(Hz); free(topw); free(botw); free(a); free(b); free(cc); free(R); free(U);
}

/* ------------------------------------------------------------------ */
/* forcing.f90:128-219 project (no buildings/inflow, 'wall')           */
/* ------------------------------------------------------------------ */
void or_project(or_ctx *c, double *u, double *v, double *w, const double *dpdx,
                const double *dpdy, const double *dpdz) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double dt = c->P.dt;
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        double R = -TADV1 * dpdx[i];
        u[i] = (u[i] + dt * (R + 0.0));
        R = -TADV1 * dpdy[i];
        v[i] = (v[i] + dt * (R + 0.0));
      }
  for (int k = 1; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        double R = -TADV1 * dpdz[i];
        w[i] = (w[i] + dt * (R + 0.0));
      }
  for (size_t i = 0; i < pl; i++) {
    u[pl * (nz - 1) + i] = u[pl * (nz - 2) + i];
    v[pl * (nz - 1) + i] = v[pl * (nz - 2) + i];
    w[pl * (nz - 1) + i] = 0.0;
    w[i] = 0.0;
  }
}

/* rmsdiv.f90:1-113 (norm_magdu=.false.) */
double or_rmsdiv(or_ctx *c, const double *du, const double *dv, const double *dw) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  double rms = 0.0;
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        rms = rms + fabs(du[i] + dv[i] + dw[i]);
      }
  return rms / ((double)nx * ny * (nz - 1));
}

/* scalars_module2.f90:22-70 (non-MPI).  out = {CFLu, CFLv, CFLw, visc_stab};
   the non-MPI reference leaves v_res_max/w_res_max undefined, here they are
   the per-component maxima of |v| and |w|. */
void or_timestep_conditions(or_ctx *c, const double *u, const double *v, const double *w,
                            const double *Nu_t, double out[4]) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  double umax = -1.0, vmax = 0.0, wmax = 0.0, numax = -INFINITY;
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        double m = u[i] * u[i] + v[i] * v[i] + w[i] * w[i];
        if (m > umax) umax = m;
        if (fabs(v[i]) > vmax) vmax = fabs(v[i]);
        if (fabs(w[i]) > wmax) wmax = fabs(w[i]);
        if (Nu_t[i] > numax) numax = Nu_t[i];
      }
  double ures = sqrt(umax), delta = fmin(c->dx, c->dy), dt = c->P.dt;
  out[0] = ures * dt / c->dx;
  out[1] = vmax * dt / c->dy;
  out[2] = wmax * dt / c->dz;
  out[3] = dt * numax / (delta * delta);
}

/* ------------------------------------------------------------------ */
/* state + whole step (main.f90:202-415, S_FLAG=.false., model=5)      */
/* ------------------------------------------------------------------ */
static double *al3(const or_ctx *c, int nlev) { return (double *)calloc((size_t)c->ld * c->ny * nlev, sizeof(double)); }

or_state *or_state_create(or_ctx *c) {
  or_state *s = (or_state *)calloc(1, sizeof(or_state));
  const int nz = c->nz;
  s->u = al3(c, nz); s->v = al3(c, nz); s->w = al3(c, nz);
  for (int q = 0; q < 9; q++) s->grad[q] = al3(c, nz);
  s->RHSx = al3(c, nz); s->RHSy = al3(c, nz); s->RHSz = al3(c, nz);
  s->RHSx_f = al3(c, nz); s->RHSy_f = al3(c, nz); s->RHSz_f = al3(c, nz);
  for (int q = 0; q < 6; q++) s->tau[q] = al3(c, nz);
  s->divtx = al3(c, nz); s->divty = al3(c, nz); s->divtz = al3(c, nz);
  s->p = al3(c, nz + 1);
  s->dpdx = al3(c, nz); s->dpdy = al3(c, nz); s->dpdz = al3(c, nz);
  size_t n2 = (size_t)c->nx * c->ny;
  s->psi_m = (double *)calloc(n2, sizeof(double)); s->psi_m0 = (double *)calloc(n2, sizeof(double));
  s->phi_m = (double *)calloc(n2, sizeof(double)); s->ustar_avg = (double *)calloc(n2, sizeof(double));
  for (size_t i = 0; i < n2; i++) s->phi_m[i] = 1.0;
  or_sgs_state *g = &s->sgs;
  g->F_LM = al3(c, nz); g->F_MM = al3(c, nz); g->F_QN = al3(c, nz); g->F_NN = al3(c, nz);
  g->Beta = al3(c, nz); g->Cs_opt2 = al3(c, nz); g->Nu_t = al3(c, nz);
  g->u_lag = al3(c, nz); g->v_lag = al3(c, nz); g->w_lag = al3(c, nz);
  g->TnLMMM = al3(c, nz); g->TnQNNN = al3(c, nz);
  g->L11t = al3(c, nz); g->L22t = al3(c, nz); g->L33t = al3(c, nz);
  g->Q11t = al3(c, nz); g->Q22t = al3(c, nz); g->Q33t = al3(c, nz);
  g->Beta_avg = (double *)calloc(nz, sizeof(double)); g->Betaclip_avg = (double *)calloc(nz, sizeof(double));
  s->theta = al3(c, nz); s->qmix = al3(c, nz);
  s->dTdx = al3(c, nz); s->dTdy = al3(c, nz); s->dTdz = al3(c, nz);
  s->dqdx = al3(c, nz); s->dqdy = al3(c, nz); s->dqdz = al3(c, nz);
  s->RHS_T = al3(c, nz); s->RHS_Tf = al3(c, nz); s->RHS_q = al3(c, nz); s->RHS_qf = al3(c, nz);
  s->beta_scal = al3(c, nz);
  s->sgs_t3 = al3(c, 1); s->sgs_q3 = al3(c, 1);
  s->psi_h = (double *)calloc(n2, sizeof(double)); s->psi_h0 = (double *)calloc(n2, sizeof(double));
  s->phi_h = (double *)calloc(n2, sizeof(double));
  for (size_t i = 0; i < n2; i++) s->phi_h[i] = 1.0;
  or_scalar_sgs *h = &s->ssgs;
  h->G_LM = al3(c, nz); h->G_MM = al3(c, nz); h->G_QN = al3(c, nz); h->G_NN = al3(c, nz);
  h->u_lag_s = al3(c, nz); h->v_lag_s = al3(c, nz); h->w_lag_s = al3(c, nz);
  h->Pr_t = al3(c, nz); h->Cs2_Pr = al3(c, nz);
  h->phi_h = s->phi_h; h->psi_m = s->psi_m;
  return s;
}

void or_state_destroy(or_state *s) {
  if (!s) return;
  free(s->u); free(s->v); free(s->w);
  for (int q = 0; q < 9; q++) free(s->grad[q]);
  free(s->RHSx); free(s->RHSy); free(s->RHSz); free(s->RHSx_f); free(s->RHSy_f); free(s->RHSz_f);
  for (int q = 0; q < 6; q++) free(s->tau[q]);
  free(s->divtx); free(s->divty); free(s->divtz); free(s->p); free(s->dpdx); free(s->dpdy); free(s->dpdz);
  free(s->psi_m); free(s->psi_m0); free(s->phi_m); free(s->ustar_avg);
  or_sgs_state *g = &s->sgs;
  free(g->F_LM); free(g->F_MM); free(g->F_QN); free(g->F_NN); free(g->Beta); free(g->Cs_opt2); free(g->Nu_t);
  free(g->u_lag); free(g->v_lag); free(g->w_lag); free(g->TnLMMM); free(g->TnQNNN);
  free(g->L11t); free(g->L22t); free(g->L33t); free(g->Q11t); free(g->Q22t); free(g->Q33t);
  free(g->Beta_avg); free(g->Betaclip_avg);
  free(s->theta); free(s->qmix); free(s->dTdx); free(s->dTdy); free(s->dTdz);
  free(s->dqdx); free(s->dqdy); free(s->dqdz);
  free(s->RHS_T); free(s->RHS_Tf); free(s->RHS_q); free(s->RHS_qf); free(s->beta_scal);
  free(s->sgs_t3); free(s->sgs_q3); free(s->psi_h); free(s->psi_h0); free(s->phi_h);
  or_scalar_sgs *h = &s->ssgs;
  free(h->G_LM); free(h->G_MM); free(h->G_QN); free(h->G_NN);
  free(h->u_lag_s); free(h->v_lag_s); free(h->w_lag_s); free(h->Pr_t); free(h->Cs2_Pr);
  free(s);
}

/* OpenMP threads the oracle's 2-D transforms use (1 without OpenMP) */
int or_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

double or_uniform(uint64_t seed, uint64_t index) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + index + 0x632BE59BD9B4E019ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void or_state_init_ic(or_ctx *c, or_state *s, uint64_t seed, double amp, double ustar, double kappa_ic) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  double *f3[3] = { s->u, s->v, s->w };
  for (int k = 0; k < nz - 1; k++) {
    double z = (k + 0.5) * c->dz;
    double ulog = (ustar / kappa_ic) * log(z / c->P.z0);
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = pl * k + x + (size_t)ld * y;
        s->u[i] = ulog; s->v[i] = 0.0; s->w[i] = 0.0;
        if (z < 0.5)
          for (int q = 0; q < 3; q++) {
            uint64_t idx = (((uint64_t)q * nz + k) * ny + y) * nx + x;
            f3[q][i] += amp * ustar * (2.0 * or_uniform(seed, idx) - 1.0);
          }
      }
  }
  for (size_t i = 0; i < pl; i++) {
    s->w[i] = 0.0;
    s->u[pl * (nz - 1) + i] = s->u[pl * (nz - 2) + i];
    s->v[pl * (nz - 1) + i] = s->v[pl * (nz - 2) + i];
    s->w[pl * (nz - 1) + i] = 0.0;
  }
  s->jt = 0; s->jt_total = 0;
  c->flm_init = 0; c->fqn_init = 0;
  c->glm_init = 0; c->gqn_init = 0;
}

static void theta_q_all_in_one(or_ctx *c, or_state *s, int which);
static void momentum_forcing(or_ctx *c, or_state *s);

void or_step(or_ctx *c, or_state *s) {
  const int nz = c->nz, nx = c->nx, ny = c->ny, ld = c->ld;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz, nv = pl * (nz - 1);
  s->jt += 1; s->jt_total += 1;
  c->jt = s->jt; c->jt_total = s->jt_total;
  memcpy(s->RHSx_f, s->RHSx, n3 * sizeof(double));
  memcpy(s->RHSy_f, s->RHSy, n3 * sizeof(double));
  memcpy(s->RHSz_f, s->RHSz, n3 * sizeof(double));
  const int scal = c->S.enable;
  if (scal) {
    or_obukhov(c, s->u, s->v, s->theta, s->qmix, s->sgs_t3, s->sgs_q3, s->psi_m, s->psi_m0,
               s->psi_h, s->psi_h0, s->phi_m, s->phi_h, s->ustar_avg);
  } else {
    /* obukhov with theta = 0: L is NaN, the neutral branch gives psi=0, phi=1
       (scalars_module.f90:642-649) */
    for (size_t i = 0; i < (size_t)nx * ny; i++) { s->psi_m[i] = 0.0; s->psi_m0[i] = 0.0; s->phi_m[i] = 1.0; }
  }
  double **g = s->grad;
  or_filt_da(c, s->u, g[0], g[1], nz);
  or_filt_da(c, s->v, g[3], g[4], nz);
  or_filt_da(c, s->w, g[6], g[7], nz);
  or_ddz_uv(c, g[2], s->u);
  or_ddz_uv(c, g[5], s->v);
  or_ddz_w(c, g[8], s->w);
  or_wallstress(c, s->u, s->v, g[2], g[5], s->tau[2], s->tau[4], s->psi_m, s->psi_m0, s->phi_m, s->ustar_avg);
  if (scal && c->S.model == 7)
    or_sgs_stag_scalar(c, (const double *const *)g, s->u, s->v, s->w, s->theta, s->tau, &s->sgs, &s->ssgs);
  else if (scal && c->sgs_model == 6)  /* model = 6: scalar_scaledep_dynamic gives Pr_t (sgs_stag.f90:149) */
    sgs_stag_impl(c, (const double *const *)g, s->u, s->v, s->w, s->theta, s->tau, &s->sgs, &s->ssgs, 6);
  else
    or_sgs_stag_model(c, (const double *const *)g, s->u, s->v, s->w, s->tau, &s->sgs, c->sgs_model);
  if (scal) {                          /* main.f90:273-276 */
    theta_q_all_in_one(c, s, 0);
    theta_q_all_in_one(c, s, 1);
  } else {
    memset(s->beta_scal, 0, n3 * sizeof(double));
  }
  or_divstress_uv(c, s->divtx, s->tau[0], s->tau[1], s->tau[2]);
  or_divstress_uv(c, s->divty, s->tau[1], s->tau[3], s->tau[4]);
  or_divstress_w(c, s->divtz, s->tau[2], s->tau[4], s->tau[5]);
  or_convec(c, s->u, s->v, s->w, g[1], g[2], g[3], g[5], g[6], g[7], s->RHSx, s->RHSy, s->RHSz);
  for (size_t i = 0; i < nv; i++) {
    s->RHSx[i] = -s->RHSx[i] - s->divtx[i];
    s->RHSy[i] = -s->RHSy[i] - s->divty[i];
    s->RHSz[i] = -s->RHSz[i] - s->divtz[i];
  }
  if (scal && !c->S.passive)           /* main.f90:312-317 */
    for (size_t i = 0; i < nv; i++) s->RHSz[i] = s->RHSz[i] + s->beta_scal[i];
  momentum_forcing(c, s);
  const double force = c->P.force;
  if (s->jt == 1 && !c->P.initu && OR_MUTANT != 8) {
    memcpy(s->RHSx_f, s->RHSx, n3 * sizeof(double));
    memcpy(s->RHSy_f, s->RHSy, n3 * sizeof(double));
    memcpy(s->RHSz_f, s->RHSz, n3 * sizeof(double));
  }
  const double dt = c->P.dt;
  for (size_t i = 0; i < nv; i++) {
    const double ta1 = OR_MUTANT == 3 ? TADV2 : TADV1, ta2 = OR_MUTANT == 3 ? TADV1 : TADV2;
    const double fu = OR_MUTANT == 4 ? 0.0 : force, fv = OR_MUTANT == 4 ? force : 0.0;
    s->u[i] = s->u[i] + dt * (ta1 * s->RHSx[i] + ta2 * s->RHSx_f[i] + fu);
    s->v[i] = s->v[i] + dt * (ta1 * s->RHSy[i] + ta2 * s->RHSy_f[i] + fv);
    s->w[i] = s->w[i] + dt * (ta1 * s->RHSz[i] + ta2 * s->RHSz_f[i]);
  }
  for (size_t i = 0; i < pl; i++) { s->u[nv + i] = BOGUS; s->v[nv + i] = BOGUS; s->w[nv + i] = BOGUS; }
  or_press_stag_array(c, s->u, s->v, s->w, s->divtz, s->p, s->dpdx, s->dpdy);
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = x + (size_t)ld * y;
        s->dpdz[pl * k + i] = (s->p[pl * (k + 1) + i] - s->p[pl * k + i]) / c->dz;
      }
  for (size_t i = 0; i < pl; i++) s->dpdz[nv + i] = BOGUS;
  for (size_t i = 0; i < nv; i++) {
    s->RHSx[i] = s->RHSx[i] - s->dpdx[i];
    s->RHSy[i] = s->RHSy[i] - s->dpdy[i];
    s->RHSz[i] = s->RHSz[i] - s->dpdz[i];
  }
  or_project(c, s->u, s->v, s->w, s->dpdx, s->dpdy, s->dpdz);
}

void or_set_time(or_ctx *c, int jt, int jt_total) { c->jt = jt; c->jt_total = jt_total; }

/* param.f90:220 coriol, main.f90:131-138 ug, vg; topbc.f90:14-74 setsponge
   (non-MPI, nproc = 1), called when ubc == 1 (main.f90:172-185) */
void or_set_forcing(or_ctx *c, const or_forcing *f) {
  const int nz = c->nz;
  c->F = *f;
  c->coriol = (double)1.3962634E-04f * f->z_i / f->u_star;   /* a default-real literal */
  c->cug = c->coriol * (f->ug0 / f->u_star);
  c->cvg = c->coriol * (f->vg0 / f->u_star);
  double *sp = c->sponge;
  for (int k = 0; k <= nz; k++) sp[k] = 0.0;
  if (f->ubc != 1) return;
  const double z_i = f->z_i, L_z = c->P.Lz * z_i, dz = c->dz;
  if (f->damping_method == 2) {
    const double z_d = 0.75 * (L_z * 1.0), cfrdmp = 3.9;
    for (int k = 1; k <= nz - 1; k++) {
      const double z_local = ((double)k - 0.5) * dz * z_i;
      if (z_local >= z_d && z_local <= L_z * 1.0)
        sp[k] = cfrdmp * 0.5 * (1.0 - cos(PI * (z_local - z_d) / (L_z - z_d)));
      else
        sp[k] = 0.0;
    }
    sp[nz] = sp[nz - 1];
  } else if (f->damping_method == 1) {
    /* nz - 0.75*nz + 1 and 0.75*nz + 1 are default-real expressions; the
       loop bound is truncated to an integer */
    const float den = (float)nz - 0.75f * (float)nz + 1.0f;
    const double factor = 9.0 / (double)den;
    sp[nz] = z_i / (50.0 * f->u_star);
    const int kend = (int)(0.75f * (float)nz + 1.0f);
    for (int k = nz - 1; k >= kend; k--) sp[k] = sp[k + 1] / pow(5.0, factor);
  }
}

void or_get_sponge(const or_ctx *c, double *sp) {
  for (int k = 0; k < c->nz; k++) sp[k] = c->sponge[k + 1];
}

/* plane mean of level k over the nx*ny points (compensated sum) */
static double level_mean(const or_ctx *c, const double *f, int k) {
  csum_t a = { 0.0, 0.0 };
  const size_t pl = (size_t)c->ld * c->ny;
  for (int y = 0; y < c->ny; y++)
    for (int x = 0; x < c->nx; x++) csum_add(&a, f[pl * k + x + (size_t)c->ld * y]);
  return csum_val(&a) / (double)(c->nx * c->ny);
}

/* main.f90:319-348: Coriolis on levels 1..nz-1 (whole rows), then the
   sponge on the nx*ny points of each level */
static void momentum_forcing(or_ctx *c, or_state *s) {
  const int nz = c->nz, nx = c->nx, ny = c->ny, ld = c->ld;
  const size_t pl = (size_t)ld * ny, nv = pl * (nz - 1);
  if (c->F.coriolis) {
    const double sy = OR_MUTANT == 9 ? -1.0 : 1.0;
    for (size_t i = 0; i < nv; i++) {
      s->RHSx[i] = s->RHSx[i] + c->coriol * s->v[i] - c->cvg;
      s->RHSy[i] = s->RHSy[i] - sy * c->coriol * s->u[i] + sy * c->cug;
    }
  }
  if (c->F.ubc != 1) return;
  const double *sp = c->sponge;
  for (int k = 0; k < nz - 1; k++) {
    const int jz = k + 1;
    const double cxy = 0.5 * (sp[jz] + sp[OR_MUTANT == 10 ? jz : jz + 1]);
    const double mu = level_mean(c, s->u, k), mv = level_mean(c, s->v, k), mw = level_mean(c, s->w, k);
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = pl * k + x + (size_t)ld * y;
        if (c->F.damping_method == 2) {
          s->RHSx[i] = s->RHSx[i] - cxy * (s->u[i] - mu);
          s->RHSy[i] = s->RHSy[i] - cxy * (s->v[i] - mv);
          s->RHSz[i] = s->RHSz[i] - 0.5 * sp[jz] * (s->w[i] - mw);
        } else if (c->F.damping_method == 1) {
          s->RHSz[i] = s->RHSz[i] - sp[jz] * s->w[i];
        }
      }
  }
}
void or_set_sgs_model(or_ctx *c, int model) { c->sgs_model = model; }
void or_set_lag_init(or_ctx *c, int flm_init, int fqn_init) { c->flm_init = flm_init; c->fqn_init = fqn_init; }

/* ================================================================== */
/* Scalars: potential temperature (theta) and humidity (q), lbc=1     */
/* ================================================================== */
void or_set_scalars(or_ctx *c, const or_scalar_params *p, const double *wt_s, const double *wq_s) {
  const size_t n2 = (size_t)c->nx * c->ny;
  c->S = *p;
  if (!c->wt_s) c->wt_s = (double *)calloc(n2, sizeof(double));
  if (!c->wq_s) c->wq_s = (double *)calloc(n2, sizeof(double));
  for (size_t i = 0; i < n2; i++) { c->wt_s[i] = wt_s ? wt_s[i] : 0.0; c->wq_s[i] = wq_s ? wq_s[i] : 0.0; }
}

void or_set_scalar_lag_init(or_ctx *c, int glm_init, int gqn_init) { c->glm_init = glm_init; c->gqn_init = gqn_init; }

/* psi_m of the unstable Businger-Dyer form, scalars_module.f90:621-622 */
static double mo_psi_m_unstable(double x) {
  return 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + PI / 2.0;
}

/* obukhov, scalars_module.f90:508-690 (patch_flag=1) */
void or_obukhov(or_ctx *c, const double *u, const double *v, const double *theta,
                const double *qmix, double *sgs_t3, double *sgs_q3, double *psi_m,
                double *psi_m0, double *psi_h, double *psi_h0, double *phi_m,
                double *phi_h, double *ustar_avg) {
  const int nx = c->nx, ny = c->ny, ld = c->ld;
  const size_t pl = (size_t)ld * ny, n2 = (size_t)nx * ny;
  const or_scalar_params *S = &c->S;
  const double dz = c->dz, z0 = c->P.z0, z0s = c->P.z0 / 7.39;
  if (c->jt == 1) {                                         /* :521-538, SCAL_init = 1 */
    for (size_t i = 0; i < n2; i++) { psi_m[i] = 0.0; psi_m0[i] = 0.0; psi_h[i] = 0.0; psi_h0[i] = 0.0; }
    if (!S->initsc)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          sgs_t3[x + (size_t)ld * y] = c->wt_s[x + (size_t)nx * y] / (S->u_star * S->T_scale);
          sgs_q3[x + (size_t)ld * y] = c->wq_s[x + (size_t)nx * y] / (S->u_star * S->q_scale);
        }
  }
  const double g_ = (S->g * S->z_i) / (S->u_star * S->u_star);                 /* :541 */
  csum_t tsum = { 0.0, 0.0 };
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) csum_add(&tsum, theta[x + (size_t)ld * y]);
  const double theta_ = csum_val(&tsum) / (double)(float)(nx * ny);             /* :556 */
  double *u1 = (double *)malloc(sizeof(double) * pl), *v1 = (double *)malloc(sizeof(double) * pl);
  memcpy(u1, u, sizeof(double) * pl); memcpy(v1, v, sizeof(double) * pl);
  or_test_filter(c, u1, c->Gtt);                                                /* :560-563 */
  or_test_filter(c, v1, c->Gtt);
  for (int y = 0; y < ny; y++)                                                  /* :565-573 */
    for (int x = 0; x < nx; x++) {
      const size_t i = x + (size_t)ld * y, j = x + (size_t)nx * y;
      const double ua = sqrt(u1[i] * u1[i] + v1[i] * v1[i]);
      ustar_avg[j] = ua * VONK / (log(0.5 * dz / z0) - psi_m[j] + psi_m0[j]);
    }
  for (int y = 0; y < ny; y++)                                                  /* :577-657 */
    for (int x = 0; x < nx; x++) {
      const size_t i = x + (size_t)ld * y, j = x + (size_t)nx * y;
      const double wt = sgs_t3[i];
      const double wtv = sgs_t3[i] + 0.61 * theta[i] * (0.001 * S->q_scale * sgs_q3[i]);
      const double us = ustar_avg[j];
      const double L = -(us * us * us) / (VONK * (g_ / theta_) * wtv);
      double zeta = 0.5 * dz / L, zeta0 = z0 / L, zeta0s = z0s / L;
      if (zeta < -30.0) zeta = -30.0;
      if (zeta > 2.0) zeta = 2.0;
      if (zeta0 < -30.0) zeta0 = -30.0;
      if (zeta0 > 2.0) zeta0 = 2.0;
      if (zeta0s < -30.0) zeta0s = -30.0;
      if (zeta0s > 2.0) zeta0s = 2.0;
      if (L < 0.0 && wt != 0.0) {
        const double xx = pow(1.0 - 16.0 * zeta, 0.25);
        psi_m[j] = mo_psi_m_unstable(xx);
        psi_h[j] = 2.0 * log((1.0 + xx * xx) / 2.0);
        phi_m[j] = 1.0 / xx;
        phi_h[j] = 1.0 / (xx * xx);
        const double x0 = pow(1.0 - 16.0 * zeta0, 0.25), x0s = pow(1.0 - 16.0 * zeta0s, 0.25);
        psi_m0[j] = mo_psi_m_unstable(x0);
        psi_h0[j] = 2.0 * log((1.0 + x0s * x0s) / 2.0);
      } else if (L > 0.0 && wt != 0.0) {
        phi_m[j] = 1.0 + 5.0 * zeta;
        phi_h[j] = 1.0 + 5.0 * zeta;
        psi_m[j] = -1.0 * 5.0 * zeta;
        psi_h[j] = -1.0 * 5.0 * zeta;
        psi_m0[j] = -1.0 * 5.0 * zeta0;
        psi_h0[j] = -1.0 * 5.0 * zeta0s;
      } else {
        psi_m[j] = 0.0; psi_m0[j] = 0.0; psi_h[j] = 0.0; psi_h0[j] = 0.0;
        phi_m[j] = 1.0; phi_h[j] = 1.0;
      }
    }
  free(u1); free(v1);
}

/* calcbeta, scalars_module.f90:175-228 (non-MPI: k = 2..nz-1) */
void or_calcbeta(or_ctx *c, const double *theta, const double *qmix, double *beta) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const or_scalar_params *S = &c->S;
  const double g_hat = S->g * (S->z_i / (S->u_star * S->u_star));
  memset(beta, 0, sizeof(double) * pl * nz);
  double *bar = (double *)calloc(nz, sizeof(double));
  for (int k = 0; k < nz; k++) {
    csum_t a = { 0.0, 0.0 };
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = pl * k + x + (size_t)ld * y;
        csum_add(&a, theta[i] * (1.0 + 0.61 * 0.001 * qmix[i] * S->q_scale));
      }
    bar[k] = csum_val(&a) / (double)(nx * ny);
  }
  for (int k = 1; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = pl * k + x + (size_t)ld * y;
        const double above = (theta[i] * (1.0 + 0.61 * 0.001 * qmix[i] * S->q_scale) - bar[k]) / bar[k];
        const double below = (theta[i - pl] * (1.0 + 0.61 * 0.001 * qmix[i - pl] * S->q_scale) - bar[k - 1]) / bar[k - 1];
        beta[i] = g_hat * (above + below) / 2.0;
      }
  free(bar);
}

/* surface flux of scalar `which` (lbc=1): scalars_module.f90:259-266,
   scalars_module_q.f90:139-145.  The reference applies it only while
   scalar(1,1,1) < 2 and leaves it undefined otherwise; here it always applies. */
static double surf_flux_of(const or_ctx *c, int which, size_t j) {
  const or_scalar_params *S = &c->S;
  return which == 0 ? c->wt_s[j] / (S->T_scale * S->u_star) : c->wq_s[j] / S->q_scale / S->u_star;
}

/* scalar_RHS_calc, scalars_module.f90:232-454 / scalars_module_q.f90:114-330.
   Level nz of RHS is not defined by the reference (temp(nz) is never set);
   here it is 0. */
void or_scalar_rhs_calc(or_ctx *c, int which, const double *dsdx, const double *dsdy,
                        double *dsdz, const double *u, const double *v, const double *w,
                        const double *Nu_t, const double *Pr_, const double *ustar_avg,
                        const double *phi_h, double *RHS, double *sgs_vert1) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz, nx2 = c->nx2, ny2 = c->ny2;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz, plb = (size_t)c->ld_big * ny2, nb = plb * nz;
  const double dz = c->dz;
  for (int y = 0; y < ny; y++)                                                   /* :278-283 */
    for (int x = 0; x < nx; x++) {
      const size_t i = x + (size_t)ld * y, j = x + (size_t)nx * y;
      dsdz[i] = -phi_h[j] * surf_flux_of(c, which, j) / (ustar_avg[j] * VONK * dz / 2.0);
    }
  double *um = (double *)calloc(nb, sizeof(double)), *vm = (double *)calloc(nb, sizeof(double));
  double *wm = (double *)calloc(nb, sizeof(double)), *sxm = (double *)calloc(nb, sizeof(double));
  double *sym = (double *)calloc(nb, sizeof(double)), *szm = (double *)calloc(nb, sizeof(double));
  double *Rm = (double *)calloc(nb, sizeof(double));
  or_dealias1(c, u, um, nz); or_dealias1(c, v, vm, nz); or_dealias1(c, w, wm, nz);   /* :286-291 */
  or_dealias1(c, dsdx, sxm, nz); or_dealias1(c, dsdy, sym, nz); or_dealias1(c, dsdz, szm, nz);
  for (int k = 0; k < nz; k++)                                                   /* :322-349 */
    for (int y = 0; y < ny2; y++)
      for (int x = 0; x < nx2; x++) {
        const size_t i = plb * k + x + (size_t)c->ld_big * y;
        if (k == 0)
          Rm[i] = um[i] * sxm[i] + vm[i] * sym[i] + (0.5 * wm[i + plb]) * szm[i + plb];
        else if (k < nz - 1)
          Rm[i] = um[i] * sxm[i] + vm[i] * sym[i] + (wm[i] * szm[i] + wm[i + plb] * szm[i + plb]) / 2.0;
        else
          Rm[i] = um[i] * sxm[i] + vm[i] * sym[i];
      }
  or_dealias2(c, RHS, Rm, nz);                                                   /* :352 */
  double *temp = (double *)calloc(n3, sizeof(double)), *dtemp = (double *)calloc(n3, sizeof(double));
  for (int pass = 0; pass < 2; pass++) {                                         /* :366-432 */
    const double *ds = pass == 0 ? dsdx : dsdy;
    memset(temp, 0, sizeof(double) * n3);
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = x + (size_t)ld * y;
        temp[i] = (1.0 / Pr_[i]) * Nu_t[i] * ds[i];
        for (int k = 1; k < nz - 1; k++) {
          const size_t q = pl * k + i;
          temp[q] = (1.0 / (0.5 * (Pr_[q] + Pr_[q + pl]))) * 0.5 * (Nu_t[q] + Nu_t[q + pl]) * ds[q];
        }
      }
    if (pass == 0) or_ddx(c, dtemp, temp, nz); else or_ddy(c, dtemp, temp, nz);
    for (int k = 0; k < (pass == 0 ? nz - 1 : nz); k++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          const size_t q = pl * k + x + (size_t)ld * y;
          RHS[q] = pass == 0 ? (-1.0 * RHS[q] + dtemp[q]) : RHS[q] + dtemp[q];
        }
  }
  memset(temp, 0, sizeof(double) * n3);                                          /* :436-452 */
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      const size_t i = x + (size_t)ld * y, j = x + (size_t)nx * y;
      const double sf = surf_flux_of(c, which, j);
      temp[i] = -1.0 * sf;
      sgs_vert1[i] = sf;
      for (int k = 1; k < nz; k++) {
        const size_t q = pl * k + i;
        temp[q] = (1.0 / Pr_[q]) * Nu_t[q] * dsdz[q];
      }
    }
  or_ddz_w(c, dtemp, temp);
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t q = pl * k + x + (size_t)ld * y;
        RHS[q] = RHS[q] + dtemp[q];
      }
  memset(RHS + pl * (nz - 1), 0, sizeof(double) * pl);
  free(temp); free(dtemp);
  free(um); free(vm); free(wm); free(sxm); free(sym); free(szm); free(Rm);
}

/* value added at the lid per level (step_scalar :491-502): the lid
   gradient applies when the domain is taller than z_i */
static double lid_gradient(const or_ctx *c, int which) {
  const or_scalar_params *S = &c->S;
  return which == 0 ? S->dTdz_top / S->T_scale * S->z_i : S->dqdz_top / S->q_scale * S->z_i;
}

/* step_scalar, scalars_module.f90:457-504 / step_scalar_q */
void or_step_scalar(or_ctx *c, int which, double *sc, const double *R, const double *Rf) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double dt = c->P.dt;
  for (int k = 0; k < nz - 1; k++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = pl * k + x + (size_t)ld * y;
        sc[i] = sc[i] + dt * (1.5 * R[i] - 0.5 * Rf[i]);
      }
  const size_t top = pl * (nz - 1);
  if (c->P.Lz > 1.0) {
    const double gz = lid_gradient(c, which) * c->dz;
    for (size_t i = 0; i < pl; i++) sc[top + i] = sc[top - pl + i] + gz;
  } else {
    for (size_t i = 0; i < pl; i++) sc[top + i] = sc[top - pl + i];
  }
}

/* theta_all_in_one (scalars_module.f90:57-171, which=0) and q_all_in_one
   (scalars_module_q.f90:31-111, which=1); lbc=1, ubc=0, jt >= SCAL_init */
static void theta_q_all_in_one(or_ctx *c, or_state *s, int which) {
  const int nz = c->nz;
  const size_t pl = (size_t)c->ld * c->ny, n3 = pl * nz;
  double *sc = which ? s->qmix : s->theta;
  double *dx = which ? s->dqdx : s->dTdx, *dy = which ? s->dqdy : s->dTdy, *dzf = which ? s->dqdz : s->dTdz;
  double *R = which ? s->RHS_q : s->RHS_T, *Rf = which ? s->RHS_qf : s->RHS_Tf;
  double *Pr_ = (double *)malloc(sizeof(double) * n3);
  const int dyn_pr = c->S.model == 7 || c->sgs_model == 6;                      /* :73-81: models 6, 7 */
  for (size_t i = 0; i < n3; i++) Pr_[i] = dyn_pr ? s->ssgs.Pr_t[i] : c->S.Pr;
  or_filt_da(c, sc, dx, dy, nz);                                                  /* :87-88 */
  or_ddz_uv(c, dzf, sc);
  const double top = lid_gradient(c, which);
  for (size_t i = 0; i < pl; i++) dzf[pl * (nz - 1) + i] = top;                   /* :90-92 */
  memcpy(Rf, R, sizeof(double) * n3);                                             /* :121 */
  or_scalar_rhs_calc(c, which, dx, dy, dzf, s->u, s->v, s->w, s->sgs.Nu_t, Pr_, s->ustar_avg, s->phi_h,
                     R, which ? s->sgs_q3 : s->sgs_t3);
  if (c->F.ubc == 1 && c->F.damping_method == 2)                                  /* :143-148 */
    for (int k = 0; k < nz - 1; k++) {
      const double cxy = 0.5 * (c->sponge[k + 1] + c->sponge[k + 2]), m = level_mean(c, sc, k);
      for (int y = 0; y < c->ny; y++)
        for (int x = 0; x < c->nx; x++) {
          const size_t i = pl * k + x + (size_t)c->ld * y;
          R[i] = R[i] - cxy * (sc[i] - m);
        }
    }
  if (which == 0) or_calcbeta(c, s->theta, s->qmix, s->beta_scal);               /* :154 */
  if (c->jt == 1 && !c->S.initsc) memcpy(Rf, R, sizeof(double) * n3);           /* :157-159 */
  or_step_scalar(c, which, sc, R, Rf);
  free(Pr_);
}

/* lagrange_scalar_Sdep.f90:1-467 (non-MPI, lbc=1, no inflow) */
void or_lagrange_scalar_sdep(or_ctx *c, const double *const S[6], const double *u,
                             const double *v, const double *w, double *theta,
                             or_sgs_state *m, or_scalar_sgs *g) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz;
  const or_scalar_params *P = &c->S;
  const double delta = c->delta, tf1 = 2.0, tf2 = 4.0;
  const double opftdelta = 1.5 * delta, powcoeff = -1.0 / 8.0;
  const double e24 = (double)F1EM24;
  const double expo = log(tf1) / (log(tf2) - log(tf1));
  double *theta_t = (double *)malloc(sizeof(double) * n3);
  memcpy(theta_t, theta, sizeof(double) * n3);                                    /* :84-86 */
  or_interpolag_scalar_sdep(c, g);                                                /* :88 */
  const double lagran_dt = c->P.dt * (double)c->P.cs_count;
  const double cst = delta * delta;
  double *dTdx = (double *)calloc(n3, sizeof(double)), *dTdy = (double *)calloc(n3, sizeof(double));
  double *dTdz = (double *)calloc(n3, sizeof(double));
  or_filt_da(c, theta, dTdx, dTdy, nz);                                           /* :95-99 */
  or_ddz_uv(c, dTdz, theta);
  const double top = lid_gradient(c, 0);
  for (size_t i = 0; i < pl; i++) dTdz[pl * (nz - 1) + i] = top;
  {                                                                               /* :108-136 */
    double slog = 0.0;
    for (int i = 0; i < nx * ny; i++) slog += log(c->P.z0);
    const double zo_avg = exp(slog / (double)(float)(nx * ny));
    csum_t su = { 0.0, 0.0 }, spsi = { 0.0, 0.0 };
    double swt = 0.0;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        csum_add(&su, sqrt(u[x + (size_t)ld * y] * u[x + (size_t)ld * y] + v[x + (size_t)ld * y] * v[x + (size_t)ld * y]));
    for (int j = 0; j < nx * ny; j++) csum_add(&spsi, g->psi_m[j]);
    for (int j = 0; j < nx * ny; j++) swt += c->wt_s[j];
    const double ustar_ = csum_val(&su) / (double)(float)(nx * ny) * VONK /
                          (log(0.5 * c->dz / zo_avg) - csum_val(&spsi) / (double)(float)(nx * ny));
    const double sflux = (swt / (double)(nx * ny)) / P->T_scale / P->u_star;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        dTdz[x + (size_t)ld * y] = -g->phi_h[x + (size_t)nx * y] * sflux / (ustar_ * VONK * c->dz / 2.0);
  }
  enum { UB, VB, WB, TB, UH, VH, WH, TH, L1, L2, L3, Q1, Q2, Q3,
         SB11, SB12, SB13, SB22, SB23, SB33, SH11, SH12, SH13, SH22, SH23, SH33,
         SXB, SYB, SZB, SXH, SYH, SZH, DXB, DYB, DZB, DXH, DYH, DZH, SMAG, NT };
  const int jt = c->jt;
  const int init_lm = c->P.inilag && !c->glm_init && (jt == c->P.cs_count || jt == 1);
  const int init_qn = c->P.inilag && !c->gqn_init && (jt == c->P.cs_count || jt == 1);
#pragma omp parallel
  {
    double *T[NT];
    for (int t = 0; t < NT; t++) T[t] = (double *)calloc(pl, sizeof(double));
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz; k++) {                                                  /* :138-398 */
    const size_t o = pl * k;
    for (size_t i = 0; i < pl; i++) {
      double ub, vb, wb, tb;
      if (k == 0) { ub = u[i]; vb = v[i]; wb = 0.25 * w[pl + i]; tb = theta_t[i]; }
      else {
        ub = 0.5 * (u[o + i] + u[o - pl + i]); vb = 0.5 * (v[o + i] + v[o - pl + i]);
        wb = w[o + i]; tb = 0.5 * (theta_t[o + i] + theta_t[o - pl + i]);
      }
      T[UB][i] = ub; T[VB][i] = vb; T[WB][i] = wb; T[TB][i] = tb;
      T[UH][i] = ub; T[VH][i] = vb; T[WH][i] = wb; T[TH][i] = tb;
      T[L1][i] = ub * tb; T[L2][i] = vb * tb; T[L3][i] = wb * tb;
      T[Q1][i] = ub * tb; T[Q2][i] = vb * tb; T[Q3][i] = wb * tb;
    }
    for (int q = 0; q < 4; q++) or_test_filter(c, T[UB + q], c->Gt);
    for (int q = 0; q < 3; q++) {
      or_test_filter(c, T[L1 + q], c->Gt);
      for (size_t i = 0; i < pl; i++) T[L1 + q][i] = T[L1 + q][i] - T[UB + q][i] * T[TB][i];
    }
    for (int q = 0; q < 4; q++) or_test_filter(c, T[UH + q], c->Gtt);
    for (int q = 0; q < 3; q++) {
      or_test_filter(c, T[Q1 + q], c->Gtt);
      for (size_t i = 0; i < pl; i++) T[Q1 + q][i] = T[Q1 + q][i] - T[UH + q][i] * T[TH][i];
    }
    for (size_t i = 0; i < pl; i++) {
      double s11 = S[0][o + i], s12 = S[1][o + i], s13 = S[2][o + i];
      double s22 = S[3][o + i], s23 = S[4][o + i], s33 = S[5][o + i];
      T[SMAG][i] = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
      for (int q = 0; q < 6; q++) { T[SB11 + q][i] = S[q][o + i]; T[SH11 + q][i] = S[q][o + i]; }
    }
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SB11 + q], c->Gt);
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SH11 + q], c->Gtt);
    for (size_t i = 0; i < pl; i++) {
      const double Sm = T[SMAG][i];
      double sx, sy, sz, dxv, dyv, dzv;
      if (k == 0) {
        sx = Sm * dTdx[i]; sy = Sm * dTdy[i]; sz = Sm * (dTdz[i] + dTdz[i + pl]) / 2.0;
        dxv = dTdx[i]; dyv = dTdy[i]; dzv = (dTdz[i] + dTdz[i + pl]) / 2.0;
      } else {
        sx = Sm * (dTdx[o + i] + dTdx[o - pl + i]) / 2.0;
        sy = Sm * (dTdy[o + i] + dTdy[o - pl + i]) / 2.0;
        sz = Sm * dTdz[o + i];
        dxv = (dTdx[o + i] + dTdx[o - pl + i]) / 2.0;
        dyv = (dTdy[o + i] + dTdy[o - pl + i]) / 2.0;
        dzv = dTdz[o + i];
      }
      T[SXB][i] = sx; T[SYB][i] = sy; T[SZB][i] = sz; T[SXH][i] = sx; T[SYH][i] = sy; T[SZH][i] = sz;
      T[DXB][i] = dxv; T[DYB][i] = dyv; T[DZB][i] = dzv; T[DXH][i] = dxv; T[DYH][i] = dyv; T[DZH][i] = dzv;
    }
    for (int q = 0; q < 3; q++) { or_test_filter(c, T[SXB + q], c->Gt); or_test_filter(c, T[SXH + q], c->Gtt); }
    for (int q = 0; q < 3; q++) { or_test_filter(c, T[DXB + q], c->Gt); or_test_filter(c, T[DXH + q], c->Gtt); }
    double *GLM = g->G_LM + o, *GMM = g->G_MM + o, *GQN = g->G_QN + o, *GNN = g->G_NN + o;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < ld; x++) {
        const size_t i = x + (size_t)ld * y;
        const double Sb = sqrt(2.0 * (T[SB11][i] * T[SB11][i] + T[SB22][i] * T[SB22][i] + T[SB33][i] * T[SB33][i] +
                                      2.0 * (T[SB12][i] * T[SB12][i] + T[SB13][i] * T[SB13][i] + T[SB23][i] * T[SB23][i])));
        const double Sh = sqrt(2.0 * (T[SH11][i] * T[SH11][i] + T[SH22][i] * T[SH22][i] + T[SH33][i] * T[SH33][i] +
                                      2.0 * (T[SH12][i] * T[SH12][i] + T[SH13][i] * T[SH13][i] + T[SH23][i] * T[SH23][i])));
        double M[3], N[3];
        for (int q = 0; q < 3; q++) {                                             /* :268-274 */
          M[q] = cst * (T[SXB + q][i] - (OR_MUTANT == 6 ? 16.0 : 4.0) * Sb * T[DXB + q][i]);
          N[q] = cst * (T[SXH + q][i] - (OR_MUTANT == 6 ? 4.0 : 16.0) * Sh * T[DXH + q][i]);
        }
        const double LM = T[L1][i] * M[0] + T[L2][i] * M[1] + T[L3][i] * M[2];  /* :284-287 */
        const double MM = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
        const double QN = T[Q1][i] * N[0] + T[Q2][i] * N[1] + T[Q3][i] * N[2];
        const double NN = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
        if (init_lm) {                                                            /* :289-299 */
          GMM[i] = MM; GLM[i] = 0.03 * MM;
          if (x >= ld - 2) { GMM[i] = 1.0; GLM[i] = 1.0; }
        }
        double Tn = m->TnLMMM[o + i];                                              /* :318-331 */
        Tn = opftdelta * pow(Tn, powcoeff);
        Tn = fmax(e24, fq(Tn));
        double dumfac = lagran_dt / Tn;
        double epsi = dumfac / (1.0 + dumfac);
        GLM[i] = (epsi * LM + (1.0 - epsi) * GLM[i]);
        GMM[i] = (epsi * MM + (1.0 - epsi) * GMM[i]);
        GLM[i] = fmax(e24, fq(GLM[i]));
        double c2 = GLM[i] / GMM[i];
        if (x >= ld - 2) c2 = e24;
        c2 = fmax(e24, fq(c2));
        if (init_qn) {                                                            /* :338-348 */
          GNN[i] = NN; GQN[i] = 0.03 * NN;
          if (x >= ld - 2) { GNN[i] = 1.0; GQN[i] = 1.0; }
        }
        Tn = m->TnQNNN[o + i];                                                     /* :350-363 */
        Tn = opftdelta * pow(Tn, powcoeff);
        Tn = fmax(e24, fq(Tn));
        dumfac = lagran_dt / Tn;
        epsi = dumfac / (1.0 + dumfac);
        GQN[i] = (epsi * QN + (1.0 - epsi) * GQN[i]);
        GNN[i] = (epsi * NN + (1.0 - epsi) * GNN[i]);
        GQN[i] = fmax(e24, fq(GQN[i]));
        double c4 = GQN[i] / GNN[i];
        if (x >= ld - 2) c4 = e24;
        c4 = fmax(e24, fq(c4));
        double beta = OR_MUTANT == 7 ? pow(c2 / c4, expo) : pow(c4 / c2, expo);  /* :370 */
        if (k == nz - 1) beta = 1.0;                                              /* :379-382 */
        m->Beta[o + i] = beta;
        const double bclip = fmax(fq(beta), fq(1.0 / (tf1 * tf2)));               /* :384-389 */
        double cpr = c2 / bclip;
        if (x >= ld - 2) cpr = e24;
        cpr = fmax(e24, fq(cpr));
        g->Cs2_Pr[o + i] = cpr;
        g->Pr_t[o + i] = m->Cs_opt2[o + i] / cpr;                                  /* :395 */
      }
  }
    for (int t = 0; t < NT; t++) free(T[t]);
  }
  if (init_lm) c->glm_init = 1;
  if (init_qn) c->gqn_init = 1;
  free(theta_t); free(dTdx); free(dTdy); free(dTdz);
}

/* seeded convective initial condition for theta and q (project IC, not a
   reference routine): theta = T_init below z_inv, + inv_strength (K/m)
   above, uniform perturbation of amplitude amp_K below z_inv; q uniform */
void or_state_init_scalar_ic(or_ctx *c, or_state *s, uint64_t seed, double T_init,
                             double z_inv, double inv_strength, double amp_K, double q_init) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz, n2 = (size_t)nx * ny;
  const or_scalar_params *S = &c->S;
  memset(s->theta, 0, n3 * sizeof(double)); memset(s->qmix, 0, n3 * sizeof(double));
  for (int k = 0; k < nz - 1; k++) {
    const double z = (k + 0.5) * c->dz;
    const double tk = T_init + (z > z_inv ? inv_strength * (z - z_inv) * S->z_i : 0.0);
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = pl * k + x + (size_t)ld * y;
        double t = tk;
        if (z < z_inv) {
          const uint64_t idx = (((uint64_t)3 * nz + k) * ny + y) * nx + x;
          t += amp_K * (2.0 * or_uniform(seed, idx) - 1.0);
        }
        s->theta[i] = t / S->T_scale;
        s->qmix[i] = q_init / S->q_scale;
      }
  }
  const size_t top = pl * (nz - 1);
  for (size_t i = 0; i < pl; i++) {
    const int x = (int)(i % ld);
    if (x >= nx) continue;
    s->theta[top + i] = s->theta[top - pl + i] + (c->P.Lz > 1.0 ? lid_gradient(c, 0) * c->dz : 0.0);
    s->qmix[top + i] = s->qmix[top - pl + i] + (c->P.Lz > 1.0 ? lid_gradient(c, 1) * c->dz : 0.0);
  }
  memset(s->RHS_T, 0, n3 * sizeof(double)); memset(s->RHS_Tf, 0, n3 * sizeof(double));
  memset(s->RHS_q, 0, n3 * sizeof(double)); memset(s->RHS_qf, 0, n3 * sizeof(double));
  memset(s->beta_scal, 0, n3 * sizeof(double));
  memset(s->dTdx, 0, n3 * sizeof(double)); memset(s->dTdy, 0, n3 * sizeof(double)); memset(s->dTdz, 0, n3 * sizeof(double));
  memset(s->dqdx, 0, n3 * sizeof(double)); memset(s->dqdy, 0, n3 * sizeof(double)); memset(s->dqdz, 0, n3 * sizeof(double));
  memset(s->sgs_t3, 0, pl * sizeof(double)); memset(s->sgs_q3, 0, pl * sizeof(double));
  for (size_t j = 0; j < n2; j++) { s->psi_h[j] = 0.0; s->psi_h0[j] = 0.0; s->phi_h[j] = 1.0; }
  or_scalar_sgs *h = &s->ssgs;
  memset(h->G_LM, 0, n3 * sizeof(double)); memset(h->G_MM, 0, n3 * sizeof(double));
  memset(h->G_QN, 0, n3 * sizeof(double)); memset(h->G_NN, 0, n3 * sizeof(double));
  memset(h->u_lag_s, 0, n3 * sizeof(double)); memset(h->v_lag_s, 0, n3 * sizeof(double));
  memset(h->w_lag_s, 0, n3 * sizeof(double)); memset(h->Cs2_Pr, 0, n3 * sizeof(double));
  for (size_t i = 0; i < n3; i++) h->Pr_t[i] = S->Pr;
  c->glm_init = 0; c->gqn_init = 0;
}

/* ------------------------------------------------------------------ */
/* rtnewt, sgsmodule.f90:37-59: Newton from x=7.5, |dx| < 1e-3, at most 100
   iterations, else beta = 1                                            */
/* ------------------------------------------------------------------ */
double or_rtnewt(const double A[6]) {
  double x = 0.5 * (0.0 + 15.0);
  for (int j = 1; j <= 100; j++) {
    double f = A[0] + x * (A[1] + x * (A[2] + x * (A[3] + x * (A[4] + x * A[5]))));
    double df = A[1] + x * (2.0 * A[2] + x * (3.0 * A[3] + x * (4.0 * A[4] + x * (5.0 * A[5]))));
    double dx = f / df;
    x = x - dx;
    if (fabs(dx) < 0.001) return x;
  }
  return 1.0;
}

/* The filtered fields of one level that scaledep_dynamic.f90:44-181 and
   new_scaledep_dynamic.f90:52-181 share: the 2-Delta bars of the w-node
   velocities and their 4-Delta hats, L_ij, Q_ij (4-Delta filter of products
   of the 2-Delta bars), the 2- and 4-Delta filtered S_ij and |S| S_ij, and
   |S| itself.  S[0..5] are the routine's S11..S33 dummies. */
enum { UB, VB, WB, UH, VH, WH, L11, L12, L13, L22, L23, L33, Q11, Q12, Q13, Q22, Q23, Q33,
       SB11, SB12, SB13, SB22, SB23, SB33, SH11, SH12, SH13, SH22, SH23, SH33,
       SSB11, SSB12, SSB13, SSB22, SSB23, SSB33, SSH11, SSH12, SSH13, SSH22, SSH23, SSH33,
       SMAG, SBAR, SHAT, NT };

static void sd_level_fields(or_ctx *c, const double *const S[6], const double *u, const double *v,
                            const double *w, int k, double *T[NT]) {
  const size_t pl = (size_t)c->ld * c->ny, o = pl * k;
  for (size_t i = 0; i < pl; i++) {                                      /* :44-73 */
    double ub, vb, wb;
    if (k == 0) { ub = u[i]; vb = v[i]; wb = 0.25 * w[pl + i]; }
    else { ub = 0.5 * (u[o + i] + u[o - pl + i]); vb = 0.5 * (v[o + i] + v[o - pl + i]); wb = w[o + i]; }
    T[UB][i] = ub; T[VB][i] = vb; T[WB][i] = wb;
    T[UH][i] = ub; T[VH][i] = vb; T[WH][i] = wb;
    T[L11][i] = ub * ub; T[L12][i] = ub * vb; T[L13][i] = ub * wb;
    T[L22][i] = vb * vb; T[L23][i] = vb * wb; T[L33][i] = wb * wb;
  }
  for (int q = 0; q < 3; q++) or_test_filter(c, T[UB + q], c->Gt);        /* :78-80 */
  const int Lp[6][3] = { { L11, UB, UB }, { L12, UB, VB }, { L13, UB, WB }, { L22, VB, VB }, { L23, VB, WB }, { L33, WB, WB } };
  for (int q = 0; q < 6; q++) {                                           /* :81-92 */
    or_test_filter(c, T[Lp[q][0]], c->Gt);
    for (size_t i = 0; i < pl; i++) T[Lp[q][0]][i] = T[Lp[q][0]][i] - T[Lp[q][1]][i] * T[Lp[q][2]][i];
  }
  for (size_t i = 0; i < pl; i++) {                                       /* :94-99 products of the 2-Delta bars */
    T[Q11][i] = T[UB][i] * T[UB][i]; T[Q12][i] = T[UB][i] * T[VB][i]; T[Q13][i] = T[UB][i] * T[WB][i];
    T[Q22][i] = T[VB][i] * T[VB][i]; T[Q23][i] = T[VB][i] * T[WB][i]; T[Q33][i] = T[WB][i] * T[WB][i];
  }
  for (int q = 0; q < 3; q++) or_test_filter(c, T[UH + q], c->Gtt);       /* :101-103 */
  const int Qp[6][3] = { { Q11, UH, UH }, { Q12, UH, VH }, { Q13, UH, WH }, { Q22, VH, VH }, { Q23, VH, WH }, { Q33, WH, WH } };
  for (int q = 0; q < 6; q++) {                                           /* :104-115 */
    or_test_filter(c, T[Qp[q][0]], c->Gtt);
    for (size_t i = 0; i < pl; i++) T[Qp[q][0]][i] = T[Qp[q][0]][i] - T[Qp[q][1]][i] * T[Qp[q][2]][i];
  }
  for (size_t i = 0; i < pl; i++) {                                       /* :117-139 */
    double s11 = S[0][o + i], s12 = S[1][o + i], s13 = S[2][o + i];
    double s22 = S[3][o + i], s23 = S[4][o + i], s33 = S[5][o + i];
    double Sm = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
    T[SMAG][i] = Sm;
    for (int q = 0; q < 6; q++) {
      T[SB11 + q][i] = S[q][o + i]; T[SH11 + q][i] = S[q][o + i];
      T[SSB11 + q][i] = Sm * S[q][o + i]; T[SSH11 + q][i] = Sm * S[q][o + i];
    }
  }
  for (int q = 0; q < 6; q++) or_test_filter(c, T[SB11 + q], c->Gt);    /* :140-151 */
  for (int q = 0; q < 6; q++) or_test_filter(c, T[SH11 + q], c->Gtt);
  for (int q = 0; q < 6; q++) or_test_filter(c, T[SSB11 + q], c->Gt);   /* :170-181 */
  for (int q = 0; q < 6; q++) or_test_filter(c, T[SSH11 + q], c->Gtt);
}

/* ------------------------------------------------------------------ */
/* scaledep_dynamic.f90:4-255 (non-MPI).  S[0..5] are the routine's dummy
   arguments S11, S12, S13, S22, S23, S33 in that order: the caller passes
   what sgs_stag.f90:132 binds to them (u, v, w, S11, S12, S13).  u, v, w are
   sim_param's.  Plane sums run over the full (ld, ny) arrays.  Outputs
   Cs_opt2(nz) and, if beta != NULL, the per-level beta of rtnewt.     */
/* ------------------------------------------------------------------ */
void or_scaledep_dynamic(or_ctx *c, const double *const S[6], const double *u,
                         const double *v, const double *w, double *Cs_opt2,
                         double *beta_out) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double delta = c->delta;
  const double nxy = (double)(nx * ny);
#pragma omp parallel
  {
    double *T[NT];
    for (int t = 0; t < NT; t++) T[t] = (double *)calloc(pl, sizeof(double));
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz; k++) {
    sd_level_fields(c, S, u, v, w, k, T);
    double sa1 = 0, sb1 = 0, sc1 = 0, sd1 = 0, se1 = 0, sa2 = 0, sb2 = 0, sc2 = 0, sd2 = 0, se2 = 0;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < ld; x++) {
        const size_t i = x + (size_t)ld * y;
        const double sb11 = T[SB11][i], sb12 = T[SB12][i], sb13 = T[SB13][i], sb22 = T[SB22][i], sb23 = T[SB23][i], sb33 = T[SB33][i];
        const double sh11 = T[SH11][i], sh12 = T[SH12][i], sh13 = T[SH13][i], sh22 = T[SH22][i], sh23 = T[SH23][i], sh33 = T[SH33][i];
        const double Sb = sqrt(2.0 * (sb11 * sb11 + sb22 * sb22 + sb33 * sb33 + 2.0 * (sb12 * sb12 + sb13 * sb13 + sb23 * sb23)));
        const double Sh = sqrt(2.0 * (sh11 * sh11 + sh22 * sh22 + sh33 * sh33 + 2.0 * (sh12 * sh12 + sh13 * sh13 + sh23 * sh23)));
        T[SBAR][i] = Sb; T[SHAT][i] = Sh;
        const double l11 = T[L11][i], l12 = T[L12][i], l13 = T[L13][i], l22 = T[L22][i], l23 = T[L23][i], l33 = T[L33][i];
        const double q11 = T[Q11][i], q12 = T[Q12][i], q13 = T[Q13][i], q22 = T[Q22][i], q23 = T[Q23][i], q33 = T[Q33][i];
        const double ssb11 = T[SSB11][i], ssb12 = T[SSB12][i], ssb13 = T[SSB13][i], ssb22 = T[SSB22][i], ssb23 = T[SSB23][i], ssb33 = T[SSB33][i];
        const double ssh11 = T[SSH11][i], ssh12 = T[SSH12][i], ssh13 = T[SSH13][i], ssh22 = T[SSH22][i], ssh23 = T[SSH23][i], ssh33 = T[SSH33][i];
        sa1 += Sb * (sb11 * l11 + sb22 * l22 + sb33 * l33 + 2.0 * (sb12 * l12 + sb13 * l13 + sb23 * l23));
        sb1 += ssb11 * l11 + ssb22 * l22 + ssb33 * l33 + 2.0 * (ssb12 * l12 + ssb13 * l13 + ssb23 * l23);
        sc1 += ssb11 * ssb11 + ssb22 * ssb22 + ssb33 * ssb33 + 2.0 * (ssb12 * ssb12 + ssb13 * ssb13 + ssb23 * ssb23);
        sd1 += 0.5 * ((Sb * Sb) * (Sb * Sb));       /* S_bar**4 as an integer power */
        se1 += Sb * (sb11 * ssb11 + sb22 * ssb22 + sb33 * ssb33 + 2.0 * (sb12 * ssb12 + sb13 * ssb13 + sb23 * ssb23));
        sa2 += Sh * (sh11 * q11 + sh22 * q22 + sh33 * q33 + 2.0 * (sh12 * q12 + sh13 * q13 + sh23 * q23));
        sb2 += ssh11 * q11 + ssh22 * q22 + ssh33 * q33 + 2.0 * (ssh12 * q12 + ssh13 * q13 + ssh23 * q23);
        sc2 += ssh11 * ssh11 + ssh22 * ssh22 + ssh33 * ssh33 + 2.0 * (ssh12 * ssh12 + ssh13 * ssh13 + ssh23 * ssh23);
        sd2 += 0.5 * ((Sh * Sh) * (Sh * Sh));
        se2 += Sh * (sh11 * ssh11 + sh22 * ssh22 + sh33 * ssh33 + 2.0 * (sh12 * ssh12 + sh13 * ssh13 + sh23 * ssh23));
      }
    const double d2 = delta * delta, cc = (2.0 * d2) * (2.0 * d2);          /* :187-215 */
    const double a1 = -(2.0 * d2 * 4.0 * sa1 / nxy);
    const double b1 = -(2.0 * d2 * sb1 / nxy);
    const double c1 = cc * sc1 / nxy;
    const double d1 = cc * 16.0 * sd1 / nxy;
    const double e1 = 2.0 * cc * 4.0 * se1 / nxy;
    const double a2 = -(2.0 * d2 * 16.0 * sa2 / nxy);
    const double b2 = -(2.0 * d2 * sb2 / nxy);
    const double c2 = cc * sc2 / nxy;
    const double d2c = cc * 256.0 * sd2 / nxy;
    const double e2 = 2.0 * cc * 16.0 * se2 / nxy;
    double A[6];                                                            /* :217-222 */
    A[0] = b2 * c1 - b1 * c2;
    A[1] = a1 * c2 - b2 * e1;
    A[2] = b2 * d1 + b1 * e2 - a2 * c1;
    A[3] = a2 * e1 - a1 * e2;
    A[4] = -a2 * d1 - b1 * d2c;
    A[5] = a1 * d2c;
    const double beta = or_rtnewt(A);
    if (beta_out) beta_out[k] = beta;
    const double cst = 2.0 * d2;                                            /* :225-244 */
    double num = 0.0, den = 0.0;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < ld; x++) {
        const size_t i = x + (size_t)ld * y;
        const double Sb = T[SBAR][i];
        double M[6];
        for (int q = 0; q < 6; q++) M[q] = cst * (T[SSB11 + q][i] - 4.0 * beta * Sb * T[SB11 + q][i]);
        /* q map: 0=11 1=12 2=13 3=22 4=23 5=33 */
        num += T[L11][i] * M[0] + T[L22][i] * M[3] + T[L33][i] * M[5] +
               2.0 * (T[L12][i] * M[1] + T[L13][i] * M[2] + T[L23][i] * M[4]);
        den += M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
      }
    Cs_opt2[k] = fmax(0.0, num / den);
  }
    for (int t = 0; t < NT; t++) free(T[t]);
  }

}

/* ------------------------------------------------------------------ */
/* new_scaledep_dynamic.f90:1-223 (non-MPI; model = 6, called at
   sgs_stag.f90:143-148 with calc_Sij's S11..S33).  Per level the plane sums
   L:M, M:M, Q:N and N:N over the (ld, ny) arrays (compensated here), with
   M_ij = 2 Delta^2 (bar(|S|S_ij) - 4 |S_bar| S_bar_ij) and
   N_ij = 2 Delta^2 (hat(|S|S_ij) - 16 |S_hat| S_hat_ij); the two
   coefficients LM/MM, QN/NN clipped at 1E-24 in single precision, Beta their
   ratio clipped below at 1/8 in single precision, Cs_opt2 = max(0, LM/MM /
   betaclip).  Output: Cs_opt2(nz).                                       */
/* ------------------------------------------------------------------ */
void or_new_scaledep_dynamic(or_ctx *c, const double *const S[6], const double *u,
                             const double *v, const double *w, double *Cs_opt2) {
  const int ld = c->ld, ny = c->ny, nz = c->nz;
  const size_t pl = (size_t)ld * ny;
  const double delta = c->delta;                                          /* :43 */
  const double tf1 = 2.0, tf2 = 4.0, tf1_2 = tf1 * tf1, tf2_2 = tf2 * tf2; /* :44-47 */
  const double cst = 2.0 * (delta * delta);                               /* :49 */
#pragma omp parallel
  {
    double *T[NT];
    for (int t = 0; t < NT; t++) T[t] = (double *)calloc(pl, sizeof(double));
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz; k++) {
    sd_level_fields(c, S, u, v, w, k, T);                                 /* :51-186 */
    csum_t lm = { 0.0, 0.0 }, mm = { 0.0, 0.0 }, qn = { 0.0, 0.0 }, nn = { 0.0, 0.0 };
    for (size_t i = 0; i < pl; i++) {
      double sb[6], sh[6], L[6], Q[6], M[6], N[6];
      for (int q = 0; q < 6; q++) {
        sb[q] = T[SB11 + q][i]; sh[q] = T[SH11 + q][i]; L[q] = T[L11 + q][i]; Q[q] = T[Q11 + q][i];
      }
      /* q: 0=11 1=12 2=13 3=22 4=23 5=33 */
      const double Sb = sqrt(2.0 * (sb[0] * sb[0] + sb[3] * sb[3] + sb[5] * sb[5] +
                                    2.0 * (sb[1] * sb[1] + sb[2] * sb[2] + sb[4] * sb[4])));   /* :154-155 */
      const double Sh = sqrt(2.0 * (sh[0] * sh[0] + sh[3] * sh[3] + sh[5] * sh[5] +
                                    2.0 * (sh[1] * sh[1] + sh[2] * sh[2] + sh[4] * sh[4])));   /* :157-158 */
      for (int q = 0; q < 6; q++) {                                       /* :188-200 */
        M[q] = cst * (T[SSB11 + q][i] - (OR_MUTANT == 13 ? tf2_2 : tf1_2) * Sb * sb[q]);
        N[q] = cst * (T[SSH11 + q][i] - (OR_MUTANT == 13 ? tf1_2 : tf2_2) * Sh * sh[q]);
      }
      csum_add(&lm, L[0] * M[0] + L[3] * M[3] + L[5] * M[5] + 2.0 * (L[1] * M[1] + L[2] * M[2] + L[4] * M[4]));
      csum_add(&mm, M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]));
      csum_add(&qn, Q[0] * N[0] + Q[3] * N[3] + Q[5] * N[5] + 2.0 * (Q[1] * N[1] + Q[2] * N[2] + Q[4] * N[4]));
      csum_add(&nn, N[0] * N[0] + N[3] * N[3] + N[5] * N[5] + 2.0 * (N[1] * N[1] + N[2] * N[2] + N[4] * N[4]));
    }
    double c2 = csum_val(&lm) / csum_val(&mm);                            /* :210-211 */
    double c4 = csum_val(&qn) / csum_val(&nn);
    c2 = (double)fmaxf(F1EM24, (float)c2);                                /* :213-214 */
    c4 = (double)fmaxf(F1EM24, (float)c4);
    const double beta = c4 / c2;                                          /* :216 */
    const double betaclip = (double)fmaxf((float)beta, (float)(1.0 / 8.0));   /* :217 */
    Cs_opt2[k] = fmax(0.0, c2 / betaclip);                                /* :219-220 */
  }
    for (int t = 0; t < NT; t++) free(T[t]);
  }
}

/* ------------------------------------------------------------------ */
/* scalar_scaledep_dynamic.f90:1-307 (non-MPI, lbc = 1; model = 6 after
   new_scaledep_dynamic, sgs_stag.f90:149).  theta is filtered in place by
   its filt_da (:60) and read filtered afterwards; dTdz from ddz_uv with the
   lid gradient (:61-64) and, on the first level (the whole (ld, ny) plane),
   the uniform phi_h_avg * surface_flux / (ustar_ vonk dz/2) with ustar_ from
   the mean first-level speed and psi_m (:73-79).  Per level 1..nz-1: L_i,
   Q_i (4-Delta filter of products of the 2-Delta bars), M_i and N_i from
   |S| dT_i and dT_i without staggering averages, const = Delta^2; the plane
   sums give Cs2_Pr2, Cs2_Pr4, Beta and Cs2_Pr as in the momentum routine
   (:227-240), and Pr_t(:,:,jz) = Cs_opt2(:,:,jz) / Cs2_Pr(jz) (:242).
   Outputs Cs2_Pr(nz-1) and Pr_t levels 1..nz-1.                          */
/* ------------------------------------------------------------------ */
void or_scalar_scaledep_dynamic(or_ctx *c, const double *const S[6], const double *u,
                                const double *v, const double *w, double *theta,
                                const double *Cs_opt2, const double *psi_m, const double *phi_h,
                                double *Cs2_Pr, double *Pr_t) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nz = c->nz;
  const size_t pl = (size_t)ld * ny, n3 = pl * nz;
  const or_scalar_params *P = &c->S;
  const double delta = c->delta, cst = delta * delta;                     /* :58, :219 */
  double *dTdx = (double *)calloc(n3, sizeof(double)), *dTdy = (double *)calloc(n3, sizeof(double));
  double *dTdz = (double *)calloc(n3, sizeof(double));
  or_filt_da(c, theta, dTdx, dTdy, nz);                                   /* :60 */
  or_ddz_uv(c, dTdz, theta);                                              /* :61 */
  const double top = lid_gradient(c, 0);
  for (size_t i = 0; i < pl; i++) dTdz[pl * (nz - 1) + i] = top;          /* :62-64 */
  {                                                                       /* :73-79 */
    double slog = 0.0, swt = 0.0;
    for (int i = 0; i < nx * ny; i++) slog += log(c->P.z0);
    const double fn = (double)(float)(nx * ny);
    const double zo_avg = exp(slog / fn);
    csum_t su = { 0.0, 0.0 }, spsi = { 0.0, 0.0 }, sphi = { 0.0, 0.0 };
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        csum_add(&su, sqrt(u[x + (size_t)ld * y] * u[x + (size_t)ld * y] + v[x + (size_t)ld * y] * v[x + (size_t)ld * y]));
    for (int j = 0; j < nx * ny; j++) csum_add(&spsi, psi_m[j]);
    for (int j = 0; j < nx * ny; j++) csum_add(&sphi, phi_h[j]);
    for (int j = 0; j < nx * ny; j++) swt += c->wt_s[j];
    const double ustar_ = csum_val(&su) / fn * VONK / (log(0.5 * c->dz / zo_avg) - csum_val(&spsi) / fn);
    const double phi_h_avg = csum_val(&sphi) / fn;
    const double sflux = (swt / (double)(nx * ny)) / P->T_scale / P->u_star;
    const double d1 = phi_h_avg * sflux / (ustar_ * VONK * c->dz / 2.0);
    for (size_t i = 0; i < pl; i++) dTdz[i] = d1;
  }
  enum { UB, VB, WB, TB, UH, VH, WH, TH, L1, L2, L3, Q1, Q2, Q3,
         SB11, SB12, SB13, SB22, SB23, SB33, SH11, SH12, SH13, SH22, SH23, SH33,
         SXB, SYB, SZB, SXH, SYH, SZH, DXB, DYB, DZB, DXH, DYH, DZH, NT };
#pragma omp parallel
  {
    double *T[NT];
    for (int t = 0; t < NT; t++) T[t] = (double *)calloc(pl, sizeof(double));
#pragma omp for schedule(dynamic)
  for (int k = 0; k < nz - 1; k++) {                                      /* :82-244 */
    const size_t o = pl * k;
    for (size_t i = 0; i < pl; i++) {                                     /* :84-105 */
      double ub, vb, wb, tb;
      if (k == 0) { ub = u[i]; vb = v[i]; wb = 0.25 * w[pl + i]; tb = theta[i]; }
      else {
        ub = 0.5 * (u[o + i] + u[o - pl + i]); vb = 0.5 * (v[o + i] + v[o - pl + i]); wb = w[o + i];
        tb = 0.5 * (theta[o + i] + theta[o - pl + i]);
      }
      T[UB][i] = ub; T[VB][i] = vb; T[WB][i] = wb; T[TB][i] = tb;
      T[UH][i] = ub; T[VH][i] = vb; T[WH][i] = wb; T[TH][i] = tb;
      T[L1][i] = ub * tb; T[L2][i] = vb * tb; T[L3][i] = wb * tb;
    }
    for (int q = 0; q < 4; q++) or_test_filter(c, T[UB + q], c->Gt);     /* :111-114 */
    for (int q = 0; q < 3; q++) {                                         /* :116-121 */
      or_test_filter(c, T[L1 + q], c->Gt);
      for (size_t i = 0; i < pl; i++) T[L1 + q][i] = T[L1 + q][i] - T[UB + q][i] * T[TB][i];
    }
    for (size_t i = 0; i < pl; i++)                                       /* :123-125 */
      for (int q = 0; q < 3; q++) T[Q1 + q][i] = T[UB + q][i] * T[TB][i];
    for (int q = 0; q < 4; q++) or_test_filter(c, T[UH + q], c->Gtt);    /* :127-130 */
    for (int q = 0; q < 3; q++) {                                         /* :132-137 */
      or_test_filter(c, T[Q1 + q], c->Gtt);
      for (size_t i = 0; i < pl; i++) T[Q1 + q][i] = T[Q1 + q][i] - T[UH + q][i] * T[TH][i];
    }
    for (size_t i = 0; i < pl; i++) {                                     /* :140-200 */
      const double s11 = S[0][o + i], s12 = S[1][o + i], s13 = S[2][o + i];
      const double s22 = S[3][o + i], s23 = S[4][o + i], s33 = S[5][o + i];
      const double Sm = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
      for (int q = 0; q < 6; q++) { T[SB11 + q][i] = S[q][o + i]; T[SH11 + q][i] = S[q][o + i]; }
      const double dt3[3] = { dTdx[o + i], dTdy[o + i], dTdz[o + i] };
      for (int q = 0; q < 3; q++) {
        T[SXB + q][i] = Sm * dt3[q]; T[SXH + q][i] = Sm * dt3[q];
        T[DXB + q][i] = dt3[q]; T[DXH + q][i] = dt3[q];
      }
    }
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SB11 + q], c->Gt);   /* :159-171 */
    for (int q = 0; q < 6; q++) or_test_filter(c, T[SH11 + q], c->Gtt);
    for (int q = 0; q < 3; q++) or_test_filter(c, T[SXB + q], c->Gt);    /* :187-193 */
    for (int q = 0; q < 3; q++) or_test_filter(c, T[SXH + q], c->Gtt);
    for (int q = 0; q < 3; q++) or_test_filter(c, T[DXB + q], c->Gt);    /* :203-209 */
    for (int q = 0; q < 3; q++) or_test_filter(c, T[DXH + q], c->Gtt);
    csum_t lm = { 0.0, 0.0 }, mm = { 0.0, 0.0 }, qn = { 0.0, 0.0 }, nn = { 0.0, 0.0 };
    for (size_t i = 0; i < pl; i++) {
      const double sb11 = T[SB11][i], sb12 = T[SB12][i], sb13 = T[SB13][i], sb22 = T[SB22][i], sb23 = T[SB23][i], sb33 = T[SB33][i];
      const double sh11 = T[SH11][i], sh12 = T[SH12][i], sh13 = T[SH13][i], sh22 = T[SH22][i], sh23 = T[SH23][i], sh33 = T[SH33][i];
      const double Sb = sqrt(2.0 * (sb11 * sb11 + sb22 * sb22 + sb33 * sb33 + 2.0 * (sb12 * sb12 + sb13 * sb13 + sb23 * sb23)));  /* :173-174 */
      const double Sh = sqrt(2.0 * (sh11 * sh11 + sh22 * sh22 + sh33 * sh33 + 2.0 * (sh12 * sh12 + sh13 * sh13 + sh23 * sh23)));  /* :176-177 */
      double M[3], N[3];
      for (int q = 0; q < 3; q++) {                                       /* :211-218 */
        M[q] = cst * (T[SXB + q][i] - 4.0 * Sb * T[DXB + q][i]);
        N[q] = cst * (T[SXH + q][i] - 16.0 * Sh * T[DXH + q][i]);
      }
      csum_add(&lm, T[L1][i] * M[0] + T[L2][i] * M[1] + T[L3][i] * M[2]);  /* :230-231 */
      csum_add(&mm, M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
      csum_add(&qn, T[Q1][i] * N[0] + T[Q2][i] * N[1] + T[Q3][i] * N[2]);
      csum_add(&nn, N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    }
    double c2 = csum_val(&lm) / csum_val(&mm), c4 = csum_val(&qn) / csum_val(&nn);
    c2 = (double)fmaxf(F1EM24, (float)c2);                                /* :233-234 */
    c4 = (double)fmaxf(F1EM24, (float)c4);
    const double beta = OR_MUTANT == 14 ? c2 / c4 : c4 / c2;             /* :236 */
    const double betaclip = (double)fmaxf((float)beta, (float)(1.0 / 8.0));
    double cpr = c2 / betaclip;                                           /* :239-240 */
    cpr = fmax(0.0, cpr);
    Cs2_Pr[k] = cpr;
    for (size_t i = 0; i < pl; i++) Pr_t[o + i] = Cs_opt2[o + i] / cpr;   /* :242 */
  }
    for (int t = 0; t < NT; t++) free(T[t]);
  }
  free(dTdx); free(dTdy); free(dTdz);
}

/* ------------------------------------------------------------------ */
/* statistics: io.f90 spectrum / avg_stats, obukhov's plane averages    */
/* ------------------------------------------------------------------ */

/* spectrum, io.f90:1613-1646: per row vel_r = f/nx, its halfcomplex
   transform, spec(1) += 0.5 c_0^2, spec(k) += |c_(k-1)|^2 (k = 2..nx/2,
   Nyquist taken as 0), then spec / ny */
void or_spectrum(or_ctx *c, const double *f, int nlev, double *spec) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nk = nx / 2;
  const size_t pl = (size_t)ld * ny;
  cpx *row = (cpx *)malloc(sizeof(cpx) * (size_t)nx), *t = (cpx *)malloc(sizeof(cpx) * (size_t)nx);
  csum_t *acc = (csum_t *)malloc(sizeof(csum_t) * (size_t)nk);
  for (int k = 0; k < nlev; k++) {
    for (int j = 0; j < nk; j++) { acc[j].s = 0.0; acc[j].c = 0.0; }
    for (int y = 0; y < ny; y++) {
      const double *r = f + pl * k + (size_t)ld * y;
      for (int x = 0; x < nx; x++) { row[x].re = r[x] / (double)nx; row[x].im = 0.0; }
      fft_exec(&c->sm.px, row, t, -1);
      csum_add(&acc[0], 0.5 * row[0].re * row[0].re);
      for (int j = 1; j < nk; j++) csum_add(&acc[j], row[j].re * row[j].re + row[j].im * row[j].im);
    }
    for (int j = 0; j < nk; j++) spec[(size_t)k * nk + j] = csum_val(&acc[j]) / (double)ny;
  }
  free(row); free(t); free(acc);
}

static double plane_mean(const or_ctx *c, const double *a, int stride) {
  csum_t s = { 0.0, 0.0 };
  for (int y = 0; y < c->ny; y++)
    for (int x = 0; x < c->nx; x++) csum_add(&s, a[x + (size_t)stride * y]);
  return csum_val(&s) / (double)(c->nx * c->ny);
}

void or_avg_stats(or_ctx *c, const double *u, const double *v, const double *w, const double *txz, double *stats) {
  const int nx = c->nx, ny = c->ny, ld = c->ld, nl = c->nz - 1, nk = nx / 2;
  const size_t pl = (size_t)ld * ny;
  double *ub = stats, *vb = stats + nl, *wb = stats + 2 * nl;
  double *u2 = stats + 3 * nl, *v2 = stats + 4 * nl, *w2 = stats + 5 * nl;
  double *uw = stats + 6 * nl, *tb = stats + 7 * nl, *st = stats + 8 * nl, *eoz = stats + 9 * nl;
  for (int k = 0; k < nl; k++) {                                             /* :1725-1727 */
    ub[k] = plane_mean(c, u + pl * k, ld);
    vb[k] = plane_mean(c, v + pl * k, ld);
    wb[k] = plane_mean(c, w + pl * k, ld);
    tb[k] = plane_mean(c, txz + pl * k, ld);                                 /* :1749 */
  }
  for (int k = 0; k < nl; k++) {
    csum_t s2u = { 0, 0 }, s2v = { 0, 0 }, s2w = { 0, 0 }, suw = { 0, 0 };
    if (k > 0)                                                               /* :1729-1742, zero at jz = 1 */
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          const size_t i = pl * k + x + (size_t)ld * y;
          const double up = 0.5 * (u[i] - ub[k] + u[i - pl] - ub[k - 1]);
          const double vp = 0.5 * (v[i] - vb[k] + v[i - pl] - vb[k - 1]);
          const double wp = w[i] - wb[k];
          csum_add(&s2u, up * up); csum_add(&s2v, vp * vp); csum_add(&s2w, wp * wp); csum_add(&suw, up * wp);
        }
    const double n = (double)(nx * ny);
    u2[k] = csum_val(&s2u) / n; v2[k] = csum_val(&s2v) / n; w2[k] = csum_val(&s2w) / n;  /* :1744-1746 */
    uw[k] = csum_val(&suw) / n;                                              /* :1748 */
    st[k] = uw[k] + tb[k];                                                   /* :1750 */
  }
  or_spectrum(c, u, nl, eoz);                                                /* :1753 */
  for (int k = 0; k < nl; k++) {
    const double z = ((double)(k + 1) - 0.5) * c->dz;                        /* :1754 */
    for (int j = 0; j < nk; j++) eoz[(size_t)k * nk + j] = eoz[(size_t)k * nk + j] / z;
  }
}

void or_surface_stats(or_ctx *c, const double *u, const double *v, const double *txz, const double *tyz,
                      const double *psi_m, const double *psi_m0, const double *ustar_avg, const double *theta,
                      const double *qmix, const double *sgs_t3, const double *sgs_q3, double out[8]) {
  const int nx = c->nx, ny = c->ny, ld = c->ld;
  const double n = (double)(float)(nx * ny);
  csum_t su = { 0, 0 };
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      const size_t i = x + (size_t)ld * y;
      csum_add(&su, sqrt(u[i] * u[i] + v[i] * v[i]));
    }
  const double mpsi = plane_mean(c, psi_m, nx), mpsi0 = plane_mean(c, psi_m0, nx);
  /* zo_ = exp(mean(log(zo))) = zo for the uniform roughness (bottombc.f90:53) */
  const double ustar_ = csum_val(&su) / n * VONK / (log(0.5 * c->dz / c->P.z0) - mpsi + mpsi0);   /* :557-558 */
  const double mtx = plane_mean(c, txz, ld), mty = plane_mean(c, tyz, ld);
  out[0] = ustar_;
  out[1] = plane_mean(c, ustar_avg, nx);                                     /* :655 */
  out[2] = pow(mtx * mtx + mty * mty, 0.25);                                 /* :656-657 */
  out[3] = out[4] = out[5] = out[6] = out[7] = 0.0;
  if (theta && qmix && sgs_t3 && sgs_q3) {
    const or_scalar_params *S = &c->S;
    csum_t stv = { 0, 0 }, swtv = { 0, 0 };
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = x + (size_t)ld * y;
        csum_add(&stv, theta[i] * (1 + 0.61 * 0.001 * qmix[i] * S->q_scale));               /* :547 */
        csum_add(&swtv, sgs_t3[i] + 0.61 * theta[i] * (0.001 * S->q_scale * sgs_q3[i]));    /* :549-550 */
      }
    const double theta_ = plane_mean(c, theta, ld), thetav_ = csum_val(&stv) / n;         /* :560-561 */
    const double wt_ = plane_mean(c, sgs_t3, ld), wtv_ = csum_val(&swtv) / n;             /* :554-555 */
    const double g_ = (S->g * S->z_i) / (S->u_star * S->u_star);                          /* :545 */
    out[3] = theta_;
    out[4] = thetav_;
    out[5] = plane_mean(c, qmix, ld);                                                      /* :660 */
    out[6] = wt_;
    out[7] = wtv_ != 0.0 ? -(ustar_ * ustar_ * ustar_) / (VONK * (g_ / theta_) * wtv_) : 0.0;  /* :661 */
  }
}
