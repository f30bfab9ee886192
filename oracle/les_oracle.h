/*
 * les_oracle.h -- CPU FP64 restatement of the KhaledGhannam/LES per-timestep
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the
 * HIP product path is compared against (tests/, __graft_entry__.smoke(),
 * bench.py's cpu_baseline leg).  Nothing in les_amd/ links or calls it.
 *
 * Array layout follows the reference exactly: real (ld, ny, nlev) Fortran
 * column-major arrays, x fastest, ld = 2*(nx/2+1) (two pad columns that
 * hold the Nyquist entry in spectral form), level index k = jz - lbz.
 * Arrays dimensioned (ld,ny,1:nz) have nz levels; p is (ld,ny,0:nz).
 *
 * Spectral planes use the FFTW-2 in-place convention of fft.f90:35-47:
 * complex (lh, ny), lh = nx/2+1, unnormalised, forward sign -1.
 * After any inverse transform the pad columns are defined to be 0.
 */
#ifndef LES_ORACLE_H
#define LES_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

/* Grid + run constants.  Mirrors param.f90:56-65,98,100,115,144-151. */
typedef struct or_params {
  int nx, ny, nz;        /* nz counts w-levels incl. the lid: levels 1..nz */
  double Lx, Ly;         /* nondimensional horizontal extent (L_x/z_i) */
  double Lz;             /* nondimensional height of the domain (L_z/z_i) */
  double dt;             /* nondimensional time step */
  double z0;             /* roughness length / z_i (zo, bottombc.f90:53) */
  double force;          /* mean pressure gradient forcing (0 disables) */
  int    inilag;         /* lagrange_Sdep F_* initialisation at jt==cs_count */
  int    initu;          /* 1: RHS_f read from IC (AB2 with given history) */
  int    cs_count;       /* steps between dynamic updates (param.f90:100) */
} or_params;

or_ctx *or_create(const or_params *p);
void    or_destroy(or_ctx *c);
/* copies of the wavenumber and filter tables (lh*ny each) */
void or_get_tables(const or_ctx *c, double *kx, double *ky, double *k2,
                   double *G_test, double *G_test_test);

/* ---- FFT layer (FFTW-2 rfftwnd in-place semantics) ---- */
void or_rfft2_fwd(or_ctx *c, double *plane, int big);   /* (ld,ny) real -> (lh,ny) complex */
void or_rfft2_inv(or_ctx *c, double *plane, int big);   /* (lh,ny) complex -> (ld,ny) real, pads=0 */
/* plain 1-D complex DFT of length n (any 2^a 3^b), sign -1 or +1, in place */
void or_cfft1(int n, double *re_im, int sign);

/* ---- spectral operators: filt_da.f90, ddx.f90, ddy.f90, ddxy.f90 ---- */
void or_filt_da(or_ctx *c, double *f, double *dfdx, double *dfdy, int nlev);
void or_ddx(or_ctx *c, double *dfdx, const double *f, int nlev);
void or_ddy(or_ctx *c, double *dfdy, const double *f, int nlev);
void or_ddxy(or_ctx *c, double *dfdx, double *dfdy, const double *f, int nlev);
void or_test_filter_init(or_ctx *c, double alpha, double *G);
void or_test_filter(or_ctx *c, double *plane, const double *G);
void or_padd(or_ctx *c, double *big_c, const double *small_c);
void or_unpadd(or_ctx *c, double *small_c, const double *big_c);
void or_dealias1(or_ctx *c, const double *u, double *u_big, int nlev);
void or_dealias2(or_ctx *c, double *u, double *u_big, int nlev);

/* ---- vertical FD: ddz_uv.f90, ddz_w.f90 (nz levels) ---- */
void or_ddz_uv(or_ctx *c, double *dfdz, const double *f);
void or_ddz_w(or_ctx *c, double *dfdz, const double *f);

/* ---- convec.f90 ---- */
void or_convec(or_ctx *c, const double *u, const double *v, const double *w,
               const double *dudy, const double *dudz, const double *dvdx,
               const double *dvdz, const double *dwdx, const double *dwdy,
               double *cx, double *cy, double *cz);

/* ---- wallstress.f90 (lbc_mom='wall', patch_flag=1) ---- */
void or_wallstress(or_ctx *c, const double *u, const double *v,
                   double *dudz, double *dvdz, double *txz, double *tyz,
                   const double *psi_m, const double *psi_m0,
                   const double *phi_m, double *ustar_avg);

/* ---- sgs_stag.f90 + calc_Sij + lagrange_Sdep + interpolag_Sdep ---- */
typedef struct or_sgs_state {
  double *F_LM, *F_MM, *F_QN, *F_NN;      /* (ld,ny,nz) */
  double *Beta, *Cs_opt2, *Nu_t;          /* (ld,ny,nz) */
  double *u_lag, *v_lag, *w_lag;          /* (ld,ny,nz) */
  double *TnLMMM, *TnQNNN;                /* (ld,ny,nz) */
  double *L11t, *L22t, *L33t, *Q11t, *Q22t, *Q33t;  /* (ld,ny,nz) */
  double *Beta_avg, *Betaclip_avg;        /* (nz) */
} or_sgs_state;

void or_calc_sij(or_ctx *c, const double *const grad[9], double *const S[6]);
void or_interpolag_sdep(or_ctx *c, or_sgs_state *s);
void or_lagrange_sdep(or_ctx *c, const double *const S[6], const double *u,
                      const double *v, const double *w, or_sgs_state *s);
/* grad order: dudx,dudy,dudz,dvdx,dvdy,dvdz,dwdx,dwdy,dwdz
   tau order : txx,txy,txz,tyy,tyz,tzz */
void or_sgs_stag(or_ctx *c, const double *const grad[9], const double *u,
                 const double *v, const double *w, double *const tau[6],
                 or_sgs_state *s);

/* ---- divstress_uv.f90, divstress_w.f90 ---- */
void or_divstress_uv(or_ctx *c, double *divt, const double *tx, const double *ty,
                     const double *tz);
void or_divstress_w(or_ctx *c, double *divt, const double *tx, const double *ty,
                    const double *tz);

/* ---- press_stag_array.f90 + tridag_array.f90 ---- */
/* p: (ld,ny,0:nz) -> nz+1 levels; dpdx,dpdy: (ld,ny,nz) */
void or_press_stag_array(or_ctx *c, const double *u, const double *v,
                         const double *w, const double *divtz, double *p,
                         double *dpdx, double *dpdy);
/* a,b,c: (lh,ny,n) real; r,u: (lh,ny,n) complex interleaved; n = nz+1 */
int  or_tridag_array(or_ctx *c, const double *a, const double *b,
                     const double *cc, const double *r, double *u);

/* ---- forcing.f90:128-219 project, rmsdiv.f90, scalars_module2.f90:22-70 ---- */
void or_project(or_ctx *c, double *u, double *v, double *w, const double *dpdx,
                const double *dpdy, const double *dpdz);
double or_rmsdiv(or_ctx *c, const double *dudx, const double *dvdy,
                 const double *dwdz);
void or_timestep_conditions(or_ctx *c, const double *u, const double *v,
                            const double *w, const double *Nu_t, double out[4]);

/* ---- whole step (main.f90:202-415, neutral model=5) ---- */
typedef struct or_state {
  double *u, *v, *w;                       /* (ld,ny,nz) */
  double *grad[9];                         /* dudx..dwdz (ld,ny,nz) */
  double *RHSx, *RHSy, *RHSz, *RHSx_f, *RHSy_f, *RHSz_f;
  double *tau[6];                          /* txx,txy,txz,tyy,tyz,tzz */
  double *divtx, *divty, *divtz;
  double *p;                               /* (ld,ny,nz+1) */
  double *dpdx, *dpdy, *dpdz;
  double *psi_m, *psi_m0, *phi_m, *ustar_avg;  /* (nx,ny) */
  or_sgs_state sgs;
  int jt, jt_total;
} or_state;

or_state *or_state_create(or_ctx *c);
void      or_state_destroy(or_state *s);
/* seeded log-law + uniform perturbation initial condition */
void or_state_init_ic(or_ctx *c, or_state *s, uint64_t seed, double amp,
                      double ustar, double kappa_ic);
void or_step(or_ctx *c, or_state *s);

/* jt / jt_total seen by routine-level calls (param.f90:147-148) and the
   lagrange_Sdep SAVE init flags (lagrange_Sdep.f90:67-68) */
void or_set_time(or_ctx *c, int jt, int jt_total);
void or_set_lag_init(or_ctx *c, int flm_init, int fqn_init);

/* counter-based uniform in [0,1): splitmix64 of (seed, index) */
double or_uniform(uint64_t seed, uint64_t index);

#ifdef __cplusplus
}
#endif
#endif
