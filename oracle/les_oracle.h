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

/* ---- scalars (S_FLAG=.true.): scalars_module.f90, scalars_module_q.f90,
   lagrange_scalar_Sdep.f90, interpolag_scalar_Sdep.f90, bottombc.f90 ---- */
/* param.f90:53,72-90,139-146,174 and the lbc=1 / patch_flag=1 surface of
   bottombc.f90:26-63 */
typedef struct or_scalar_params {
  int    enable;     /* S_FLAG: theta_all_in_one + q_all_in_one every step */
  int    model;      /* 7: Pr_ = Pr_t from lagrange_scalar_Sdep; else Pr_ = Pr */
  int    initsc;     /* scalar history given: no Euler scalar step at jt==1 */
  int    passive;    /* passive_scalar: no buoyancy in RHSz */
  double Pr;         /* 0.55 */
  double z_i, u_star, T_scale, q_scale, g;
  double dTdz_top;   /* K/m at the lid */
  double dqdz_top;   /* (g/kg)/m at the lid */
} or_scalar_params;

/* wt_s (K m/s) and wq_s ((g/kg) m/s): (nx,ny) kinematic surface fluxes,
   i.e. after the unit conversion of bottombc.f90:57-58.  z_os = z0/7.39. */
void or_set_scalars(or_ctx *c, const or_scalar_params *p, const double *wt_s, const double *wq_s);
/* lagrange_scalar_Sdep.f90:65-66 SAVE flags */
void or_set_scalar_lag_init(or_ctx *c, int glm_init, int gqn_init);

/* Scalar Lagrangian SGS state, sgsmodule.f90 (G_*, u_lag_s, Pr_t) plus the
   surface functions the scalar model reads (bottombc phi_h, psi_m). */
typedef struct or_scalar_sgs {
  double *G_LM, *G_MM, *G_QN, *G_NN;       /* (ld,ny,nz) */
  double *u_lag_s, *v_lag_s, *w_lag_s;     /* (ld,ny,nz) */
  double *Pr_t, *Cs2_Pr;                   /* (ld,ny,nz) */
  double *phi_h, *psi_m;                   /* (nx,ny) */
} or_scalar_sgs;

/* obukhov, scalars_module.f90:508-690.  theta/qmix: (ld,ny,nz) (level 1 read);
   sgs_t3/sgs_q3: the level-1 plane (ld,ny); the rest (nx,ny). */
void or_obukhov(or_ctx *c, const double *u, const double *v, const double *theta,
                const double *qmix, double *sgs_t3, double *sgs_q3, double *psi_m,
                double *psi_m0, double *psi_h, double *psi_h0, double *phi_m,
                double *phi_h, double *ustar_avg);
/* calcbeta, scalars_module.f90:175-228 -> beta_scal (ld,ny,nz) */
void or_calcbeta(or_ctx *c, const double *theta, const double *qmix, double *beta_scal);
/* scalar_RHS_calc, scalars_module.f90:232-454 (which=0, theta) and
   scalars_module_q.f90:114-330 (which=1, q); lbc=1.  dsdz level 1 is
   overwritten with the surface gradient; sgs_vert1 = level-1 plane of
   sgs_t3/sgs_q3 (ld,ny). */
void or_scalar_rhs_calc(or_ctx *c, int which, const double *dsdx, const double *dsdy,
                        double *dsdz, const double *u, const double *v, const double *w,
                        const double *Nu_t, const double *Pr_, const double *ustar_avg,
                        const double *phi_h, double *RHS, double *sgs_vert1);
/* step_scalar, scalars_module.f90:457-504 / step_scalar_q */
void or_step_scalar(or_ctx *c, int which, double *scalar, const double *RHS, const double *RHS_f);
/* interpolag_scalar_Sdep.f90:1-316 */
void or_interpolag_scalar_sdep(or_ctx *c, or_scalar_sgs *g);
/* lagrange_scalar_Sdep.f90:1-467; theta is Nyquist-filtered in place
   (filt_da at :90); reads TnLMMM, TnQNNN, Cs_opt2 of m, overwrites m->Beta */
void or_lagrange_scalar_sdep(or_ctx *c, const double *const S[6], const double *u,
                             const double *v, const double *w, double *theta,
                             or_sgs_state *m, or_scalar_sgs *g);
/* sgs_stag with model=7 (sgs_stag.f90:59-69,150-153,166-173) */
void or_sgs_stag_scalar(or_ctx *c, const double *const grad[9], const double *u,
                        const double *v, const double *w, double *theta,
                        double *const tau[6], or_sgs_state *s, or_scalar_sgs *g);

/* ---- scaledep_dynamic.f90:4-255 + rtnewt (sgsmodule.f90:37-59), model=3.
   S[0..5] = the dummy arguments S11..S33; sgs_stag.f90:132 binds them to
   (u, v, w, S11, S12, S13) (argument mismatch kept).  Cs_opt2, beta: (nz). */
double or_rtnewt(const double A[6]);
void or_scaledep_dynamic(or_ctx *c, const double *const S[6], const double *u,
                         const double *v, const double *w, double *Cs_opt2,
                         double *beta);

void or_sgs_stag_model3(or_ctx *c, const double *const grad[9], const double *u,
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
  /* scalars (allocated always, used when or_set_scalars enabled them) */
  double *theta, *qmix;                    /* (ld,ny,nz) */
  double *dTdx, *dTdy, *dTdz, *dqdx, *dqdy, *dqdz;
  double *RHS_T, *RHS_Tf, *RHS_q, *RHS_qf, *beta_scal;
  double *sgs_t3, *sgs_q3;                 /* level-1 plane (ld,ny) */
  double *psi_h, *psi_h0, *phi_h;          /* (nx,ny) */
  or_scalar_sgs ssgs;
} or_state;

or_state *or_state_create(or_ctx *c);
void      or_state_destroy(or_state *s);
/* seeded log-law + uniform perturbation initial condition */
void or_state_init_ic(or_ctx *c, or_state *s, uint64_t seed, double amp,
                      double ustar, double kappa_ic);
void or_step(or_ctx *c, or_state *s);
/* theta (K), inversion (K/m above z_inv, z in units of z_i), uniform
   perturbation amplitude amp_K where z < z_inv, humidity q_init (g/kg) */
void or_state_init_scalar_ic(or_ctx *c, or_state *s, uint64_t seed, double T_init,
                             double z_inv, double inv_strength, double amp_K, double q_init);

/* jt / jt_total seen by routine-level calls (param.f90:147-148) and the
   lagrange_Sdep SAVE init flags (lagrange_Sdep.f90:67-68) */
void or_set_time(or_ctx *c, int jt, int jt_total);
/* SGS model of or_step (param.f90 'model'): 1 Smagorinsky with Mason's wall
   damping, 2 standard dynamic, 3 plane-averaged scale-dependent, 4 Lagrangian
   scale-invariant, 5 Lagrangian scale-dependent (default) */
void or_set_sgs_model(or_ctx *c, int model);
void or_set_lag_init(or_ctx *c, int flm_init, int fqn_init);

/* Coriolis forcing and the top sponge (param.f90:68-75,220, main.f90:131-138,
   319-348, topbc.f90:14-95).  ug0/vg0 in m/s, u_star in m/s, z_i in m; the
   domain height is Lz z_i. */
typedef struct or_forcing {
  int coriolis;          /* coriolis_forcing */
  double ug0, vg0;       /* geostrophic wind (m/s) */
  double u_star, z_i;
  int ubc;               /* 1: sponge layer */
  int damping_method;    /* 1: relax w in the top quarter; 2: cosine profile above 0.75 L_z */
} or_forcing;
void or_set_forcing(or_ctx *c, const or_forcing *f);
/* setsponge's profile sponge(1..nz) into sp[0..nz-1] (zeros when ubc != 1) */
void or_get_sponge(const or_ctx *c, double *sp);

/* std_dynamic.f90:4-140: Cs_opt2(nz) of the standard dynamic model */
void or_std_dynamic(or_ctx *c, const double *const S[6], const double *u, const double *v,
                    const double *w, double *Cs_opt2);
/* interpolag_Ssim.f90 and lagrange_Ssim.f90:4-300 (model 4, F_LM and F_MM) */
void or_interpolag_ssim(or_ctx *c, or_sgs_state *s);
void or_lagrange_ssim(or_ctx *c, const double *const S[6], const double *u,
                      const double *v, const double *w, or_sgs_state *s);
/* new_scaledep_dynamic.f90:1-223 (model 6): Cs_opt2(nz) from the plane sums
   of L:M, M:M, Q:N, N:N with single-precision clips; S = calc_Sij's S11..S33 */
void or_new_scaledep_dynamic(or_ctx *c, const double *const S[6], const double *u,
                             const double *v, const double *w, double *Cs_opt2);
/* scalar_scaledep_dynamic.f90:1-307 (model 6 with scalars): filters theta in
   place (its filt_da), writes Cs2_Pr(nz-1) and Pr_t (ld,ny,nz) levels
   1..nz-1 = Cs_opt2 / Cs2_Pr; psi_m, phi_h: (nx,ny) */
void or_scalar_scaledep_dynamic(or_ctx *c, const double *const S[6], const double *u,
                                const double *v, const double *w, double *theta,
                                const double *Cs_opt2, const double *psi_m, const double *phi_h,
                                double *Cs2_Pr, double *Pr_t);
/* sgs_stag.f90:3-366 with any model 1..6 (neutral) */
void or_sgs_stag_model(or_ctx *c, const double *const grad[9], const double *u,
                       const double *v, const double *w, double *const tau[6],
                       or_sgs_state *s, int model);

/* ---- statistics (all plane sums compensated) ----
   spectrum, io.f90:1613-1646: spec (nlev, nx/2), spec(k) = <|c_k|^2>_y with
   c_k the x coefficients of f/nx, the mean (k = 0) counted with 1/2 */
void or_spectrum(or_ctx *c, const double *f, int nlev, double *spec);
/* avg_stats, io.f90:1722-1755 (current statistics of levels 1..nz-1, nl =
   nz-1): stats = [ubar | vbar | wbar | <u'^2> | <v'^2> | <w'^2> | <u'w'> |
   <txz> | <u'w'> + <txz>] (nl each) then Eozbar (nl, nx/2) = spectrum(u)/z */
void or_avg_stats(or_ctx *c, const double *u, const double *v, const double *w, const double *txz, double *stats);
/* obukhov's plane averages, scalars_module.f90:545-561,655-661: out = [ustar_,
   ustar_avg_time, ustar_time, theta_, thetav_, q_time, wt_, L_]; theta, qmix,
   sgs_t3, sgs_q3 may be NULL (neutral: the scalar entries and L_ are 0) */
void or_surface_stats(or_ctx *c, const double *u, const double *v, const double *txz, const double *tyz,
                      const double *psi_m, const double *psi_m0, const double *ustar_avg, const double *theta,
                      const double *qmix, const double *sgs_t3, const double *sgs_q3, double out[8]);

/* OpenMP threads of the 2-D transforms (results do not depend on them) */
int  or_threads(void);
void or_set_threads(int n);

/* counter-based uniform in [0,1): splitmix64 of (seed, index) */
double or_uniform(uint64_t seed, uint64_t index);

#ifdef __cplusplus
}
#endif

/* ---- learned SGS closure reference (les_mlp_oracle.c): 8 -> 64 -> 64 -> 64
   -> 2 MLP with SiLU; theta flat W1 b1 W2 b2 W3 b3 W4 b4 ([out][in]); x
   (P,8), y/t (P,2); bf16 != 0 rounds operands where the GPU's MFMA does ---- */
#define OR_MLP_NIN 8
#define OR_MLP_NH 64
#define OR_MLP_NOUT 2
int    or_mlp_nparam(void);
void   or_mlp_forward(const double *theta, const double *x, long P, double *y, int bf16);
double or_mlp_loss_grad(const double *theta, const double *x, const double *t, long P, double *grad, int bf16);

#endif
