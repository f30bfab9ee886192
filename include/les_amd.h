/*
 * les_amd.h -- C ABI of the MI355X (gfx950) LES hot-path library
 * (les_amd/libles_amd.so).
 *
 * Each routine-level entry point is a drop-in for one external subroutine of
 * the KhaledGhannam/LES reference (implicit-interface Fortran, arrays passed
 * by base address).  Arrays are DEVICE pointers in the reference's layout:
 * real (ld, ny, nlev) column-major, x fastest, ld = 2*(nx/2+1); arrays
 * dimensioned (ld,ny,1:nz) hold nz levels, p holds nz+1 levels (0:nz).
 * The module globals the Fortran routines read (grid, kx/ky, G_test,
 * G_test_test, dt, jt, jt_total, the lagrange_Sdep SAVE flags) live in the
 * les_ctx.  All calls enqueue on the context's stream and return 0 on
 * success; a non-zero status means les_last_error() holds the message (the
 * Fortran shim maps it to `error stop`, where the reference used `stop`).
 *
 * Suffix _f32: the same routine on float arrays (FP32-bulk mode).
 */
#ifndef LES_AMD_H
#define LES_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct les_ctx les_ctx;

/* Grid + run constants: param.f90:56-65 (grid), :100 (cs_count), :144-145
   (mean pressure forcing), :79 (initu, inilag), param.f90:219 (dt),
   bottombc.f90:53 (zo = zom/z_i). */
typedef struct les_grid {
  int nx, ny, nz;
  double Lx, Ly, Lz; /* nondimensional (L/z_i) */
  double dt, z0, force;
  int inilag, initu, cs_count;
} les_grid;

/* Lagrangian SGS state, sgsmodule.f90:302-326 / sim_param.f90:289 (device
   pointers; (ld,ny,nz) in the call's precision, Beta_avg/Betaclip_avg (nz)
   always double). */
typedef struct les_sgs_fields {
  void *F_LM, *F_MM, *F_QN, *F_NN, *Beta, *Cs_opt2, *Nu_t, *u_lag, *v_lag, *w_lag;
  void *TnLMMM, *TnQNNN, *L11t, *L22t, *L33t, *Q11t, *Q22t, *Q33t;
  double *Beta_avg, *Betaclip_avg;
} les_sgs_fields;

/* ---- context (replaces init_fft fft.f90:35-51, init_wavenumber
   fft.f90:54-81 and the two test_filter_init calls main.f90:166-170).
   Grids: nx, ny multiples of 8 of the form 2^a 3^b up to 1024, nz >= 3, and
   (nx + 2) ny (nz + 1) < 2^31 points per array; returns 2 otherwise. ---- */
int les_create(const les_grid *g, les_ctx **out);
void les_destroy(les_ctx *c);
const char *les_last_error(les_ctx *c);
int les_set_stream(les_ctx *c, void *hip_stream);
int les_get_grid(les_ctx *c, les_grid *out);
int les_sync(les_ctx *c);
int les_set_time(les_ctx *c, int jt, int jt_total);        /* param.f90:147-148 */
int les_get_time(les_ctx *c, int *jt, int *jt_total);
int les_set_lag_init(les_ctx *c, int flm_init, int fqn_init); /* lagrange_Sdep.f90:67-68 */
/* host copies of kx, ky, k2 (fft.f90:19), G_test, G_test_test
   (test_filtermodule.f90:9), each (lh, ny); any pointer may be NULL */
int les_get_tables(les_ctx *c, double *kx, double *ky, double *k2, double *G_test, double *G_test_test);

/* ---- horizontal spectral operators ---- */
/* filt_da.f90:3-42  filt_da(f, dfdx, dfdy): f is overwritten (Nyquist killed) */
int les_filt_da(les_ctx *c, double *f, double *dfdx, double *dfdy, int nlev);
/* ddx.f90:1-31, ddy.f90:1-31: f untouched */
int les_ddx(les_ctx *c, double *dfdx, const double *f, int nlev);
int les_ddy(les_ctx *c, double *dfdy, const double *f, int nlev);
/* ddxy.f90:1-29 */
int les_ddxy(les_ctx *c, double *dfdx, double *dfdy, const double *f, int nlev);
/* test_filtermodule.f90:32-63: G (lh,ny) device array */
int les_test_filter_init(les_ctx *c, double alpha, double *G);
/* test_filtermodule.f90:13-30: one (ld,ny) plane filtered in place */
int les_test_filter(les_ctx *c, double *plane, const double *G);
/* padd.f90:1-26 / unpadd.f90:1-22: complex (lh,ny) <-> (lh_big,ny2) planes */
int les_padd(les_ctx *c, double *u_big, const double *u);
int les_unpadd(les_ctx *c, double *cc, const double *cc_big);
/* dealias1.f90:1-31 (u untouched) / dealias2.f90:1-33 */
int les_dealias1(les_ctx *c, const double *u, double *u_big, int nlev);
int les_dealias2(les_ctx *c, double *u, double *u_big, int nlev);

/* ---- vertical FD: ddz_uv.f90:5-57, ddz_w.f90:6-51 (nz levels) ---- */
int les_ddz_uv(les_ctx *c, double *dfdz, const double *f);
int les_ddz_w(les_ctx *c, double *dfdz, const double *f);

/* ---- convec.f90:1-248  convec(cx,cy,cz) with its sim_param inputs ---- */
int les_convec(les_ctx *c, const double *u, const double *v, const double *w, const double *dudy,
               const double *dudz, const double *dvdx, const double *dvdz, const double *dwdx,
               const double *dwdy, double *cx, double *cy, double *cz);

/* ---- wallstress.f90:5-69 ('wall'); psi_m, psi_m0, phi_m, ustar_avg (nx,ny) ---- */
int les_wallstress(les_ctx *c, const double *u, const double *v, double *dudz, double *dvdz, double *txz,
                   double *tyz, const double *psi_m, const double *psi_m0, const double *phi_m,
                   double *ustar_avg);

/* ---- SGS: sgs_stag.f90:3-500 (model 5), lagrange_Sdep.f90:4-478,
   interpolag_Sdep.f90:2-380.  grad: dudx,dudy,dudz,dvdx,dvdy,dvdz,dwdx,dwdy,dwdz;
   tau: txx,txy,txz,tyy,tyz,tzz; S: S11,S12,S13,S22,S23,S33 ---- */
int les_calc_sij(les_ctx *c, double *const grad[9], double *const S[6]);
int les_sgs_stag(les_ctx *c, double *const grad[9], const double *u, const double *v, const double *w,
                 double *const tau[6], const les_sgs_fields *s);
int les_lagrange_sdep(les_ctx *c, double *const S[6], const double *u, const double *v, const double *w,
                      const les_sgs_fields *s);
int les_interpolag_sdep(les_ctx *c, const les_sgs_fields *s);

/* ---- SGS model selection, param.f90 'model' (sgs_stag.f90:56-175):
   1 = Smagorinsky, Cs_opt2 = Co^2 = 0.16^2 every step with Mason's
       wall-damped length l(jz) (sgs_stag.f90:73-104);
   2 = standard dynamic, plane-averaged (std_dynamic.f90);
   3 = scale-dependent dynamic with plane-averaged coefficients
       (scaledep_dynamic.f90);
   4 = Lagrangian scale-invariant (lagrange_Ssim.f90, interpolag_Ssim.f90);
   5 = Lagrangian scale-dependent (default);
   6 = scale-dependent dynamic with plane-averaged coefficients and no
       quintic (new_scaledep_dynamic.f90), with scalars also Pr_t of
       scalar_scaledep_dynamic.f90 (:143-149).
   u_lag accumulates under models 4 and 5 only; the jt_total < 500 constant-Cs
   period applies to models 5 and 6 (:165-174).  The scalar Lagrangian model
   (scalar model 7) needs SGS model 5, and scalar model 6 goes with SGS model
   6 and only with it; with models 1-4 the scalars use the constant Pr
   (scalars_module.f90:77-84).  Returns 1 for another model. ---- */
int les_set_sgs_model(les_ctx *c, int model);
/* ---- std_dynamic.f90:4-140: Cs_opt2 (nz device doubles) of the standard
   dynamic model from S11..S33 and u, v, w (ld,ny,nz), Delta = (dx dy dz)^(1/3) ---- */
int les_std_dynamic(les_ctx *c, double *const S[6], const double *u, const double *v, const double *w,
                    double *Cs_opt2);
/* ---- new_scaledep_dynamic.f90:1-223 (model 6): Cs_opt2 (nz device doubles)
   from per-level plane sums of L:M, M:M, Q:N, N:N with the single-precision
   clips of :208-220; S: calc_Sij's S11..S33, u, v, w (ld,ny,nz) ---- */
int les_new_scaledep_dynamic(les_ctx *c, double *const S[6], const double *u, const double *v, const double *w,
                             double *Cs_opt2);
/* ---- lagrange_Ssim.f90:4-300 (calls interpolag_Ssim first) and
   interpolag_Ssim.f90 on the F_LM, F_MM, Beta, Cs_opt2, u_lag.. members of s;
   Beta_avg and Betaclip_avg receive nx*ny, the plane sums of Beta = 1 ---- */
int les_lagrange_ssim(les_ctx *c, double *const S[6], const double *u, const double *v, const double *w,
                      const les_sgs_fields *s);
int les_interpolag_ssim(les_ctx *c, const les_sgs_fields *s);

/* ---- Coriolis forcing and the top sponge layer (param.f90:68-75,220,
   main.f90:131-138,172-185,319-348, topbc.f90:14-95, scalars_module.f90:143-148,
   scalars_module_q.f90:116-121).  With coriolis_forcing the step adds
   coriol (v - vg) to RHSx and -coriol (u - ug) to RHSy, coriol =
   1.3962634e-4 z_i/u_star, ug = ug0/u_star, vg = vg0/u_star.  With ubc = 1
   the sponge of damping_method 1 relaxes w in the top quarter, that of
   damping_method 2 relaxes u, v, w (and theta, q when scalars run) to their
   plane means above 0.75 L_z, L_z = Lz z_i.  Returns 1 on an invalid switch. ---- */
typedef struct les_forcing {
  int coriolis_forcing;
  double ug0, vg0;      /* geostrophic wind, m/s */
  double u_star, z_i;   /* m/s, m */
  int ubc;              /* 0: no sponge, 1: sponge */
  int damping_method;   /* 1 or 2 */
} les_forcing;
int les_set_forcing(les_ctx *c, const les_forcing *f);
/* setsponge's profile sponge(1..nz) (host doubles, nz) */
int les_get_sponge(les_ctx *c, double *sponge);
/* ---- scaledep_dynamic.f90:4-255 with rtnewt (sgsmodule.f90:37-59).
   S: the routine's dummy arguments S11, S12, S13, S22, S23, S33 (ld,ny,nz);
   sgs_stag.f90:132 binds them to u, v, w, S11, S12, S13.  u, v, w are
   sim_param's.  Cs_opt2: (nz) device doubles; beta: (nz) device doubles or
   NULL (the routine's per-level root, which it writes to output/beta). ---- */
int les_scaledep_dynamic(les_ctx *c, double *const S[6], const double *u, const double *v, const double *w,
                         double *Cs_opt2, double *beta);

/* ---- scalars: potential temperature theta and humidity q (S_FLAG) ----
   Replaces the scalar parameters of param.f90:53,72-90,139-146,174 and the
   surface set-up of bottombc.f90:26-63 (lbc=1, patch_flag=1; z_os = z0/7.39).
   Must precede les_state_alloc for the state to carry the scalars. */
typedef struct les_scalar_params {
  int enable;      /* S_FLAG: theta_all_in_one + q_all_in_one every step */
  int model;       /* 7: Pr_ = Pr_t of lagrange_scalar_Sdep (SGS model 5); 6: Pr_t of
                      scalar_scaledep_dynamic (SGS model 6); otherwise Pr */
  int initsc;      /* scalar history given (no Euler scalar step at jt==1) */
  int passive;     /* passive_scalar: no buoyancy in RHSz */
  double Pr, z_i, u_star, T_scale, q_scale, g;
  double dTdz_top; /* K/m at the lid */
  double dqdz_top; /* (g/kg)/m at the lid */
} les_scalar_params;
/* wt_s (K m/s), wq_s ((g/kg) m/s): host (nx,ny) kinematic surface fluxes after
   the conversion of bottombc.f90:57-58; NULL means 0 */
int les_set_scalars(les_ctx *c, const les_scalar_params *p, const double *wt_s, const double *wq_s);
int les_set_scalar_lag_init(les_ctx *c, int glm_init, int gqn_init); /* lagrange_scalar_Sdep.f90:65-66 */
/* FP32-bulk state only (no reference counterpart): theta and qmix are stored
   as deviations from theta_ref = T_init/T_scale and q_ref = q_init/q_scale of
   les_state_init_scalar_ic; both are 0 in FP64 mode */
int les_scalar_ref(les_ctx *c, double *theta_ref, double *q_ref);

/* sgsmodule.f90 G_LM..G_NN, u_lag_s..w_lag_s, Pr_t (ld,ny,nz) and the
   Cs2_Pr output of lagrange_scalar_Sdep; phi_h, psi_m (nx,ny) from bottombc */
typedef struct les_scalar_sgs_fields {
  void *G_LM, *G_MM, *G_QN, *G_NN, *u_lag_s, *v_lag_s, *w_lag_s, *Pr_t, *Cs2_Pr;
  const double *phi_h, *psi_m;
} les_scalar_sgs_fields;

/* obukhov, scalars_module.f90:508-690: theta/qmix (ld,ny,nz) (level 1 read);
   sgs_t3/sgs_q3 are the level-1 planes (ld,ny); psi/phi/ustar (nx,ny) */
int les_obukhov(les_ctx *c, const double *u, const double *v, const double *theta, const double *qmix,
                double *sgs_t3, double *sgs_q3, double *psi_m, double *psi_m0, double *psi_h, double *psi_h0,
                double *phi_m, double *phi_h, double *ustar_avg);
/* calcbeta, scalars_module.f90:175-228 */
int les_calcbeta(les_ctx *c, const double *theta, const double *qmix, double *beta_scal);
/* scalar_RHS_calc, scalars_module.f90:232-454 (which=0) / scalars_module_q.f90:114-330
   (which=1); dsdz level 1 is overwritten with the surface gradient, sgs_vert1 is
   the level-1 plane of sgs_t3/sgs_q3; RHS level nz is set to 0 */
int les_scalar_rhs_calc(les_ctx *c, int which, const double *dsdx, const double *dsdy, double *dsdz,
                        const double *u, const double *v, const double *w, const double *Nu_t, const double *Pr_,
                        const double *ustar_avg, const double *phi_h, double *RHS, double *sgs_vert1);
/* step_scalar, scalars_module.f90:457-504 / step_scalar_q, scalars_module_q.f90:357-397 */
int les_step_scalar(les_ctx *c, int which, double *scalar, const double *RHS, const double *RHS_f);
/* interpolag_scalar_Sdep.f90:1-316 */
int les_interpolag_scalar_sdep(les_ctx *c, const les_scalar_sgs_fields *s);
/* lagrange_scalar_Sdep.f90:1-467 (theta is Nyquist-filtered in place, :95);
   reads TnLMMM, TnQNNN, Cs_opt2 of m and overwrites m->Beta */
int les_lagrange_scalar_sdep(les_ctx *c, double *const S[6], const double *u, const double *v, const double *w,
                             double *theta, const les_sgs_fields *m, const les_scalar_sgs_fields *s);
/* scalar_scaledep_dynamic.f90:1-307 (model 6 with scalars, lbc = 1): theta
   is Nyquist-filtered in place (:60); Cs_opt2 is the (ld,ny,nz) field; writes
   s->Pr_t on levels 1..nz-1 and the routine's Cs2_Pr(1..nz-1) into the first
   nz-1 doubles of s->Cs2_Pr (sgs_stag.f90:149 binds its (ld,ny,nz) array);
   reads s->phi_h, s->psi_m ---- */
int les_scalar_scaledep_dynamic(les_ctx *c, double *const S[6], const double *u, const double *v, const double *w,
                                double *theta, const double *Cs_opt2, const les_scalar_sgs_fields *s);
/* sgs_stag.f90:3-500 with model=7 (:59-69,150-153,166-173), or with model 6
   (:143-149, scalar_scaledep_dynamic) when SGS model 6 is selected */
int les_sgs_stag_scalar(les_ctx *c, double *const grad[9], const double *u, const double *v, const double *w,
                        double *theta, double *const tau[6], const les_sgs_fields *m,
                        const les_scalar_sgs_fields *s);

/* ---- divstress_uv.f90:1-63, divstress_w.f90:4-86 ---- */
int les_divstress_uv(les_ctx *c, double *divt, const double *tx, const double *ty, const double *tz);
int les_divstress_w(les_ctx *c, double *divt, const double *tx, const double *ty, const double *tz);

/* ---- press_stag_array.f90:1-294 (p is (ld,ny,0:nz)); tridag_array.f90:1-109
   (a,b,c real (lh,ny,nz+1), r,u complex (lh,ny,nz+1)) ---- */
int les_press_stag_array(les_ctx *c, const double *u, const double *v, const double *w,
                         const double *divtz, double *p, double *dpdx, double *dpdy);
int les_tridag_array(les_ctx *c, const double *a, const double *b, const double *cc, const double *r,
                     double *u);

/* Allocation-free variant of les_tridag_array (no counterpart in the
   reference code), for callers that cannot allocate device memory: code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
   Same arguments, results and errors as les_tridag_array; the Thomas factors
   ((nz+1)*lh*ny doubles) and the error flag live in one file-scope device
   array of 64 MiB per GPU, held until the process ends (grids up to e.g.
   256 x 256 x 128; larger ones return 2 with a message).  Every call on a GPU
   shares that array: calls that overlap on one GPU race. */
int les_tridag_array_noalloc(les_ctx *c, const double *a, const double *b, const double *cc, const double *r,
                             double *u);

/* ---- forcing.f90:128-219 project; rmsdiv.f90:1-113; scalars_module2.f90:22-70 ---- */
int les_project(les_ctx *c, double *u, double *v, double *w, const double *dpdx, const double *dpdy,
                const double *dpdz);
int les_rmsdiv(les_ctx *c, const double *dudx, const double *dvdy, const double *dwdz, double *rms);
int les_timestep_conditions(les_ctx *c, const double *u, const double *v, const double *w,
                            const double *Nu_t, double out[4]);

/* FP32 variants (FP32-bulk mode) */
int les_filt_da_f32(les_ctx *c, float *f, float *dfdx, float *dfdy, int nlev);
int les_ddx_f32(les_ctx *c, float *dfdx, const float *f, int nlev);
int les_ddy_f32(les_ctx *c, float *dfdy, const float *f, int nlev);
int les_test_filter_f32(les_ctx *c, float *plane, const float *G);
int les_dealias1_f32(les_ctx *c, const float *u, float *u_big, int nlev);
int les_dealias2_f32(les_ctx *c, float *u, float *u_big, int nlev);
int les_convec_f32(les_ctx *c, const float *u, const float *v, const float *w, const float *dudy,
                   const float *dudz, const float *dvdx, const float *dvdz, const float *dwdx,
                   const float *dwdy, float *cx, float *cy, float *cz);
int les_divstress_uv_f32(les_ctx *c, float *divt, const float *tx, const float *ty, const float *tz);
int les_divstress_w_f32(les_ctx *c, float *divt, const float *tx, const float *ty, const float *tz);
int les_press_stag_array_f32(les_ctx *c, const float *u, const float *v, const float *w,
                             const float *divtz, float *p, float *dpdx, float *dpdy);

/* ---- whole time step (main.f90:202-415) on context-owned device state:
   neutral model=5, or with les_set_scalars(enable) the convective step with
   obukhov, theta/q transport, buoyancy and (model=7) the scalar Lagrangian
   model.  precision: 0 = FP64, 1 = FP32-bulk (FP64 kept for the tridiagonal
   solve, wall stress, surface layer and reductions) ---- */
int les_state_alloc(les_ctx *c, int precision);
/* device pointer of a state field by reference name (u, v, w, dudx ..., RHSx,
   RHSx_f, txx ..., divtx ..., p, dpdx ..., F_LM ..., Cs_opt2, Nu_t, Beta,
   u_lag ..., psi_m, ustar_avg, ...); nlev and elsize (may be NULL) receive
   its level count and element size.  NULL if unknown. */
void *les_state_field(les_ctx *c, const char *name, int *nlev, int *elsize);
/* seeded initial condition: u = (ustar/kappa_ic) ln(z/z0) at z=(k-1/2)dz,
   v = w = 0, plus amp*ustar*(2U-1) on u, v, w where z < 0.5, U = splitmix64
   uniform of (seed, ((field*nz + k)*ny + y)*nx + x); w(1)=w(nz)=0,
   u(nz)=u(nz-1), v(nz)=v(nz-1), RHS=RHS_f=0.  Resets jt, jt_total and the
   SGS state. */
int les_state_init_ic(les_ctx *c, uint64_t seed, double amp, double ustar, double kappa_ic);
/* seeded scalar initial condition (after les_state_init_ic): theta = T_init
   (K) below z_inv (units of z_i), + inv_strength (K/m) above, plus amp_K*(2U-1)
   below z_inv (U of field index 3); q = q_init (g/kg); Pr_t = Pr, phi_h = 1 */
int les_state_init_scalar_ic(les_ctx *c, uint64_t seed, double T_init, double z_inv, double inv_strength,
                             double amp_K, double q_init);
int les_step(les_ctx *c, int nsteps);
/* Replay the step through HIP graphs (no reference counterpart): windows of
   lcm(2, cs_count) steps whose host-side branches repeat (no Euler start, no
   lagrange_Sdep initialisation, not crossing jt_total = 500) are captured
   once and relaunched; other steps run eagerly.  Results are identical. */
int les_set_graph(les_ctx *c, int enable);
/* capture the graph for the window starting at the current step without
   running it (keeps graph capture out of a timed region) */
int les_graph_prepare(les_ctx *c);
/* rmsdiv of the state's current derivatives and timestep_conditions */
int les_state_diag(les_ctx *c, double *rms, double out4[4]);

/* ---- statistics, FP64 reductions on the device (plane sums compensated,
   so independent of the summation order).  stats layout, nl = nz-1 levels:
   ubar, vbar, wbar, <u'^2>, <v'^2>, <w'^2>, <u'w'>, <txz>, <u'w'>+<txz> (nl
   doubles each; fluctuations on w-nodes, zero on the first level), then
   E11(k1)/z (nl rows of nx/2): 9 nl + nl nx/2 doubles.  surface: ustar_,
   ustar_avg_time, ustar_time, theta_, thetav_, q_time, wt_, L_ (the
   scalar entries and L_ are 0 without scalars). ---- */
/* spectrum, io.f90:1613-1646: spec (nlev, nx/2) device doubles */
int les_spectrum(les_ctx *c, const double *f, int nlev, double *spec);
/* the current statistics of avg_stats, io.f90:1722-1755 (stats: device) */
int les_avg_stats(les_ctx *c, const double *u, const double *v, const double *w, const double *txz,
                  double *stats);
/* obukhov's plane averages, scalars_module.f90:545-561,655-661: first-level
   u, v, txz, tyz (ld,ny), psi_m, psi_m0, ustar_avg (nx,ny); theta, qmix
   (ld,ny) and sgs_t3, sgs_q3 (ld,ny) or all four NULL; out8: device */
int les_surface_stats(les_ctx *c, const double *u, const double *v, const double *txz, const double *tyz,
                      const double *psi_m, const double *psi_m0, const double *ustar_avg, const double *theta,
                      const double *qmix, const double *sgs_t3, const double *sgs_q3, double *out8);
/* the same on the whole-step state (either precision) into host arrays
   (either may be NULL); _accumulate folds them into running means held on
   the device (avg_stats' cumulative averages, io.f90:1800-1813) and
   _average reads those back (n: samples; reset != 0 restarts them) */
int les_state_stats(les_ctx *c, double *stats, double *surface);
int les_state_stats_accumulate(les_ctx *c);
int les_state_stats_average(les_ctx *c, double *stats, double *surface, int *n, int reset);

/* ---- run files (host side; the state entry points copy through host
   buffers).  Fortran sequential unformatted records: 4-byte little-endian
   byte counts around each record; records over 2^31-9 bytes are split into
   subrecords (leading count negative while more follow, trailing count
   negative after the first), as the gfortran / flang runtimes do. ---- */
/* one record of the concatenated segments; append != 0 adds it at the end */
int les_record_write(const char *path, int append, int nseg, const void *const *seg, const uint64_t *bytes);
/* record `index` (0 = first): its length, and up to buf_bytes of it into buf
   (buf may be NULL) */
int les_record_read(const char *path, int index, void *buf, uint64_t buf_bytes, uint64_t *record_bytes);
/* checkpoint, io.f90:1314-1338 (output_final's vel_sc.out / vel.out): the
   state as one record, real(8) throughout (FP32-bulk fields widened, theta
   and q in the reference convention); the layout with scalars when the state
   carries them */
int les_state_checkpoint(les_ctx *c, const char *path);
/* initial.f90:46-175 (initu): the first record of path read with the list
   for model (1..7), initsc and inilag, into the state fields it names (the
   model-7 list with initsc and without inilag does not match checkpoint's
   layout: the reference reads such a file with its fields shifted, and so
   does this reader).  Returns 2 for an unknown model, 3 for a record shorter
   than the list. */
int les_state_read_initial(les_ctx *c, const char *path, int model, int initsc, int inilag);
/* total_time.dat: param.f90:200-210 reads jt_total (returns 1 if the file is
   missing: the reference then takes 0); io.f90:1375-1379 writes it */
int les_read_total_time(const char *path, int *jt_total);
int les_write_total_time(const char *path, int jt_total);
/* patches, bottombc.f90:23-57: wts.dat, wqs.dat, zom.dat in dir (ny lines of
   nx free-format values, W/m^2 and m) -> wt_s (K m/s), wq_s ((g/kg) m/s), zo
   (units of z_i), each (nx,ny) x fastest */
int les_read_surface(const char *dir, int nx, int ny, double z_i, double *wt_s, double *wq_s, double *zo);
/* param_nml of input.nml (param.f90:174-230): defaults of param.f90:63-85,
   the namelist's values, jt_total from total_time.dat (NULL or missing: 0),
   and for jt_total > 0 initu = initsc = 1, inilag = 0; dt = dt_dim u_star /
   z_i.  Returns 3 at an unknown or malformed item (the values before it are
   kept, as a Fortran namelist READ with iostat leaves them). */
typedef struct les_param_nml {
  double q_scale, T_scale, u_star, ug0, vg0, inv_strength, inv_strength_q, dTdz_top, dqdz_top, T_init, q_init,
      dt_dim;
  double z_i, dt;
  int coriolis_forcing, nsteps, c_count, p_count, ubc, damping_method, initu, initsc, inilag, jt_total;
} les_param_nml;
int les_read_param_nml(const char *nml_path, const char *total_time_path, les_param_nml *p);
/* one numeric or logical item of any namelist group (case-insensitive) */
int les_read_nml_value(const char *path, const char *group, const char *key, double *value);

/* Device memory and host mirrors (no reference counterpart): the Fortran
   drop-in shims keep one device mirror per host array, keyed by its address,
   and push inputs / pull outputs around each routine-level call. */
int les_dev_alloc(les_ctx *c, uint64_t bytes, void **ptr);
int les_dev_free(les_ctx *c, void *ptr);
int les_memcpy_h2d(les_ctx *c, void *dst, const void *src, uint64_t bytes);
int les_memcpy_d2h(les_ctx *c, void *dst, const void *src, uint64_t bytes);
/* device mirror of `host` (allocated on first use, grown if bytes grows) */
void *les_mirror(les_ctx *c, const void *host, uint64_t bytes);
int les_mirror_push(les_ctx *c, const void *host, uint64_t bytes);
int les_mirror_pull(les_ctx *c, void *host, uint64_t bytes);

/* Per-launch HIP-event timing on the context stream (no reference
   counterpart; measurement support).  les_profile_report writes a JSON object
   {"routine/kernel": [launches, total_ms, algorithmic_bytes], ...} and clears
   the records. */
int les_profile(les_ctx *c, int enable);
int les_profile_report(les_ctx *c, char *buf, int buflen);

/* ---- autotuner (no counterpart in the reference code): times the column-FFT
   variant (what & 1; kx columns per block -- FP32 8/16/32, FP64 4/8/16 -- and
   the FFT plan, radix-4 or radix-8 passes; in FP32 the zero-pad and
   truncation passes always take radix-8), the tridiagonal schedule
   (what & 2; where the Thomas sweep keeps gamma and U -- 0 HBM, 1 gamma in
   LDS, 2 both in LDS -- and modes per block, 3: parallel cyclic reduction,
   one wavefront per mode, or 4: the sweep from both ends, four threads per
   mode, at most 256 modes per block) and the padded-product pass of the
   convective step (what & 4; 0: groups of levels per workgroup, 1:
   streaming runs of levels with the radix-8 plan, 2: streaming runs with one
   transform per wavefront, for the convective step with both scalars at
   nx2 = 384, elsewhere the same as 1) on this context's state, and keeps
   the fastest.  Widths and Thomas schedules give identical results; the
   plans, the streaming products, cyclic reduction and the two-ended sweep
   change only the rounding.  report: JSON with the measured ms per
   candidate (may be NULL).  les_get_tune / les_set_tune: {col_tx FP32,
   col_tx FP64, tri_mode, tri_block, col_plan FP32, col_plan FP64,
   pad_mode FP32, pad_mode FP64}. ---- */
int les_autotune(les_ctx *c, int what, char *report, int report_len);
int les_get_tune(les_ctx *c, int *out8);
int les_set_tune(les_ctx *c, const int *in8);
/* Layout of the x-spectral intermediates that the row passes write for the
   column passes of filt_da / ddx / ddy (also read by the zero-pad pass of
   the convective inputs formed from filt_da's spectra) and of divstress; no
   counterpart in the reference code, whose FFTW plans write row-major
   spectra.  0 row-major, 1 tiles of 8 coefficient columns, each contiguous
   over a plane's ny rows (needs nx a multiple of 16; the default there).
   Results are bit-identical in either layout; the tiles give the column
   passes whole 128-byte lines.  les_get_layout returns the current one (-1
   for a null context). */
int les_set_layout(les_ctx *c, int xs_tiled);
int les_get_layout(les_ctx *c);

/* ---- ensembles (no counterpart in the reference code): independent
   contexts, each with its own grid, surface parameters, precision and state
   (les_state_alloc + initial condition done), advanced together.  A window of
   W steps (W = lcm of the members' windows) of every member is one HIP graph
   whose branches run on the members' streams; windows in which a member has
   a one-off condition run eagerly.  Each member's results are bit-identical
   to stepping it alone.  Members must outlive the ensemble. ---- */
typedef struct les_ensemble les_ensemble;
int les_ensemble_create(les_ctx *const *members, int n, les_ensemble **out);
void les_ensemble_destroy(les_ensemble *e);
const char *les_ensemble_last_error(les_ensemble *e);
int les_ensemble_window(les_ensemble *e);
int les_ensemble_step(les_ensemble *e, int nsteps);
int les_ensemble_sync(les_ensemble *e);

/* ---- learned SGS closure (no counterpart in the reference code).  An MLP
   8 -> 64 -> 64 -> 64 -> 2 (SiLU) maps local invariants of the strain rate to
   the (Cs_opt2, Beta) that lagrange_Sdep.f90:390-421 produces, with BF16
   MFMA / FP32-accumulate kernels for inference and training.  Inputs per
   point (levels 1..nz-1 of lagrange_Sdep's w-nodes): asinh of |S|, |S_bar|,
   |S_hat|, S:S_bar, S_bar:(|S|S)_bar, S_hat:(|S|S)_hat and the rotation
   magnitude |Omega|, then z/Delta.  Arrays are device pointers; X (P,8) raw
   features, Y/T (P,2) raw (Cs_opt2, Beta); the model normalises internally
   with les_mlp_set_norm / les_mlp_fit_norm.  theta: flat parameters W1 (64x8)
   b1 W2 (64x64) b2 W3 b3 W4 (2x64) b4, row-major [out][in], host floats. ---- */
int les_mlp_nparam(void);
int les_mlp_init(les_ctx *c, uint64_t seed);
int les_mlp_set_params(les_ctx *c, const float *theta);
int les_mlp_get_params(les_ctx *c, float *theta);
int les_mlp_set_norm(les_ctx *c, const float *mu, const float *sd, const float *ymu, const float *ysd);
int les_mlp_get_norm(les_ctx *c, float *mu, float *sd, float *ymu, float *ysd);
int les_mlp_fit_norm(les_ctx *c, const float *X, const float *Y, int64_t P);
int les_mlp_forward(les_ctx *c, const float *X, int64_t P, float *Y);
/* gradient (device, les_mlp_nparam() floats) of L = 1/(2P) sum |y_n - t_n|^2
   in normalised units over all P samples, and L (device float) */
int les_mlp_grad(les_ctx *c, const float *X, const float *T, int64_t P, float *grad, float *loss);
/* `steps` Adam steps on batches of `batch` samples drawn from the P;
   loss_hist: device floats (steps) or NULL */
int les_mlp_train(les_ctx *c, const float *X, const float *T, int64_t P, int batch, int steps, float lr,
                  uint64_t seed, float *loss_hist);
/* features and labels of the state's last step, which must have run the
   Lagrangian update (jt a multiple of the update period); X: (nx*ny*(nz-1), 8),
   Y: (nx*ny*(nz-1), 2) or NULL; point order x fastest, then y, then level */
int les_mlp_harvest(les_ctx *c, float *X, float *Y);
/* mode 0: the dynamic procedure every cs_count steps (the reference).  mode 1:
   lagrange_Sdep every `period` steps (a multiple of cs_count, which is then its
   averaging interval), the network at the other multiples of cs_count.  Each
   dynamic update also judges the network: relative L2 error of its Cs_opt2
   against the dynamic one; above `threshold` the network's output is not
   used (the dynamic coefficients are kept) until a later update passes. */
int les_set_closure(les_ctx *c, int mode, int period, double threshold);
/* out[4]: sum (net - dyn)^2, sum dyn^2, relative error, 1 if the network is in use */
int les_closure_judge(les_ctx *c, double *out);

#ifdef __cplusplus
}
#endif
#endif
