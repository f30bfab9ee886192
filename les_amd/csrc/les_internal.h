// les_internal.h -- device context, constant tables and launcher declarations
// shared by the HIP kernels (les_kernels.hip) and the C-ABI (les_capi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "les_fft.h"

namespace les {

// Level k and column x of the linear index i of an (ld, ny, nlev) array, in
// 32-bit arithmetic (les_create keeps every array below 2^31 points; 64-bit
// division is a long software sequence on the GPU).
struct LevelCol {
  int k, x;
};
__device__ __forceinline__ LevelCol level_col(long i, long pl, int ld) {
  const unsigned ii = (unsigned)i, p = (unsigned)pl;
  const unsigned k = ii / p;
  return {(int)k, (int)((ii - k * p) % (unsigned)ld)};
}

// Compensated summation: s + c is the running sum, every addition an
// error-free TwoSum whose rounding error is carried in c.  A sum formed this
// way is the exact sum rounded once, up to a term of order n eps^2, in any
// order of the additions (plane means of the statistics and of calcbeta).
// No contraction: a product fused into the TwoSum's first addition would
// break the error term.
struct CSum {
  double s, c;
};
__device__ __forceinline__ CSum csum_add(CSum a, double x) {
#pragma clang fp contract(off)
  const double t = a.s + x, bb = t - a.s;
  return CSum{t, a.c + ((a.s - (t - bb)) + (x - bb))};
}
__device__ __forceinline__ CSum csum_merge(CSum a, CSum b) {
#pragma clang fp contract(off)
  const double t = a.s + b.s, bb = t - a.s;
  return CSum{t, (a.c + b.c) + ((a.s - (t - bb)) + (b.s - bb))};
}

// Grid and run constants (param.f90:56-65,98,100,144-151 of the reference),
// plus the derived sizes used by the kernels.
struct Grid {
  int nx, ny, nz;       // nz = w-levels incl. lid (arrays (ld,ny,1:nz))
  int lh, ld, nxh;      // lh = nx/2+1, ld = 2*lh, nxh = nx/2 (kept kx count)
  int nx2, ny2;         // 3/2-padded sizes
  double Lx, Ly, Lz, dx, dy, dz, dt, z0, force, delta;
  int cs_count, inilag, initu;
};

// Per-precision device tables.
template <typename T>
struct Tables {
  cx<T>* tw_nx = nullptr;   // exp(-2 pi i k/n) tables
  cx<T>* tw_ny = nullptr;
  cx<T>* tw_nx2 = nullptr;
  cx<T>* tw_ny2 = nullptr;
  T* kx = nullptr;          // (nxh)   kx(jx) for jx < nx/2  (fft.f90:60-76)
  T* ky = nullptr;          // (ny)    ky(jy), Nyquist row 0
  T* G2 = nullptr;          // (lh*ny) G_test      (2 Delta)
  T* G4 = nullptr;          // (lh*ny) G_test_test (4 Delta)
};

// Device scratch: named slots that grow on demand and live with the context.
struct Arena {
  std::vector<void*> ptr;
  std::vector<size_t> cap;
};

// Launch shapes chosen by les_autotune; the FFT plan and the PCR tridiagonal
// schedule (tri_mode 3) change the rounding, the other choices do not.
struct Tune {
  int col_tx[2] = {16, 8};  // column-FFT tile width (kx columns per block): FP32, FP64
  int col_plan[2] = {0, 0}; // column-FFT plan (les_fft.h): 0 radix-4 passes, 2 radix-8 passes
  int tri_mode = 0;         // Thomas sweep storage: 0 gamma/U in HBM, 1 gamma in LDS, 2 both in LDS;
                            // 3 parallel cyclic reduction, one wavefront per mode
  int tri_block = 128;      // horizontal modes per block of the tridiagonal solve
  int pad_mode[2] = {0, 0}; // padded products: 0 level groups per workgroup, 1 streaming level runs,
                            // 2 streaming with one transform per wavefront (convective step with
                            // both scalars at nx2 = 384; elsewhere as 1)
  int xs_tiled = 0;         // x-spectral intermediates between the row and column passes: 0 row-major,
                            // 1 8-coefficient column tiles (xs_off; set at creation where nx/2 % 8 == 0)
};

// Scalar transport switches and constants (param.f90:53,72-90,139-146,174),
// the lbc=1 surface of bottombc.f90:26-63.
struct ScalarParams {
  int enable = 0;   // S_FLAG: theta and q transported
  int model = 5;    // 7: Pr_ = Pr_t from lagrange_scalar_Sdep, 6: from scalar_scaledep_dynamic; else Pr
  int initsc = 1;   // scalar history given: no Euler scalar step at jt == 1
  int passive = 0;  // passive_scalar: no buoyancy in RHSz
  double Pr = 0.55, z_i = 1000.0, u_star = 1.0, T_scale = 400.0, q_scale = 100.0, g = 9.81;
  double dTdz_top = 0.0, dqdz_top = 0.0;  // K/m, (g/kg)/m at the lid
  // FP32-bulk state stores theta - theta_ref and q - q_ref: a float around
  // 0.7 keeps only ~4 digits of a 1e-4 perturbation.  Every equation but
  // calcbeta and obukhov sees gradients only; those two add the reference
  // back in FP64.  0 in FP64 mode and for the routine-level entry points.
  double theta_ref = 0.0, q_ref = 0.0;
};

// Coriolis forcing and the top sponge (param.f90:68-75,220, main.f90:131-138,
// 319-348, topbc.f90:14-95), set by les_set_forcing
struct Forcing {
  int coriolis = 0, ubc = 0, damping_method = 1;
  double ug0 = 0.0, vg0 = 0.0, u_star = 1.0, z_i = 1000.0;
  double coriol = 0.0, cug = 0.0, cvg = 0.0;  // coriol, coriol*ug, coriol*vg
  std::vector<double> sponge;                 // sponge(1..nz) at [1..nz]
  double* dcoef = nullptr;  // device, 2 (nz-1): 0.5 (sponge(jz) + sponge(jz+1)) | the RHSz factor
};

// What the AB2 epilogues add after the buoyancy term, per step: the
// Coriolis products of the velocities before the update (formed by
// op_forcing_pre, because the epilogue updates u and v in place) and the
// sponge terms about the plane means of the same fields.
template <typename T>
struct ForcingArgs {
  const T* cor[2] = {nullptr, nullptr};  // coriol*v (for RHSx), coriol*u (for RHSy); levels < nz-1
  T cvg = T(0), cug = T(0);
  int damp = 0;                  // damping_method with ubc = 1, else 0
  const double* coef = nullptr;  // Forcing::dcoef
  const double* sums = nullptr;  // plane sums of u, v, w, theta, q on levels < nz-1 ([5][nz-1])
  double nxy = 1.0;
};

struct State;  // whole-step device state (les_capi.hip)

struct ProfRec {
  std::string tag;
  hipEvent_t a, b;
  double bytes;
};

}  // namespace les

struct les_ctx {
  les::Grid g;
  hipStream_t stream = nullptr;
  int device = 0;
  int jt = 0, jt_total = 0;        // param.f90:147-148
  int flm_init = 0, fqn_init = 0;  // lagrange_Sdep.f90:67-68 SAVE flags
  int keep2 = 0, keep4 = 0;        // kx columns with non-zero G_test / G_test_test
  les::FftPlan p_nx, p_ny, p_nx2, p_ny2;
  les::Tables<double> td;
  les::Tables<float> tf;
  double* G2_host = nullptr;       // host copies used for pruning decisions
  les::Arena arena;
  les::State* st = nullptr;
  std::string err;
  // optional per-launch HIP-event timing (les_profile)
  bool prof = false;
  std::vector<les::ProfRec> prof_recs;
  std::string op_tag;
  // captured multi-step window of the whole step (les_set_graph)
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  int g_start_mod = -1, g_use_const = -1, g_swap = -1;
  long eager_steps = 0;  // steps run eagerly since the state was initialised
  // scalars (les_set_scalars)
  les::ScalarParams scal;
  double* wt_s = nullptr;          // (nx,ny) kinematic surface fluxes, device
  double* wq_s = nullptr;
  double zo_avg = 0.0;             // exp(mean log zo), lagrange_scalar_Sdep.f90:110
  double wt_mean = 0.0;            // sum(wt_s)/(nx*ny), lagrange_scalar_Sdep.f90:114
  int glm_init = 0, gqn_init = 0;  // lagrange_scalar_Sdep.f90:65-66 SAVE flags
  int sgs_model = 5;               // param.f90 'model': 1 Smagorinsky-Mason, 2 standard dynamic,
                                   // 3 plane-averaged scale-dependent, 4 Lagrangian scale-invariant,
                                   // 5 Lagrangian scale-dependent, 6 plane-averaged scale-dependent
                                   // without the quintic (new_scaledep_dynamic)
  double* l2z = nullptr;           // model 1: l(jz)^2 of sgs_stag.f90:84-94, device (nz)
  les::Forcing frc;
  les::Tune tune;
  int lag_cs = 0;                  // steps between Lagrangian updates when it differs from cs_count (closure)
  int last_lag_jt = -1;            // jt of the last lagrange_Sdep update
  // learned closure (les_mlp.hip): the network between dynamic updates
  void* mlp = nullptr;             // les::MlpModel
  int closure_mode = 0, closure_period = 10;
  double closure_thresh = 0.2;
  long stats_n = 0;                // samples in the running statistics (les_state_stats_accumulate)
  // host-address -> device mirror registry (les_mirror)
  std::vector<std::pair<const void*, std::pair<void*, size_t>>> mirrors;
};

namespace les {

// scratch slots
enum Slot : int {
  S_XS0 = 0, S_XS1, S_XS2, S_XS3, S_BIG0, S_BIG1, S_BIG2, S_SIJ, S_FF, S_LAGF, S_LAGX,
  S_PSPEC, S_PHAT, S_TRI, S_TMP0, S_TMP1, S_TMP2, S_RED, S_PLANE,
  S_SC0, S_SC1, S_SC2, S_SC3, S_LAGF2, S_RED2, S_SDQ, S_SDLEV, S_MLPP, S_TUNE0, S_TUNE1, S_TUNE2,
  S_STATP, S_STATX, S_STAT, S_STATAVG, S_FORCE, S_FSUM, S_FSPEC0, S_FSPEC1, S_FSPEC2, S_FWALL, S_NSLOTS
};
void* scratch(les_ctx* c, int slot, size_t bytes);
void scratch_release(les_ctx* c, int slot);  // give a slot's memory back (stream drained first)

template <typename T> const Tables<T>& tables(const les_ctx* c);
template <> const Tables<double>& tables<double>(const les_ctx* c);
template <> const Tables<float>& tables<float>(const les_ctx* c);

// ---- launchers (les_kernels.hip); all enqueue on c->stream ----
// horizontal spectral operators, fields (ld,ny,nlev)
// with keep_slot >= 0 the x-spectral rows of (f, df/dx, df/dy) stay in that
// scratch slot (kind-major, field, level, y; scale 1/nx) and *spec points there
template <typename T>
void op_filt_da(les_ctx* c, T* f, T* dfdx, T* dfdy, int nlev, int keep_slot = -1, const cx<T>** spec = nullptr);
template <typename T>
void op_filt_da3(les_ctx* c, T* u, T* v, T* w, T* const gr[9], int nlev, int keep_slot = -1,
                 const cx<T>** spec = nullptr);
// filt_da's kept spectra of u, v, w and of theta, q for the convective inputs
template <typename T>
struct SpecIn {
  const cx<T>* uvw = nullptr;
  const cx<T>* s[2] = {nullptr, nullptr};
};
template <typename T> void op_ddxy(les_ctx* c, T* dfdx, T* dfdy, const T* f, int nlev);  // null outputs skipped
template <typename T> void op_test_filter(les_ctx* c, T* planes, int nplanes, const T* G, int keep);
template <typename T> void op_test_filter_level0(les_ctx* c, const T* a, const T* b, T* dst, const T* G, int keep);
template <typename T> void op_dealias1(les_ctx* c, const T* u, T* u_big, int nlev);
template <typename T> void op_dealias2(les_ctx* c, T* u, T* u_big, int nlev);
template <typename T> void op_padd(les_ctx* c, T* big_c, const T* small_c, int nplanes);
template <typename T> void op_unpadd(les_ctx* c, T* small_c, const T* big_c, int nplanes);
template <typename T> void op_ddz_uv(les_ctx* c, T* dfdz, const T* f);
template <typename T> void op_ddz_w(les_ctx* c, T* dfdz, const T* f);
template <typename T> void op_ddz3(les_ctx* c, T* dudz, T* dvdz, T* dwdz, const T* u, const T* v, const T* w);
template <typename T> void op_divstress(les_ctx* c, T* divt, const T* tx, const T* ty, const T* tz, int wnode);
// divtx, divty, divtz of the step in one pipeline; tau: txx, txy, txz, tyy, tyz, tzz
template <typename T> void op_divstress3(les_ctx* c, T* const divt[3], T* const tau[6]);
template <typename T>
void op_convec(les_ctx* c, const T* u, const T* v, const T* w, const T* dudy, const T* dudz,
               const T* dvdx, const T* dvdz, const T* dwdx, const T* dwdy, T* cx, T* cy, T* cz);
template <typename T>
void op_convec_ab2(les_ctx* c, T* u, T* v, T* w, const T* dudy, const T* dudz, const T* dvdx, const T* dvdz,
                   const T* dwdx, const T* dwdy, T* RHSx, T* RHSy, T* RHSz, T* RHSx_f, T* RHSy_f, T* RHSz_f,
                   const T* divtx, const T* divty, const T* divtz, int euler_first, const T* beta = nullptr,
                   const ForcingArgs<T>* fo = nullptr, const SpecIn<T>* si = nullptr);
// the per-step inputs of the forcing terms (les_scalars.hip): Coriolis
// products and plane sums of the fields before the AB2 update; theta/q may be null
template <typename T>
ForcingArgs<T> op_forcing_pre(les_ctx* c, const T* u, const T* v, const T* w, const T* theta, const T* q);
template <typename T>
void op_wallstress(les_ctx* c, const T* u, const T* v, T* dudz, T* dvdz, T* txz, T* tyz,
                   const double* psi_m, const double* psi_m0, const double* phi_m, double* ustar_avg);

struct SgsPtrs {  // device pointers, (ld,ny,nz) unless noted
  void *F_LM, *F_MM, *F_QN, *F_NN, *Beta, *Cs_opt2, *Nu_t, *u_lag, *v_lag, *w_lag;
  void *TnLMMM, *TnQNNN, *L11t, *L22t, *L33t, *Q11t, *Q22t, *Q33t;
  void *Beta_avg, *Betaclip_avg;  // (nz), double in every precision
};
// scalar Lagrangian SGS state (sgsmodule.f90 G_*, u_lag_s, Pr_t) plus the
// surface functions the scalar model reads (bottombc phi_h, psi_m, (nx,ny) double)
struct ScalarSgsPtrs {
  void *G_LM, *G_MM, *G_QN, *G_NN, *u_lag_s, *v_lag_s, *w_lag_s, *Pr_t, *Cs2_Pr;
  const double *phi_h, *psi_m;
};
template <typename T> void op_calc_sij(les_ctx* c, const T* const g[9], T* const S[6]);
template <typename T> void op_interpolag(les_ctx* c, const SgsPtrs& s);
template <typename T> void op_interpolag4(les_ctx* c, T* const F[4], T* ul, T* vl, T* wl);
// the interpolation of F into scratch (returned, 4 fields) for an epilogue
// that reads it and writes F
template <typename T> T* op_interpolag4_into(les_ctx* c, const T* const F[4], T* ul, T* vl, T* wl);
// interpolag_Ssim.f90 on F_LM, F_MM (model 4)
template <typename T> void op_interpolag_ssim(les_ctx* c, const SgsPtrs& s);
// lagrange_Ssim.f90:4-300 (model 4)
template <typename T>
void op_lagrange_ssim(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, const SgsPtrs& s);
// fields f0..20 of the lagrange_Sdep filter set (f0 = 9: strain rates only)
template <typename T>
T* lag_filter_mom(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, int f0 = 0);
// steps between Lagrangian updates: the averaging count of u_lag and the
// Lagrangian time step (interpolag_Sdep.f90:228-262, lagrange_Sdep.f90:321-331)
inline int lag_count(const les_ctx* c) { return c->lag_cs > 0 ? c->lag_cs : c->g.cs_count; }
template <typename T>
void op_lagrange(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, const SgsPtrs& s);
// theta/ss may be null (model 5); with ss the model-7 scalar model runs after lagrange_Sdep
template <typename T>
void op_sgs_stag(les_ctx* c, const T* const g[9], const T* u, const T* v, const T* w, T* const tau[6],
                 const SgsPtrs& s, T* theta = nullptr, const ScalarSgsPtrs* ss = nullptr);

// SGS model 3 (les_dynamic.hip): scaledep_dynamic.f90:4-255 with rtnewt.
// S: the routine's dummies S11..S33; per-level Cs_opt2 and beta go to
// cs_lev/beta_lev (device doubles, nz; null: internal scratch); with
// Cs_planes every point of level k receives Cs_opt2(k) (sgs_stag.f90:133-135)
template <typename T>
void op_scaledep_dynamic(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, double* cs_lev,
                         double* beta_lev, T* Cs_planes);
// SGS model 2 (les_dynamic.hip): std_dynamic.f90:4-140, per-level Cs_opt2 to
// cs_lev (device, nz; null: scratch) and, with Cs_planes, to every point
template <typename T>
void op_std_dynamic(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, double* cs_lev,
                    T* Cs_planes);
// SGS model 6 (les_dynamic.hip): new_scaledep_dynamic.f90, the same outputs
template <typename T>
void op_new_scaledep_dynamic(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, double* cs_lev,
                             T* Cs_planes);
// its scalar part (les_scalars.hip): scalar_scaledep_dynamic.f90 filters theta
// in place, writes Cs2_Pr(1..nz-1) into the first nz-1 elements of ss.Cs2_Pr
// (as the reference's call binds sgs_stag's (ld,ny,nz) array) and
// Pr_t = Cs_opt2 / Cs2_Pr on levels 1..nz-1; reuse_mom: the momentum filter
// set of this update is already in its scratch slot
template <typename T>
void op_scalar_scaledep(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, T* theta,
                        const T* Cs_opt2, const ScalarSgsPtrs& ss, bool reuse_mom);

// ---- learned closure (les_mlp.hip) ----
int mlp_nparam();
void mlp_init(les_ctx* c, uint64_t seed);
void mlp_free(les_ctx* c);
void mlp_set_params(les_ctx* c, const float* host);
void mlp_get_params(les_ctx* c, float* host);
void mlp_set_norm(les_ctx* c, const float* mu, const float* sd, const float* ymu, const float* ysd);
void mlp_get_norm(les_ctx* c, float* mu, float* sd, float* ymu, float* ysd);
void mlp_fit_norm(les_ctx* c, const float* X, const float* Y, long P);
void mlp_judge_read(les_ctx* c, double out[4]);
void op_mlp_forward(les_ctx* c, const float* X, long P, float* Y);
void op_mlp_grad(les_ctx* c, const float* X, const float* T, long P, float* grad, float* loss);
void op_mlp_train(les_ctx* c, const float* X, const float* T, long P, int batch, int steps, float lr,
                  uint64_t seed, float* loss_hist);
template <typename T>
void op_mlp_harvest(les_ctx* c, const T* F, const T* const S[6], const T* const gr[9], const T* Cs, const T* Beta,
                    float* X, float* Y);
template <typename T>
void op_mlp_closure(les_ctx* c, const T* F, const T* const S[6], const T* const gr[9], T* Cs, T* Beta, int judge,
                    double thresh);

// ---- scalars (les_scalars.hip) ----
// obukhov, scalars_module.f90:508-690.  sgs_t3/sgs_q3: level-1 planes (ld,ny)
// double; psi/phi/ustar (nx,ny) double.
// th_ref/q_ref: references added to the stored scalars (see ScalarParams)
template <typename T>
void op_obukhov(les_ctx* c, const T* u, const T* v, const T* theta, double* sgs_t3, double* sgs_q3,
                double* psi_m, double* psi_m0, double* psi_h, double* psi_h0, double* phi_m, double* phi_h,
                double* ustar_avg, double th_ref = 0.0);
// calcbeta, scalars_module.f90:175-228
template <typename T>
void op_calcbeta(les_ctx* c, const T* theta, const T* qmix, T* beta, double th_ref = 0.0, double q_ref = 0.0);
// one or two scalars through scalar_RHS_calc (+ step_scalar when do_step)
struct ScalarJob {
  void *s, *dsdx, *dsdy, *dsdz, *RHS, *RHS_f;  // (ld,ny,nz)
  double* sgs_vert1;                           // level-1 plane (ld,ny)
  int which;                                   // 0 theta, 1 q
};
template <typename T>
void op_scalar_dz(les_ctx* c, const ScalarJob* jobs, int ns, const double* phi_h, const double* ustar_avg,
                  int surface_only);
template <typename T>
void op_scalar_rhs(les_ctx* c, const ScalarJob* jobs, int ns, const T* u, const T* v, const T* w, const T* Nu_t,
                   const T* Pr_t, int do_step, int euler);
template <typename T> void op_step_scalar(les_ctx* c, int which, T* s, const T* RHS, const T* RHS_f);
// convec + RHS assembly + AB2 of the momentum and both scalars' advection,
// SGS flux divergence and step, sharing one padded-grid pass
template <typename T>
void op_pad_products_probe(les_ctx* c, int scal);
template <typename T>
void op_convec_scalar_ab2(les_ctx* c, T* u, T* v, T* w, const T* const gr[9], T* const R[3], T* const Rf[3],
                          const T* const divt[3], int euler_first, const T* beta, const ScalarJob* jobs,
                          const T* Nu_t, const T* Pr_t, int scal_euler, const ForcingArgs<T>* fo = nullptr,
                          const SpecIn<T>* si = nullptr);
// lagrange_scalar_Sdep.f90:1-467 (theta Nyquist-filtered in place); with
// reuse_mom the momentum model's filtered fields of this step are reused
template <typename T>
void op_lagrange_scalar(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, T* theta,
                        const SgsPtrs& m, const ScalarSgsPtrs& ss, bool reuse_mom);

template <typename T>
void op_press_stag(les_ctx* c, const T* u, const T* v, const T* w, const T* divtz, T* p, T* dpdx, T* dpdy);
int op_tridag_array(les_ctx* c, const double* a, const double* b, const double* cc, const double* r, double* u);
// the same on caller-provided device scratch: gam of tridag_array_gamma_bytes,
// err one int; returns 0, or 1 / 2 for tridag_array.f90's two stops
size_t tridag_array_gamma_bytes(const les_ctx* c);
int tridag_array_with(les_ctx* c, const double* a, const double* b, const double* cc, const double* r, double* u,
                      double* gam, int* err);
// les_stats.hip: statistics as FP64 reductions (all outputs device doubles).
// stats layout over nl = nz - 1 levels: ubar vbar wbar <u'^2> <v'^2> <w'^2>
// <u'w'> <txz> <u'w'>+<txz> (nl each), then E11(k1)/z (nl rows of nx/2)
size_t stats_len(const les_ctx* c);
template <typename T>
void op_spectrum(les_ctx* c, const T* f, int nlev, double* spec, int eoz);
template <typename T>
void op_avg_stats(les_ctx* c, const T* u, const T* v, const T* w, const T* txz, double* stats);
// surface: ustar_, ustar_avg_time, ustar_time, theta_, thetav_, q_time, wt_,
// L_; theta..sgs_q3 may be null; th_ref, q_ref are added to stored theta, q
template <typename T>
void op_surface_stats(les_ctx* c, const T* u, const T* v, const T* txz, const T* tyz, const double* psi_m,
                      const double* psi_m0, const double* ustar_avg, const T* theta, const T* qmix,
                      const double* sgs_t3, const double* sgs_q3, double th_ref, double q_ref, double* out8);
// running means (n avg + cur) / (n + 1) over len doubles
void op_running_mean(les_ctx* c, double* avg, const double* cur, size_t len, long n);
// les_noalloc.hip: device address of the allocation-free entry points'
// scratch array on `device` (the current device), looked up once per device
int noalloc_init(int device, std::string* err);
// the pressure solve's tridiagonal sweep under the schedule in c->tune
size_t tridiag_lds(int mode, int block, int nz);
template <typename T>
void launch_tridiag(les_ctx* c, const cx<T>* H, const cx<T>* BT, cx<T>* PH, double* gam, cx<double>* Ub, int* err);
template <typename T>
void op_project(les_ctx* c, T* u, T* v, T* w, const T* dpdx, const T* dpdy, const T* dpdz);
template <typename T> double op_rmsdiv(les_ctx* c, const T* dudx, const T* dvdy, const T* dwdz);
template <typename T>
void op_timestep_conditions(les_ctx* c, const T* u, const T* v, const T* w, const T* nu, double out[4]);

// step pieces
template <typename T>
void op_ab2(les_ctx* c, T* u, T* v, T* w, T* RHSx, T* RHSy, T* RHSz, const T* RHSx_f, const T* RHSy_f,
            const T* RHSz_f, const T* divtx, const T* divty, const T* divtz, int euler_first);
template <typename T>
void op_pressure_update(les_ctx* c, T* u, T* v, T* w, T* RHSx, T* RHSy, T* RHSz, const T* p,
                        const T* dpdx, const T* dpdy, T* dpdz);
template <typename T> void op_fill(les_ctx* c, T* a, size_t n, T val);
template <typename T> void op_copy(les_ctx* c, T* dst, const T* src, size_t n);

void check(les_ctx* c, hipError_t e, const char* what);

// Records HIP events around one launch on the context stream when profiling
// is enabled; `bytes` is the launch's algorithmic HBM traffic (in + out).
struct ProfScope {
  les_ctx* c;
  int idx = -1;
  ProfScope(les_ctx* cc, const char* kind, double bytes) : c(cc) {
    if (!c->prof) return;
    ProfRec r;
    r.tag = c->op_tag + "/" + kind;
    r.bytes = bytes;
    (void)hipEventCreate(&r.a);
    (void)hipEventCreate(&r.b);
    (void)hipEventRecord(r.a, c->stream);
    c->prof_recs.push_back(r);
    idx = (int)c->prof_recs.size() - 1;
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(c->prof_recs[idx].b, c->stream);
  }
};
// Names the routine whose launches follow (profiling tags).
struct OpTag {
  les_ctx* c;
  std::string old;
  OpTag(les_ctx* cc, const char* name) : c(cc), old(cc->op_tag) { c->op_tag = name; }
  ~OpTag() { c->op_tag = old; }
};

}  // namespace les
