// les_capi.hip -- C ABI (include/les_amd.h): context set-up, device scratch,
// routine-level drop-in entry points and the whole-step driver.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/les_amd.h"
#include "les_internal.h"

using namespace les;

namespace les {

constexpr double PI = 3.1415926535897932384626433;  // param.f90:51

void* scratch(les_ctx* c, int slot, size_t bytes) {
  Arena& a = c->arena;
  if ((int)a.ptr.size() < S_NSLOTS) {
    a.ptr.assign(S_NSLOTS, nullptr);
    a.cap.assign(S_NSLOTS, 0);
  }
  if (bytes == 0) bytes = 16;
  if (a.cap[slot] < bytes) {
    // growth happens on first use of a size; in-flight work may still read
    // the old buffer, so drain the stream before releasing it
    check(c, hipStreamSynchronize(c->stream), "scratch sync");
    if (a.ptr[slot]) check(c, hipFree(a.ptr[slot]), "scratch free");
    a.ptr[slot] = nullptr;
    size_t want = bytes + bytes / 8;
    check(c, hipMalloc(&a.ptr[slot], want), "scratch alloc");
    a.cap[slot] = a.ptr[slot] ? want : 0;
  }
  return a.ptr[slot];
}

void scratch_release(les_ctx* c, int slot) {
  Arena& a = c->arena;
  if ((int)a.ptr.size() <= slot || !a.ptr[slot]) return;
  check(c, hipStreamSynchronize(c->stream), "scratch release sync");
  check(c, hipFree(a.ptr[slot]), "scratch release");
  a.ptr[slot] = nullptr;
  a.cap[slot] = 0;
}

static FftPlan make_plan(int n) {
  FftPlan p;
  std::memset(&p, 0, sizeof(p));
  p.n = n;
  int m = n;
  while (m % 4 == 0) { p.rad[p.ns++] = 4; m /= 4; }
  while (m % 2 == 0) { p.rad[p.ns++] = 2; m /= 2; }
  while (m % 3 == 0) { p.rad[p.ns++] = 3; m /= 3; }
  if (m != 1 || n % 4 != 0) p.n = -1;
  return p;
}

// exp(-2 pi i k/n) for k < n (run-time-length kernels), then the
// stage-major table of the compile-time-length kernels (les_fft.h
// fft_radix / fft_tw_off): stage s of radix R after NS = R_0..R_{s-1} holds
// W^(r k n/(NS R)), r = 1..R-1, k < NS.  Computed in long double.
template <typename T>
static cx<T>* upload_twiddles(les_ctx* c, int n) {
  const size_t len = (size_t)(1 + FFT_NPLANS) * n;
  std::vector<cx<T>> h(len);
  auto w = [&](long k) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    return cx<T>{(T)cosl(a), (T)sinl(a)};
  };
  for (int k = 0; k < n; ++k) h[k] = w(k);
  for (int plan = 0; plan < FFT_NPLANS; ++plan) {  // the compile-time plans (les_fft.h)
    const int nst = fft_nstages(n, plan);
    if (nst <= 0) continue;
    int off = n + plan * n, ns = 1;
    for (int s = 0; s < nst; ++s) {
      const int R = fft_radix(n, s, plan), tws = n / (ns * R);
      for (int r = 1; r < R; ++r)
        for (int k = 0; k < ns; ++k) h[off + (r - 1) * ns + k] = w((long)r * k * tws);
      off += (R - 1) * ns;
      ns *= R;
    }
  }
  cx<T>* d = nullptr;
  check(c, hipMalloc(&d, len * sizeof(cx<T>)), "twiddle alloc");
  check(c, hipMemcpy(d, h.data(), len * sizeof(cx<T>), hipMemcpyHostToDevice), "twiddle upload");
  return d;
}
template <typename T>
static T* upload(les_ctx* c, const std::vector<double>& v) {
  std::vector<T> h(v.size());
  for (size_t i = 0; i < v.size(); ++i) h[i] = (T)v[i];
  T* d = nullptr;
  check(c, hipMalloc(&d, h.size() * sizeof(T)), "table alloc");
  check(c, hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "table upload");
  return d;
}

// fft.f90:54-81 init_wavenumber (incl. the ky(lh,:)=0 quirk) and
// test_filtermodule.f90:32-63 test_filter_init (ifilter=1)
struct HostTables {
  std::vector<double> kx, ky, k2, G2, G4;
};
static void host_tables(const Grid& g, HostTables& t) {
  const int lh = g.lh, ny = g.ny;
  t.kx.assign((size_t)lh * ny, 0.0);
  t.ky.assign((size_t)lh * ny, 0.0);
  t.k2.assign((size_t)lh * ny, 0.0);
  for (int jy = 0; jy < ny; ++jy)
    for (int jx = 0; jx < lh; ++jx) {
      double kx = jx < lh - 1 ? (double)jx : 0.0;
      double ky = (double)(((jy + ny / 2) % ny) - ny / 2);
      if (jx == lh - 1) { kx = 0.0; ky = 0.0; }
      if (jy == ny / 2) { kx = 0.0; ky = 0.0; }
      kx = 2.0 * PI / g.Lx * kx;
      ky = 2.0 * PI / g.Ly * ky;
      t.kx[jx + (size_t)lh * jy] = kx;
      t.ky[jx + (size_t)lh * jy] = ky;
      t.k2[jx + (size_t)lh * jy] = kx * kx + ky * ky;
    }
  auto filt = [&](double alpha, std::vector<double>& G) {
    G.assign((size_t)lh * ny, 0.0);
    const double delta = alpha * sqrt(g.dx * g.dy);
    const double q = PI / delta, kc2 = q * q;
    for (size_t i = 0; i < G.size(); ++i) G[i] = t.k2[i] >= kc2 ? 0.0 : 1.0 / ((double)g.nx * g.ny);
    for (int y = 0; y < ny; ++y) G[lh - 1 + (size_t)lh * y] = 0.0;
    for (int x = 0; x < lh; ++x) G[x + (size_t)lh * (ny / 2)] = 0.0;
  };
  filt(2.0, t.G2);
  filt(4.0, t.G4);
}
static int keep_cols(const Grid& g, const std::vector<double>& G) {
  int keep = 1;
  for (int y = 0; y < g.ny; ++y)
    for (int x = 0; x < g.nxh; ++x)
      if (G[x + (size_t)g.lh * y] != 0.0 && x + 1 > keep) keep = x + 1;
  return keep;
}

template <typename T>
static void make_tables(les_ctx* c, Tables<T>& t, const HostTables& h) {
  const Grid& g = c->g;
  t.tw_nx = upload_twiddles<T>(c, g.nx);
  t.tw_ny = upload_twiddles<T>(c, g.ny);
  t.tw_nx2 = upload_twiddles<T>(c, g.nx2);
  t.tw_ny2 = upload_twiddles<T>(c, g.ny2);
  std::vector<double> kx(g.nxh), ky(g.ny);
  for (int x = 0; x < g.nxh; ++x) kx[x] = h.kx[x];                   // row jy=1 (ky=0)
  for (int y = 0; y < g.ny; ++y) ky[y] = h.ky[(size_t)g.lh * y];     // column jx=1
  t.kx = upload<T>(c, kx);
  t.ky = upload<T>(c, ky);
  t.G2 = upload<T>(c, h.G2);
  t.G4 = upload<T>(c, h.G4);
}
template <typename T>
static void free_tables(Tables<T>& t) {
  void* ps[] = {t.tw_nx, t.tw_ny, t.tw_nx2, t.tw_ny2, t.kx, t.ky, t.G2, t.G4};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

// ------------------------------------------------------------------------
// whole-step state
// ------------------------------------------------------------------------
struct FieldInfo {
  void* p;
  int nlev;
  int elsize;
};
struct State {
  int prec = 0;  // 0 fp64, 1 fp32-bulk
  size_t es = 8;
  std::map<std::string, FieldInfo> f;
  int rhs_swap = 0;  // RHS <-> RHS_f buffer parity
  int scal = 0;      // scalar fields allocated and stepped
  std::vector<void*> owned;
};

static const char* F3[] = {"u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz",
                           "RHSx", "RHSy", "RHSz", "RHSx_f", "RHSy_f", "RHSz_f", "txx", "txy", "txz", "tyy", "tyz",
                           "tzz", "divtx", "divty", "divtz", "dpdx", "dpdy", "dpdz", "F_LM", "F_MM", "F_QN", "F_NN",
                           "Beta", "Cs_opt2", "Nu_t", "u_lag", "v_lag", "w_lag", "TnLMMM", "TnQNNN", "L11t", "L22t",
                           "L33t", "Q11t", "Q22t", "Q33t"};

// scalar fields allocated when the scalars are enabled (sim_param theta/qmix,
// scalars_module(_q) gradients/RHS, calcbeta's beta_scal, sgsmodule G_*,
// u_lag_s, Pr_t, and lagrange_scalar_Sdep's Cs2_Pr)
static const char* F3S[] = {"theta", "qmix", "dTdx", "dTdy", "dTdz", "dqdx", "dqdy", "dqdz", "RHS_T", "RHS_Tf",
                            "RHS_q", "RHS_qf", "beta_scal", "G_LM", "G_MM", "G_QN", "G_NN", "u_lag_s", "v_lag_s",
                            "w_lag_s", "Pr_t", "Cs2_Pr"};

// RHS <-> RHS_f pairs that exchange buffers each step instead of copying
static const char* RHS_PAIRS[][2] = {{"RHSx", "RHSx_f"}, {"RHSy", "RHSy_f"}, {"RHSz", "RHSz_f"},
                                     {"RHS_T", "RHS_Tf"}, {"RHS_q", "RHS_qf"}};

static void* state_ptr(les_ctx* c, const char* name) {
  State* s = c->st;
  std::string n(name);
  if (s->rhs_swap) {
    for (auto& pr : RHS_PAIRS) {
      if (n == pr[0]) { n = pr[1]; break; }
      if (n == pr[1]) { n = pr[0]; break; }
    }
  }
  auto it = s->f.find(n);
  return it == s->f.end() ? nullptr : it->second.p;
}

template <typename T>
static T* F(les_ctx* c, const char* name) {
  return (T*)state_ptr(c, name);
}

template <typename T>
static SgsPtrs sgs_ptrs(les_ctx* c) {
  SgsPtrs s;
  s.F_LM = F<T>(c, "F_LM"); s.F_MM = F<T>(c, "F_MM"); s.F_QN = F<T>(c, "F_QN"); s.F_NN = F<T>(c, "F_NN");
  s.Beta = F<T>(c, "Beta"); s.Cs_opt2 = F<T>(c, "Cs_opt2"); s.Nu_t = F<T>(c, "Nu_t");
  s.u_lag = F<T>(c, "u_lag"); s.v_lag = F<T>(c, "v_lag"); s.w_lag = F<T>(c, "w_lag");
  s.TnLMMM = F<T>(c, "TnLMMM"); s.TnQNNN = F<T>(c, "TnQNNN");
  s.L11t = F<T>(c, "L11t"); s.L22t = F<T>(c, "L22t"); s.L33t = F<T>(c, "L33t");
  s.Q11t = F<T>(c, "Q11t"); s.Q22t = F<T>(c, "Q22t"); s.Q33t = F<T>(c, "Q33t");
  s.Beta_avg = state_ptr(c, "Beta_avg");
  s.Betaclip_avg = state_ptr(c, "Betaclip_avg");
  return s;
}

template <typename T>
static ScalarSgsPtrs scalar_sgs_ptrs(les_ctx* c) {
  ScalarSgsPtrs p;
  p.G_LM = F<T>(c, "G_LM"); p.G_MM = F<T>(c, "G_MM"); p.G_QN = F<T>(c, "G_QN"); p.G_NN = F<T>(c, "G_NN");
  p.u_lag_s = F<T>(c, "u_lag_s"); p.v_lag_s = F<T>(c, "v_lag_s"); p.w_lag_s = F<T>(c, "w_lag_s");
  p.Pr_t = F<T>(c, "Pr_t"); p.Cs2_Pr = F<T>(c, "Cs2_Pr");
  p.phi_h = (const double*)state_ptr(c, "phi_h");
  p.psi_m = (const double*)state_ptr(c, "psi_m");
  return p;
}

template <typename T>
static void scalar_jobs(les_ctx* c, ScalarJob jobs[2]) {
  jobs[0] = ScalarJob{F<T>(c, "theta"), F<T>(c, "dTdx"), F<T>(c, "dTdy"), F<T>(c, "dTdz"), F<T>(c, "RHS_T"),
                      F<T>(c, "RHS_Tf"), (double*)state_ptr(c, "sgs_t3"), 0};
  jobs[1] = ScalarJob{F<T>(c, "qmix"), F<T>(c, "dqdx"), F<T>(c, "dqdy"), F<T>(c, "dqdz"), F<T>(c, "RHS_q"),
                      F<T>(c, "RHS_qf"), (double*)state_ptr(c, "sgs_q3"), 1};
}

// theta_all_in_one (scalars_module.f90:57-171) and q_all_in_one
// (scalars_module_q.f90:31-111) up to scalar_RHS_calc: filt_da, ddz_uv and
// the surface gradient of both scalars and calcbeta, which reads theta after
// its filt_da and q before its own, as in the reference.  The advection, SGS
// flux divergence and step_scalar follow inside op_convec_scalar_ab2.
template <typename T>
static void scalars_prep(les_ctx* c, const ScalarJob jobs[2], SpecIn<T>* si) {
  const Grid& g = c->g;
  const ScalarParams& P = c->scal;
  op_filt_da<T>(c, (T*)jobs[0].s, (T*)jobs[0].dsdx, (T*)jobs[0].dsdy, g.nz, S_FSPEC1, &si->s[0]);
  op_calcbeta<T>(c, (T*)jobs[0].s, (T*)jobs[1].s, F<T>(c, "beta_scal"), P.theta_ref, P.q_ref);
  op_filt_da<T>(c, (T*)jobs[1].s, (T*)jobs[1].dsdx, (T*)jobs[1].dsdy, g.nz, S_FSPEC2, &si->s[1]);
  op_scalar_dz<T>(c, jobs, 2, (const double*)state_ptr(c, "phi_h"), (const double*)state_ptr(c, "ustar_avg"), 0);
}

template <typename T>
static void step_once(les_ctx* c) {
  const Grid& g = c->g;
  State* s = c->st;
  const int scal = s->scal;
  c->jt += 1;
  c->jt_total += 1;
  s->rhs_swap ^= 1;  // RHS_f = RHS (main.f90:212-214, scalars_module.f90:121) as a buffer swap
  T *u = F<T>(c, "u"), *v = F<T>(c, "v"), *w = F<T>(c, "w");
  T* gr[9] = {F<T>(c, "dudx"), F<T>(c, "dudy"), F<T>(c, "dudz"), F<T>(c, "dvdx"), F<T>(c, "dvdy"),
              F<T>(c, "dvdz"), F<T>(c, "dwdx"), F<T>(c, "dwdy"), F<T>(c, "dwdz")};
  T* tau[6] = {F<T>(c, "txx"), F<T>(c, "txy"), F<T>(c, "txz"), F<T>(c, "tyy"), F<T>(c, "tyz"), F<T>(c, "tzz")};
  if (scal) {  // main.f90:217
    op_obukhov<T>(c, u, v, F<T>(c, "theta"), (double*)state_ptr(c, "sgs_t3"), (double*)state_ptr(c, "sgs_q3"),
                  (double*)state_ptr(c, "psi_m"), (double*)state_ptr(c, "psi_m0"), (double*)state_ptr(c, "psi_h"),
                  (double*)state_ptr(c, "psi_h0"), (double*)state_ptr(c, "phi_m"), (double*)state_ptr(c, "phi_h"),
                  (double*)state_ptr(c, "ustar_avg"), c->scal.theta_ref);
  }
  // without scalars obukhov's neutral branch leaves psi = 0, phi = 1: the
  // state's constant contents
  // the x-spectral rows of the filtered u, v, w and their x, y derivatives
  // stay in S_FSPEC0 for the convective inputs (nothing changes u, v, w
  // before convec)
  SpecIn<T> si;
  op_filt_da3<T>(c, u, v, w, gr, g.nz, S_FSPEC0, &si.uvw);
  op_ddz3<T>(c, gr[2], gr[5], gr[8], u, v, w);
  op_wallstress<T>(c, u, v, gr[2], gr[5], tau[2], tau[4], (const double*)state_ptr(c, "psi_m"),
                   (const double*)state_ptr(c, "psi_m0"), (const double*)state_ptr(c, "phi_m"),
                   (double*)state_ptr(c, "ustar_avg"));
  const T* grc[9] = {gr[0], gr[1], gr[2], gr[3], gr[4], gr[5], gr[6], gr[7], gr[8]};
  if (scal && (c->scal.model == 7 || c->sgs_model == 6)) {
    const ScalarSgsPtrs ss = scalar_sgs_ptrs<T>(c);
    op_sgs_stag<T>(c, grc, u, v, w, tau, sgs_ptrs<T>(c), F<T>(c, "theta"), &ss);
  } else {
    op_sgs_stag<T>(c, grc, u, v, w, tau, sgs_ptrs<T>(c));
  }
  ScalarJob jobs[2];
  if (scal) {  // main.f90:273-276
    scalar_jobs<T>(c, jobs);
    scalars_prep<T>(c, jobs, &si);
  }
  T *dx = F<T>(c, "divtx"), *dy = F<T>(c, "divty"), *dz = F<T>(c, "divtz");
  T* const dv3[3] = {dx, dy, dz};
  op_divstress3<T>(c, dv3, tau);
  T *Rx = F<T>(c, "RHSx"), *Ry = F<T>(c, "RHSy"), *Rz = F<T>(c, "RHSz");
  const int euler = (c->jt == 1 && !g.initu) ? 1 : 0;
  // Coriolis and sponge (main.f90:319-348, scalars_module.f90:143-148): the
  // products and plane means of the fields before the update
  const ForcingArgs<T> fo = op_forcing_pre<T>(c, u, v, w, scal ? (const T*)F<T>(c, "theta") : (const T*)nullptr,
                                              scal ? (const T*)F<T>(c, "qmix") : (const T*)nullptr);
  const ForcingArgs<T>* fop = (fo.cor[0] || fo.damp) ? &fo : nullptr;
  if (scal) {
    // convec + RHS assembly + AB2 (main.f90:304-390) and the scalars'
    // scalar_RHS_calc + step_scalar in one padded-grid pass
    const ScalarParams& P = c->scal;
    T* R3[3] = {Rx, Ry, Rz};
    T* Rf3[3] = {F<T>(c, "RHSx_f"), F<T>(c, "RHSy_f"), F<T>(c, "RHSz_f")};
    const T* dv3[3] = {dx, dy, dz};
    const T* grc2[9] = {gr[0], gr[1], gr[2], gr[3], gr[4], gr[5], gr[6], gr[7], gr[8]};
    op_convec_scalar_ab2<T>(c, u, v, w, grc2, R3, Rf3, dv3, euler,
                            P.passive ? (const T*)nullptr : (const T*)F<T>(c, "beta_scal"), jobs, F<T>(c, "Nu_t"),
                            (P.model == 7 || c->sgs_model == 6) ? (const T*)F<T>(c, "Pr_t") : (const T*)nullptr,
                            (c->jt == 1 && !P.initsc) ? 1 : 0, fop, &si);
  } else {
    // convec + RHS assembly + AB2 predictor in one epilogue (main.f90:304-390)
    op_convec_ab2<T>(c, u, v, w, gr[1], gr[2], gr[3], gr[5], gr[6], gr[7], Rx, Ry, Rz, F<T>(c, "RHSx_f"),
                     F<T>(c, "RHSy_f"), F<T>(c, "RHSz_f"), dx, dy, dz, euler, nullptr, fop, &si);
  }
  T *p = (T*)state_ptr(c, "p"), *px = F<T>(c, "dpdx"), *py = F<T>(c, "dpdy"), *pz = F<T>(c, "dpdz");
  op_press_stag<T>(c, u, v, w, dz, p, px, py);
  op_pressure_update<T>(c, u, v, w, Rx, Ry, Rz, p, px, py, pz);
}

// Mason's wall-damped mixing length of model 1, sgs_stag.f90:84-94
// (lbc_mom 'wall', Co = 0.16, nnn = 2, vonk = 0.41): l(jz)^2 per level, z on
// the first uv-node at jz = 1 and on w-nodes above.  Integer powers as
// products, the -1/nnn power by pow.
std::vector<double> mason_l2(const Grid& g) {
#pragma clang fp contract(off)
  const double Co = 0.16, vonk = 0.41;
  std::vector<double> l2(g.nz);
  for (int k = 0; k < g.nz; ++k) {
    const double zz = k == 0 ? (1.0 - 0.5) * g.dz : (double)k * g.dz;
    const double a = vonk * zz;
    const double l = pow(Co * Co * (1.0 / (a * a)) + 1.0 / (g.delta * g.delta), -1.0 / 2.0);
    l2[k] = l * l;
  }
  return l2;
}

// coriol (param.f90:220, a default-real literal), ug, vg (main.f90:131-138)
// and setsponge (topbc.f90:14-74, nproc = 1) with L_z = Lz z_i
void set_forcing_constants(const Grid& g, Forcing& F) {
#pragma clang fp contract(off)
  const int nz = g.nz;
  F.coriol = (double)1.3962634E-04f * F.z_i / F.u_star;
  F.cug = F.coriol * (F.ug0 / F.u_star);
  F.cvg = F.coriol * (F.vg0 / F.u_star);
  F.sponge.assign(nz + 1, 0.0);
  if (F.ubc != 1) return;
  double* sp = F.sponge.data();
  const double z_i = F.z_i, L_z = g.Lz * z_i, dz = g.dz;
  if (F.damping_method == 2) {
    const double z_d = 0.75 * (L_z * 1.0), cfrdmp = 3.9;
    for (int k = 1; k <= nz - 1; ++k) {
      const double z_local = ((double)k - 0.5) * dz * z_i;
      sp[k] = (z_local >= z_d && z_local <= L_z * 1.0)
                  ? cfrdmp * 0.5 * (1.0 - cos(PI * (z_local - z_d) / (L_z - z_d)))
                  : 0.0;
    }
    sp[nz] = sp[nz - 1];
  } else {
    // nz - 0.75*nz + 1 and the loop bound 0.75*nz + 1 are default-real
    // expressions; the bound is truncated to an integer
    const float den = (float)nz - 0.75f * (float)nz + 1.0f;
    const double factor = 9.0 / (double)den;
    sp[nz] = z_i / (50.0 * F.u_star);
    const int kend = (int)(0.75f * (float)nz + 1.0f);
    for (int k = nz - 1; k >= kend; --k) sp[k] = sp[k + 1] / pow(5.0, factor);
  }
}

// a captured step window bakes the model and forcing switches in
static void drop_graph(les_ctx* c) {
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
}

}  // namespace les

// ------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------
#define GUARD(c)                     \
  if (!(c)) return -1;               \
  (c)->err.clear();
#define DONE(c) return (c)->err.empty() ? 0 : 1;

extern "C" {

int les_create(const les_grid* gi, les_ctx** out) {
  if (!gi || !out) return -1;
  *out = nullptr;
  les_ctx* c = new les_ctx();
  Grid& g = c->g;
  g.nx = gi->nx; g.ny = gi->ny; g.nz = gi->nz;
  g.lh = g.nx / 2 + 1; g.ld = 2 * g.lh; g.nxh = g.nx / 2;
  g.nx2 = 3 * g.nx / 2; g.ny2 = 3 * g.ny / 2;
  g.Lx = gi->Lx; g.Ly = gi->Ly; g.Lz = gi->Lz;
  g.dx = g.Lx / g.nx; g.dy = g.Ly / g.ny; g.dz = g.Lz / (g.nz - 1);
  g.dt = gi->dt; g.z0 = gi->z0; g.force = gi->force;
  g.delta = sqrt(g.dx * g.dy);
  g.cs_count = gi->cs_count > 0 ? gi->cs_count : 5;
  g.inilag = gi->inilag; g.initu = gi->initu;
  c->tune.xs_tiled = g.nxh % 8 == 0;  // tiled x-spectral intermediates wherever rows allow them
  c->p_nx = make_plan(g.nx); c->p_ny = make_plan(g.ny);
  c->p_nx2 = make_plan(g.nx2); c->p_ny2 = make_plan(g.ny2);
  if (c->p_nx.n < 0 || c->p_ny.n < 0 || c->p_nx2.n < 0 || c->p_ny2.n < 0 || g.nx > 1024 || g.ny > 1024 ||
      g.nz < 3 || g.nx % 8 || g.ny % 8 || (long)g.ld * g.ny * (g.nz + 1) >= (1L << 31)) {
    fprintf(stderr, "les_create: nx, ny must be multiples of 8 of the form 2^a 3^b (<= 1024), nz >= 3, "
                    "and ld*ny*(nz+1) < 2^31\n");
    delete c;
    return 2;
  }
  check(c, hipGetDevice(&c->device), "hipGetDevice");
  if (c->err.empty()) noalloc_init(c->device, &c->err);
  HostTables h;
  host_tables(g, h);
  c->keep2 = keep_cols(g, h.G2);
  c->keep4 = keep_cols(g, h.G4);
  make_tables<double>(c, c->td, h);
  make_tables<float>(c, c->tf, h);
  c->G2_host = new double[h.G2.size()];
  std::memcpy(c->G2_host, h.G2.data(), h.G2.size() * sizeof(double));
  if (!c->err.empty()) {
    fprintf(stderr, "les_create: %s\n", c->err.c_str());
    les_destroy(c);
    return 1;
  }
  *out = c;
  return 0;
}

void les_destroy(les_ctx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  free_tables(c->td);
  free_tables(c->tf);
  for (void* p : c->arena.ptr)
    if (p) (void)hipFree(p);
  for (auto& m : c->mirrors)
    if (m.second.first) (void)hipFree(m.second.first);
  if (c->st) {
    for (void* p : c->st->owned) (void)hipFree(p);
    delete c->st;
  }
  delete[] c->G2_host;
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  les::mlp_free(c);
  if (c->wt_s) (void)hipFree(c->wt_s);
  if (c->wq_s) (void)hipFree(c->wq_s);
  if (c->l2z) (void)hipFree(c->l2z);
  if (c->frc.dcoef) (void)hipFree(c->frc.dcoef);
  delete c;
}

const char* les_last_error(les_ctx* c) { return c ? c->err.c_str() : "null context"; }

int les_set_stream(les_ctx* c, void* s) {
  GUARD(c);
  c->stream = (hipStream_t)s;
  DONE(c);
}
int les_get_grid(les_ctx* c, les_grid* o) {
  GUARD(c);
  if (!o) return -1;
  const Grid& g = c->g;
  o->nx = g.nx; o->ny = g.ny; o->nz = g.nz;
  o->Lx = g.Lx; o->Ly = g.Ly; o->Lz = g.Lz;
  o->dt = g.dt; o->z0 = g.z0; o->force = g.force;
  o->inilag = g.inilag; o->initu = g.initu; o->cs_count = g.cs_count;
  DONE(c);
}
int les_sync(les_ctx* c) {
  GUARD(c);
  check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  DONE(c);
}
int les_set_time(les_ctx* c, int jt, int jt_total) {
  GUARD(c);
  c->jt = jt;
  c->jt_total = jt_total;
  DONE(c);
}
int les_get_time(les_ctx* c, int* jt, int* jt_total) {
  GUARD(c);
  if (jt) *jt = c->jt;
  if (jt_total) *jt_total = c->jt_total;
  DONE(c);
}
int les_set_lag_init(les_ctx* c, int a, int b) {
  GUARD(c);
  c->flm_init = a;
  c->fqn_init = b;
  DONE(c);
}
int les_get_tables(les_ctx* c, double* kx, double* ky, double* k2, double* G, double* GG) {
  GUARD(c);
  HostTables h;
  host_tables(c->g, h);
  const size_t n = h.kx.size() * sizeof(double);
  if (kx) std::memcpy(kx, h.kx.data(), n);
  if (ky) std::memcpy(ky, h.ky.data(), n);
  if (k2) std::memcpy(k2, h.k2.data(), n);
  if (G) std::memcpy(G, h.G2.data(), n);
  if (GG) std::memcpy(GG, h.G4.data(), n);
  DONE(c);
}

// ---- routine-level entry points ----
int les_filt_da(les_ctx* c, double* f, double* dx, double* dy, int nlev) {
  GUARD(c);
  op_filt_da<double>(c, f, dx, dy, nlev);
  DONE(c);
}
int les_filt_da_f32(les_ctx* c, float* f, float* dx, float* dy, int nlev) {
  GUARD(c);
  op_filt_da<float>(c, f, dx, dy, nlev);
  DONE(c);
}
int les_ddx(les_ctx* c, double* dfdx, const double* f, int nlev) {
  GUARD(c);
  op_ddxy<double>(c, dfdx, nullptr, f, nlev);
  DONE(c);
}
int les_ddy(les_ctx* c, double* dfdy, const double* f, int nlev) {
  GUARD(c);
  op_ddxy<double>(c, nullptr, dfdy, f, nlev);
  DONE(c);
}
int les_ddx_f32(les_ctx* c, float* dfdx, const float* f, int nlev) {
  GUARD(c);
  op_ddxy<float>(c, dfdx, nullptr, f, nlev);
  DONE(c);
}
int les_ddy_f32(les_ctx* c, float* dfdy, const float* f, int nlev) {
  GUARD(c);
  op_ddxy<float>(c, nullptr, dfdy, f, nlev);
  DONE(c);
}
int les_ddxy(les_ctx* c, double* dfdx, double* dfdy, const double* f, int nlev) {
  GUARD(c);
  op_ddxy<double>(c, dfdx, dfdy, f, nlev);
  DONE(c);
}
int les_test_filter_init(les_ctx* c, double alpha, double* G) {
  GUARD(c);
  HostTables h;
  host_tables(c->g, h);
  std::vector<double> out((size_t)c->g.lh * c->g.ny);
  const double delta = alpha * sqrt(c->g.dx * c->g.dy);
  const double q = PI / delta, kc2 = q * q;
  for (size_t i = 0; i < out.size(); ++i) out[i] = h.k2[i] >= kc2 ? 0.0 : 1.0 / ((double)c->g.nx * c->g.ny);
  for (int y = 0; y < c->g.ny; ++y) out[c->g.lh - 1 + (size_t)c->g.lh * y] = 0.0;
  for (int x = 0; x < c->g.lh; ++x) out[x + (size_t)c->g.lh * (c->g.ny / 2)] = 0.0;
  check(c, hipMemcpyAsync(G, out.data(), out.size() * sizeof(double), hipMemcpyHostToDevice, c->stream),
        "test_filter_init upload");
  check(c, hipStreamSynchronize(c->stream), "sync");
  DONE(c);
}
int les_test_filter(les_ctx* c, double* plane, const double* G) {
  GUARD(c);
  op_test_filter<double>(c, plane, 1, G, c->g.nxh);
  DONE(c);
}
int les_test_filter_f32(les_ctx* c, float* plane, const float* G) {
  GUARD(c);
  op_test_filter<float>(c, plane, 1, G, c->g.nxh);
  DONE(c);
}
int les_padd(les_ctx* c, double* u_big, const double* u) {
  GUARD(c);
  op_padd<double>(c, u_big, u, 1);
  DONE(c);
}
int les_unpadd(les_ctx* c, double* cc, const double* cc_big) {
  GUARD(c);
  op_unpadd<double>(c, cc, cc_big, 1);
  DONE(c);
}
int les_dealias1(les_ctx* c, const double* u, double* u_big, int nlev) {
  GUARD(c);
  op_dealias1<double>(c, u, u_big, nlev);
  DONE(c);
}
int les_dealias2(les_ctx* c, double* u, double* u_big, int nlev) {
  GUARD(c);
  op_dealias2<double>(c, u, u_big, nlev);
  DONE(c);
}
int les_dealias1_f32(les_ctx* c, const float* u, float* u_big, int nlev) {
  GUARD(c);
  op_dealias1<float>(c, u, u_big, nlev);
  DONE(c);
}
int les_dealias2_f32(les_ctx* c, float* u, float* u_big, int nlev) {
  GUARD(c);
  op_dealias2<float>(c, u, u_big, nlev);
  DONE(c);
}
int les_ddz_uv(les_ctx* c, double* dfdz, const double* f) {
  GUARD(c);
  op_ddz_uv<double>(c, dfdz, f);
  DONE(c);
}
int les_ddz_w(les_ctx* c, double* dfdz, const double* f) {
  GUARD(c);
  op_ddz_w<double>(c, dfdz, f);
  DONE(c);
}
int les_convec(les_ctx* c, const double* u, const double* v, const double* w, const double* dudy,
               const double* dudz, const double* dvdx, const double* dvdz, const double* dwdx, const double* dwdy,
               double* cx, double* cy, double* cz) {
  GUARD(c);
  op_convec<double>(c, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz);
  DONE(c);
}
int les_convec_f32(les_ctx* c, const float* u, const float* v, const float* w, const float* dudy,
                   const float* dudz, const float* dvdx, const float* dvdz, const float* dwdx, const float* dwdy,
                   float* cx, float* cy, float* cz) {
  GUARD(c);
  op_convec<float>(c, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz);
  DONE(c);
}
int les_wallstress(les_ctx* c, const double* u, const double* v, double* dudz, double* dvdz, double* txz,
                   double* tyz, const double* psi_m, const double* psi_m0, const double* phi_m, double* ustar) {
  GUARD(c);
  op_wallstress<double>(c, u, v, dudz, dvdz, txz, tyz, psi_m, psi_m0, phi_m, ustar);
  DONE(c);
}
static SgsPtrs to_ptrs(const les_sgs_fields* s) {
  SgsPtrs p;
  p.F_LM = s->F_LM; p.F_MM = s->F_MM; p.F_QN = s->F_QN; p.F_NN = s->F_NN; p.Beta = s->Beta;
  p.Cs_opt2 = s->Cs_opt2; p.Nu_t = s->Nu_t; p.u_lag = s->u_lag; p.v_lag = s->v_lag; p.w_lag = s->w_lag;
  p.TnLMMM = s->TnLMMM; p.TnQNNN = s->TnQNNN; p.L11t = s->L11t; p.L22t = s->L22t; p.L33t = s->L33t;
  p.Q11t = s->Q11t; p.Q22t = s->Q22t; p.Q33t = s->Q33t; p.Beta_avg = s->Beta_avg; p.Betaclip_avg = s->Betaclip_avg;
  return p;
}
int les_calc_sij(les_ctx* c, double* const grad[9], double* const S[6]) {
  GUARD(c);
  const double* g[9];
  for (int i = 0; i < 9; ++i) g[i] = grad[i];
  op_calc_sij<double>(c, g, S);
  DONE(c);
}
int les_sgs_stag(les_ctx* c, double* const grad[9], const double* u, const double* v, const double* w,
                 double* const tau[6], const les_sgs_fields* s) {
  GUARD(c);
  if (!s) return -1;
  const double* g[9];
  for (int i = 0; i < 9; ++i) g[i] = grad[i];
  op_sgs_stag<double>(c, g, u, v, w, tau, to_ptrs(s));
  DONE(c);
}
int les_lagrange_sdep(les_ctx* c, double* const S[6], const double* u, const double* v, const double* w,
                      const les_sgs_fields* s) {
  GUARD(c);
  if (!s) return -1;
  const double* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
  op_lagrange<double>(c, Sc, u, v, w, to_ptrs(s));
  DONE(c);
}
int les_set_sgs_model(les_ctx* c, int model) {
  GUARD(c);
  if (model < 1 || model > 6) {
    c->err = "les_set_sgs_model: model must be 1 (Smagorinsky-Mason), 2 (standard dynamic), 3 (plane-averaged "
             "scale-dependent), 4 (Lagrangian scale-invariant), 5 (Lagrangian scale-dependent) or 6 "
             "(plane-averaged scale-dependent, new_scaledep_dynamic)";
    return 1;
  }
  if (model != 5 && c->scal.enable && c->scal.model == 7) {
    c->err = "les_set_sgs_model: the scalar Lagrangian model (model 7) runs with lagrange_Sdep, SGS model 5 "
             "(sgs_stag.f90:150-153)";
    return 1;
  }
  if (model == 6 && c->scal.enable && c->scal.model != 6) {
    c->err = "les_set_sgs_model: SGS model 6 takes Pr_t from scalar_scaledep_dynamic (scalars_module.f90:77-84); "
             "set the scalar model to 6";
    return 1;
  }
  if (model != 5 && c->closure_mode) {
    c->err = "les_set_sgs_model: the learned closure replaces the model-5 update; switch it off first";
    return 1;
  }
  if (model == 1) {
    const Grid& g = c->g;
    std::vector<double> l2 = mason_l2(g);
    if (!c->l2z) check(c, hipMalloc((void**)&c->l2z, g.nz * sizeof(double)), "hipMalloc l2z");
    if (c->err.empty()) check(c, hipMemcpy(c->l2z, l2.data(), g.nz * sizeof(double), hipMemcpyHostToDevice), "l2z");
    if (!c->err.empty()) return 1;
  }
  c->sgs_model = model;
  drop_graph(c);
  DONE(c);
}
int les_std_dynamic(les_ctx* c, double* const S[6], const double* u, const double* v, const double* w,
                    double* Cs_opt2) {
  GUARD(c);
  if (!S || !Cs_opt2) return -1;
  const double* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
  op_std_dynamic<double>(c, Sc, u, v, w, Cs_opt2, (double*)nullptr);
  DONE(c);
}
int les_new_scaledep_dynamic(les_ctx* c, double* const S[6], const double* u, const double* v, const double* w,
                             double* Cs_opt2) {
  GUARD(c);
  if (!S || !Cs_opt2) return -1;
  const double* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
  op_new_scaledep_dynamic<double>(c, Sc, u, v, w, Cs_opt2, (double*)nullptr);
  DONE(c);
}
int les_lagrange_ssim(les_ctx* c, double* const S[6], const double* u, const double* v, const double* w,
                      const les_sgs_fields* s) {
  GUARD(c);
  if (!S || !s) return -1;
  const double* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
  op_lagrange_ssim<double>(c, Sc, u, v, w, to_ptrs(s));
  DONE(c);
}
int les_interpolag_ssim(les_ctx* c, const les_sgs_fields* s) {
  GUARD(c);
  if (!s) return -1;
  op_interpolag_ssim<double>(c, to_ptrs(s));
  DONE(c);
}

// Coriolis forcing and the top sponge
int les_set_forcing(les_ctx* c, const les_forcing* f) {
  GUARD(c);
  if (!f) return -1;
  if (f->ubc != 0 && f->ubc != 1) { c->err = "les_set_forcing: ubc must be 0 or 1"; return 1; }
  if (f->ubc == 1 && f->damping_method != 1 && f->damping_method != 2) {
    c->err = "les_set_forcing: damping_method must be 1 or 2";
    return 1;
  }
  if ((f->coriolis_forcing || f->ubc == 1) && !(f->u_star > 0.0 && f->z_i > 0.0)) {
    c->err = "les_set_forcing: u_star and z_i must be positive";
    return 1;
  }
  const Grid& g = c->g;
  const int nl = g.nz - 1;
  Forcing& F = c->frc;
  F.coriolis = f->coriolis_forcing ? 1 : 0;
  F.ubc = f->ubc;
  F.damping_method = f->damping_method;
  F.ug0 = f->ug0; F.vg0 = f->vg0; F.u_star = f->u_star; F.z_i = f->z_i;
  set_forcing_constants(g, F);
  std::vector<double> coef(2 * (size_t)nl, 0.0);
  for (int k = 0; k < nl; ++k) {
    coef[k] = 0.5 * (F.sponge[k + 1] + F.sponge[k + 2]);
    coef[nl + k] = F.damping_method == 2 ? 0.5 * F.sponge[k + 1] : F.sponge[k + 1];
  }
  if (!F.dcoef) check(c, hipMalloc((void**)&F.dcoef, coef.size() * sizeof(double)), "hipMalloc forcing");
  if (c->err.empty())
    check(c, hipMemcpy(F.dcoef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice), "forcing");
  if (!c->err.empty()) return 1;
  drop_graph(c);
  DONE(c);
}
int les_get_sponge(les_ctx* c, double* sponge) {
  GUARD(c);
  if (!sponge) return -1;
  const int nz = c->g.nz;
  for (int k = 0; k < nz; ++k) sponge[k] = c->frc.sponge.empty() ? 0.0 : c->frc.sponge[k + 1];
  DONE(c);
}
int les_scaledep_dynamic(les_ctx* c, double* const S[6], const double* u, const double* v, const double* w,
                         double* Cs_opt2, double* beta) {
  GUARD(c);
  if (!S || !Cs_opt2) return -1;
  const double* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
  op_scaledep_dynamic<double>(c, Sc, u, v, w, Cs_opt2, beta, (double*)nullptr);
  DONE(c);
}
int les_interpolag_sdep(les_ctx* c, const les_sgs_fields* s) {
  GUARD(c);
  if (!s) return -1;
  op_interpolag<double>(c, to_ptrs(s));
  DONE(c);
}
int les_divstress_uv(les_ctx* c, double* divt, const double* tx, const double* ty, const double* tz) {
  GUARD(c);
  op_divstress<double>(c, divt, tx, ty, tz, 0);
  DONE(c);
}
int les_divstress_w(les_ctx* c, double* divt, const double* tx, const double* ty, const double* tz) {
  GUARD(c);
  op_divstress<double>(c, divt, tx, ty, tz, 1);
  DONE(c);
}
int les_divstress_uv_f32(les_ctx* c, float* divt, const float* tx, const float* ty, const float* tz) {
  GUARD(c);
  op_divstress<float>(c, divt, tx, ty, tz, 0);
  DONE(c);
}
int les_divstress_w_f32(les_ctx* c, float* divt, const float* tx, const float* ty, const float* tz) {
  GUARD(c);
  op_divstress<float>(c, divt, tx, ty, tz, 1);
  DONE(c);
}
int les_press_stag_array(les_ctx* c, const double* u, const double* v, const double* w, const double* divtz,
                         double* p, double* dpdx, double* dpdy) {
  GUARD(c);
  op_press_stag<double>(c, u, v, w, divtz, p, dpdx, dpdy);
  DONE(c);
}
int les_press_stag_array_f32(les_ctx* c, const float* u, const float* v, const float* w, const float* divtz,
                             float* p, float* dpdx, float* dpdy) {
  GUARD(c);
  op_press_stag<float>(c, u, v, w, divtz, p, dpdx, dpdy);
  DONE(c);
}
int les_tridag_array(les_ctx* c, const double* a, const double* b, const double* cc, const double* r, double* u) {
  GUARD(c);
  int e = op_tridag_array(c, a, b, cc, r, u);
  if (e && c->err.empty()) c->err = e == 1 ? "tridag_array: rewrite eqs (b(1) == 0)" : "tridag_array failed (zero pivot)";
  DONE(c);
}
int les_project(les_ctx* c, double* u, double* v, double* w, const double* dpdx, const double* dpdy,
                const double* dpdz) {
  GUARD(c);
  op_project<double>(c, u, v, w, dpdx, dpdy, dpdz);
  DONE(c);
}
int les_rmsdiv(les_ctx* c, const double* dudx, const double* dvdy, const double* dwdz, double* rms) {
  GUARD(c);
  double r = op_rmsdiv<double>(c, dudx, dvdy, dwdz);
  if (rms) *rms = r;
  DONE(c);
}
int les_timestep_conditions(les_ctx* c, const double* u, const double* v, const double* w, const double* nu,
                            double out[4]) {
  GUARD(c);
  op_timestep_conditions<double>(c, u, v, w, nu, out);
  DONE(c);
}

// ---- whole-step state ----
int les_state_alloc(les_ctx* c, int precision) {
  GUARD(c);
  if (c->st) {
    for (void* p : c->st->owned) (void)hipFree(p);
    delete c->st;
  }
  State* s = new State();
  s->prec = precision ? 1 : 0;
  s->es = s->prec ? 4 : 8;
  c->st = s;
  const Grid& g = c->g;
  const size_t pl = (size_t)g.ld * g.ny;
  auto add = [&](const char* name, int nlev, size_t es) {
    void* p = nullptr;
    check(c, hipMalloc(&p, pl * nlev * es), "state alloc");
    if (p) check(c, hipMemset(p, 0, pl * nlev * es), "state zero");
    s->owned.push_back(p);
    s->f[name] = FieldInfo{p, nlev, (int)es};
  };
  for (const char* n : F3) add(n, g.nz, s->es);
  add("p", g.nz + 1, s->es);
  auto add2 = [&](const char* name, size_t n, double val) {
    void* p = nullptr;
    check(c, hipMalloc(&p, n * sizeof(double)), "state alloc");
    std::vector<double> h(n, val);
    if (p) check(c, hipMemcpy(p, h.data(), n * sizeof(double), hipMemcpyHostToDevice), "state init");
    s->owned.push_back(p);
    s->f[name] = FieldInfo{p, 1, 8};
  };
  const size_t n2 = (size_t)g.nx * g.ny;
  add2("psi_m", n2, 0.0);
  add2("psi_m0", n2, 0.0);
  add2("phi_m", n2, 1.0);
  add2("ustar_avg", n2, 0.0);
  add2("Beta_avg", g.nz, 0.0);
  add2("Betaclip_avg", g.nz, 0.0);
  if (c->scal.enable) {
    s->scal = 1;
    for (const char* n : F3S) add(n, g.nz, s->es);
    add2("psi_h", n2, 0.0);
    add2("psi_h0", n2, 0.0);
    add2("phi_h", n2, 1.0);
    add2("sgs_t3", pl, 0.0);  // level-1 planes (ld,ny) of sgs_t3 / sgs_q3, FP64
    add2("sgs_q3", pl, 0.0);
  }
  DONE(c);
}

void* les_state_field(les_ctx* c, const char* name, int* nlev, int* elsize) {
  if (!c || !c->st || !name) return nullptr;
  void* p = state_ptr(c, name);
  if (!p) return nullptr;
  std::string n(name);
  auto it = c->st->f.begin();
  for (; it != c->st->f.end(); ++it)
    if (it->second.p == p) break;
  if (nlev) *nlev = it->second.nlev;
  if (elsize) *elsize = it->second.elsize;
  return p;
}

static uint64_t splitmix(uint64_t seed, uint64_t index) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + index + 0x632BE59BD9B4E019ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int les_state_init_ic(les_ctx* c, uint64_t seed, double amp, double ustar, double kappa_ic) {
  GUARD(c);
  if (!c->st) return -1;
  State* s = c->st;
  const Grid& g = c->g;
  const size_t pl = (size_t)g.ld * g.ny, n3 = pl * g.nz;
  std::vector<double> f[3];
  for (int q = 0; q < 3; ++q) f[q].assign(n3, 0.0);
  for (int k = 0; k < g.nz - 1; ++k) {
    const double z = (k + 0.5) * g.dz;
    const double ulog = (ustar / kappa_ic) * log(z / g.z0);
    for (int y = 0; y < g.ny; ++y)
      for (int x = 0; x < g.nx; ++x) {
        const size_t i = pl * k + x + (size_t)g.ld * y;
        f[0][i] = ulog;
        if (z < 0.5)
          for (int q = 0; q < 3; ++q) {
            const uint64_t idx = (((uint64_t)q * g.nz + k) * g.ny + y) * g.nx + x;
            const double U = (double)(splitmix(seed, idx) >> 11) * (1.0 / 9007199254740992.0);
            f[q][i] += amp * ustar * (2.0 * U - 1.0);
          }
      }
  }
  for (size_t i = 0; i < pl; ++i) {
    f[2][i] = 0.0;
    f[0][pl * (g.nz - 1) + i] = f[0][pl * (g.nz - 2) + i];
    f[1][pl * (g.nz - 1) + i] = f[1][pl * (g.nz - 2) + i];
    f[2][pl * (g.nz - 1) + i] = 0.0;
  }
  for (auto& kv : s->f) {
    const std::string& n = kv.first;
    if (n == "phi_m" || n == "psi_m" || n == "psi_m0" || n == "phi_h" || n == "psi_h" || n == "psi_h0") continue;
    const size_t bytes = (n == "ustar_avg") ? (size_t)g.nx * g.ny * 8
                         : (n == "Beta_avg" || n == "Betaclip_avg") ? (size_t)g.nz * 8
                         : (n == "sgs_t3" || n == "sgs_q3") ? pl * 8
                         : pl * kv.second.nlev * s->es;
    check(c, hipMemsetAsync(kv.second.p, 0, bytes, c->stream), "state reset");
  }
  s->rhs_swap = 0;
  const char* names[3] = {"u", "v", "w"};
  for (int q = 0; q < 3; ++q) {
    void* d = state_ptr(c, names[q]);
    if (s->prec) {
      std::vector<float> h(n3);
      for (size_t i = 0; i < n3; ++i) h[i] = (float)f[q][i];
      check(c, hipMemcpyAsync(d, h.data(), n3 * 4, hipMemcpyHostToDevice, c->stream), "ic upload");
      check(c, hipStreamSynchronize(c->stream), "ic sync");
    } else {
      check(c, hipMemcpyAsync(d, f[q].data(), n3 * 8, hipMemcpyHostToDevice, c->stream), "ic upload");
      check(c, hipStreamSynchronize(c->stream), "ic sync");
    }
  }
  c->jt = 0;
  c->jt_total = 0;
  c->flm_init = 0;
  c->fqn_init = 0;
  c->glm_init = 0;
  c->gqn_init = 0;
  c->eager_steps = 0;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  DONE(c);
}

static void step_eager(les_ctx* c) {
  if (c->st->prec)
    step_once<float>(c);
  else
    step_once<double>(c);
  c->eager_steps += 1;
}

// The host-side branches of one step depend on jt mod cs_count (dynamic
// update), the RHS buffer parity, jt == 1 (Euler start), the lagrange_Sdep
// initialisation (jt == cs_count) and jt_total < 500 (constant Cs,
// sgs_stag.f90:166-173).  A window of W = lcm(2, cs_count) steps therefore
// repeats exactly while none of the one-off conditions occurs inside it; such
// a window is captured once into a HIP graph and replayed.
static int graph_window(const les_ctx* c) {
  int cs = c->g.cs_count;
  if (c->closure_mode && c->closure_period > cs) cs = c->closure_period;  // a multiple of cs_count
  return cs % 2 == 0 ? cs : 2 * cs;
}
static bool window_is_regular(const les_ctx* c, int W, int* use_const) {
  const Grid& g = c->g;
  if (c->jt < 1 && !g.initu) return false;                      // Euler first step inside
  const bool lag_init = g.inilag && (!c->flm_init || !c->fqn_init ||
                                     (c->st->scal && c->scal.model == 7 && (!c->glm_init || !c->gqn_init)));
  if (lag_init && c->jt < lag_count(c)) return false;           // initialisation inside
  const int first = c->jt_total + 1, last = c->jt_total + W;
  if (first < 500 && last >= 500) return false;                // constant-Cs switch inside
  // every routine must have run eagerly once, so scratch is at full size
  if (c->eager_steps < W) return false;
  *use_const = last < 500 ? 1 : 0;
  return true;
}
static bool graph_matches(const les_ctx* c, int W, int uc) {
  return c->gexec && c->g_start_mod == c->jt % W && c->g_use_const == uc && c->g_swap == c->st->rhs_swap;
}
// capture W steps from the current position into c->gexec without running
// them; host-side counters are restored afterwards
static void graph_capture(les_ctx* c, int W, int uc) {
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  const int jt0 = c->jt, jtt0 = c->jt_total, swap0 = c->st->rhs_swap, fl0 = c->flm_init, fq0 = c->fqn_init;
  const int gl0 = c->glm_init, gq0 = c->gqn_init;
  const long es0 = c->eager_steps;
  check(c, hipStreamSynchronize(c->stream), "pre-capture sync");
  hipGraph_t graph = nullptr;
  check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  for (int i = 0; i < W; ++i) step_eager(c);
  check(c, hipStreamEndCapture(c->stream, &graph), "hipStreamEndCapture");
  c->jt = jt0;
  c->jt_total = jtt0;
  c->st->rhs_swap = swap0;
  c->flm_init = fl0;
  c->glm_init = gl0;
  c->gqn_init = gq0;
  c->fqn_init = fq0;
  c->eager_steps = es0;
  if (graph && c->err.empty()) {
    check(c, hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    c->g_start_mod = jt0 % W;
    c->g_use_const = uc;
    c->g_swap = swap0;
  }
  if (graph) (void)hipGraphDestroy(graph);
}

// host-side counters after a replayed window of W steps (the captured steps
// advanced and then restored them): W is even, so the RHS buffer parity is
// unchanged; the last Lagrangian update is the last multiple of its period
static void advance_after_replay(les_ctx* c, int W) {
  const int jt0 = c->jt;
  c->jt += W;
  c->jt_total += W;
  const int lc = lag_count(c), last = (c->jt / lc) * lc;
  if (c->sgs_model == 5 && last > jt0 && last > 1) c->last_lag_jt = last;
}

}  // extern "C"

// ------------------------------------------------------------------------
// ensembles: independent contexts (grids, surface parameters, precisions)
// advanced together.  One HIP graph holds a window of W steps of every
// member: the capture stream forks to each member's stream, the member's
// steps are recorded there, and the branches join back, so a window of the
// whole ensemble is one hipGraphLaunch and independent branches may run
// concurrently.  Windows in which any member has a one-off condition run
// eagerly, member by member, each on its own stream.
// ------------------------------------------------------------------------
struct les_ensemble {
  std::vector<les_ctx*> m;
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> own;  // streams created for members that had none
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t gexec = nullptr;
  std::vector<int> key_mod, key_uc, key_swap;
  int W = 2;
  std::string err;
};

static int gcd_i(int a, int b) { return b ? gcd_i(b, a % b) : a; }

static bool ens_check(les_ensemble* e) {
  for (les_ctx* c : e->m)
    if (!c->err.empty()) {
      e->err = "member: " + c->err;
      return false;
    }
  return e->err.empty();
}
static bool ens_regular(les_ensemble* e, std::vector<int>& uc) {
  uc.assign(e->m.size(), 0);
  for (size_t i = 0; i < e->m.size(); ++i)
    if (!window_is_regular(e->m[i], e->W, &uc[i])) return false;
  return true;
}
static bool ens_matches(const les_ensemble* e, const std::vector<int>& uc) {
  if (!e->gexec) return false;
  for (size_t i = 0; i < e->m.size(); ++i) {
    const les_ctx* c = e->m[i];
    if (e->key_mod[i] != c->jt % e->W || e->key_uc[i] != uc[i] || e->key_swap[i] != c->st->rhs_swap) return false;
  }
  return true;
}
// ensemble stream waits for every member's queued work, or every member
// waits for the ensemble stream
static void ens_join_members(les_ensemble* e) {
  for (size_t i = 0; i < e->m.size(); ++i) {
    (void)hipEventRecord(e->join[i], e->m[i]->stream);
    (void)hipStreamWaitEvent(e->stream, e->join[i], 0);
  }
}
static void ens_fork_members(les_ensemble* e) {
  (void)hipEventRecord(e->fork, e->stream);
  for (les_ctx* c : e->m) (void)hipStreamWaitEvent(c->stream, e->fork, 0);
}
static void ens_capture(les_ensemble* e, const std::vector<int>& uc) {
  if (e->gexec) {
    (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
  }
  struct Saved { int jt, jtt, swap, fl, fq, gl, gq, lag; long es; };
  std::vector<Saved> sv;
  for (les_ctx* c : e->m) {
    check(c, hipStreamSynchronize(c->stream), "ensemble pre-capture sync");
    sv.push_back({c->jt, c->jt_total, c->st->rhs_swap, c->flm_init, c->fqn_init, c->glm_init, c->gqn_init,
                  c->last_lag_jt, c->eager_steps});
  }
  (void)hipStreamSynchronize(e->stream);
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    e->err = "ensemble: hipStreamBeginCapture failed";
    return;
  }
  ens_fork_members(e);
  for (size_t i = 0; i < e->m.size(); ++i)
    for (int s = 0; s < e->W; ++s) step_eager(e->m[i]);
  ens_join_members(e);
  if (hipStreamEndCapture(e->stream, &graph) != hipSuccess) e->err = "ensemble: hipStreamEndCapture failed";
  for (size_t i = 0; i < e->m.size(); ++i) {
    les_ctx* c = e->m[i];
    const Saved& v = sv[i];
    c->jt = v.jt;
    c->jt_total = v.jtt;
    c->st->rhs_swap = v.swap;
    c->flm_init = v.fl;
    c->fqn_init = v.fq;
    c->glm_init = v.gl;
    c->gqn_init = v.gq;
    c->last_lag_jt = v.lag;
    c->eager_steps = v.es;
  }
  if (graph && ens_check(e)) {
    if (hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
      e->err = "ensemble: hipGraphInstantiate failed";
      e->gexec = nullptr;
    }
    e->key_mod.clear(); e->key_uc.clear(); e->key_swap.clear();
    for (size_t i = 0; i < e->m.size(); ++i) {
      e->key_mod.push_back(e->m[i]->jt % e->W);
      e->key_uc.push_back(uc[i]);
      e->key_swap.push_back(e->m[i]->st->rhs_swap);
    }
  }
  if (graph) (void)hipGraphDestroy(graph);
}

extern "C" {

int les_ensemble_create(les_ctx* const* members, int n, les_ensemble** out) {
  if (!members || n < 1 || !out) return -1;
  *out = nullptr;
  les_ensemble* e = new les_ensemble();
  for (int i = 0; i < n; ++i) {
    les_ctx* c = members[i];
    if (!c || !c->st) {
      delete e;
      return -1;
    }
    for (les_ctx* d : e->m)
      if (d == c) {
        delete e;
        return -1;
      }
    if (c->device != members[0]->device) {
      delete e;
      return -1;
    }
    e->m.push_back(c);
  }
  int W = 2;
  for (les_ctx* c : e->m) {
    const int w = graph_window(c);
    W = W / gcd_i(W, w) * w;
  }
  e->W = W;
  (void)hipSetDevice(members[0]->device);
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return 1;
  }
  for (les_ctx* c : e->m)
    if (!c->stream) {  // the legacy stream cannot join a capture
      hipStream_t s = nullptr;
      (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
      c->stream = s;
      e->own.push_back(s);
    }
  (void)hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  e->join.resize(e->m.size());
  for (auto& j : e->join) (void)hipEventCreateWithFlags(&j, hipEventDisableTiming);
  *out = e;
  return 0;
}

void les_ensemble_destroy(les_ensemble* e) {
  if (!e) return;
  (void)hipStreamSynchronize(e->stream);
  for (les_ctx* c : e->m) (void)hipStreamSynchronize(c->stream);
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  (void)hipEventDestroy(e->fork);
  for (auto j : e->join) (void)hipEventDestroy(j);
  // members keep working after the ensemble is gone: give back the streams
  // created here only when no member still uses them
  for (hipStream_t s : e->own) {
    bool used = false;
    for (les_ctx* c : e->m) used = used || c->stream == s;
    if (!used) (void)hipStreamDestroy(s);
  }
  (void)hipStreamDestroy(e->stream);
  delete e;
}

const char* les_ensemble_last_error(les_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int les_ensemble_window(les_ensemble* e) { return e ? e->W : -1; }

int les_ensemble_step(les_ensemble* e, int nsteps) {
  if (!e) return -1;
  e->err.clear();
  for (les_ctx* c : e->m) c->err.clear();
  int done = 0;
  std::vector<int> uc;
  while (done < nsteps && ens_check(e)) {
    if (nsteps - done >= e->W && ens_regular(e, uc)) {
      if (!ens_matches(e, uc)) ens_capture(e, uc);
      if (e->gexec && ens_check(e)) {
        ens_join_members(e);  // eager work queued on the members comes first
        if (hipGraphLaunch(e->gexec, e->stream) != hipSuccess) {
          e->err = "ensemble: hipGraphLaunch failed";
          break;
        }
        ens_fork_members(e);  // and later member work waits for the window
        for (les_ctx* c : e->m) advance_after_replay(c, e->W);
        done += e->W;
        continue;
      }
    }
    for (les_ctx* c : e->m) step_eager(c);
    ++done;
  }
  return ens_check(e) ? 0 : 1;
}

int les_ensemble_sync(les_ensemble* e) {
  if (!e) return -1;
  bool ok = hipStreamSynchronize(e->stream) == hipSuccess;
  for (les_ctx* c : e->m) ok = (hipStreamSynchronize(c->stream) == hipSuccess) && ok;
  return ok ? 0 : 1;
}

// ---- autotuner: launch shapes timed on this context's own state ----
extern "C++" {
template <class Fn>
static double tune_time(les_ctx* c, Fn&& fn, int reps = 3) {
  fn();  // warm-up: scratch sizes, kernel attributes
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, c->stream);
  for (int r = 0; r < reps; ++r) fn();
  (void)hipEventRecord(b, c->stream);
  (void)hipEventSynchronize(b);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return (double)ms / reps;
}

template <typename T>
static void tune_run(les_ctx* c, int what, std::string& rep) {
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1;
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const size_t ldb = 2 * (size_t)(g.nx2 / 2 + 1), nbig = ldb * g.ny2 * nz;
  const T* u = F<T>(c, "u");
  char buf[160];
  rep += "{";
  if (what & 1) {
    // column-FFT tile: x/y derivatives (ny -> ny), zero padding (ny -> 3ny/2)
    // and truncation (3ny/2 -> ny), the three kinds of column pass of a step
    T* A = (T*)scratch(c, S_TUNE0, n3 * sizeof(T));
    T* B = (T*)scratch(c, S_TUNE1, n3 * sizeof(T));
    T* Bg = (T*)scratch(c, S_TUNE2, nbig * sizeof(T));
    // (tile width, plan) candidates: plan 0 at three widths, plan 2 at two
    const int p = sizeof(T) == 8 ? 1 : 0;
    const int c32[5][2] = {{8, 0}, {16, 0}, {32, 0}, {8, 2}, {16, 2}};
    const int c64[5][2] = {{4, 0}, {8, 0}, {16, 0}, {4, 2}, {8, 2}};
    int best_tx = c->tune.col_tx[p], best_plan = c->tune.col_plan[p];
    double tbest = 1e30;
    rep += "\"col\": {";
    for (int i = 0; i < 5 && c->err.empty(); ++i) {
      c->tune.col_tx[p] = p ? c64[i][0] : c32[i][0];
      c->tune.col_plan[p] = p ? c64[i][1] : c32[i][1];
      const double ms = tune_time(c, [&] {
        op_ddxy<T>(c, A, B, u, nz);
        op_dealias1<T>(c, u, Bg, nz);
        op_dealias2<T>(c, A, Bg, nz);
      }, 6);  // the candidates lie within a few percent of each other
      std::snprintf(buf, sizeof buf, "%s\"%d/%d\": %.4f", i ? ", " : "", c->tune.col_tx[p], c->tune.col_plan[p], ms);
      rep += buf;
      if (ms < tbest) {
        tbest = ms;
        best_tx = c->tune.col_tx[p];
        best_plan = c->tune.col_plan[p];
      }
    }
    c->tune.col_tx[p] = best_tx;
    c->tune.col_plan[p] = best_plan;
    rep += "}";
    scratch_release(c, S_TUNE0);
    scratch_release(c, S_TUNE1);
    scratch_release(c, S_TUNE2);
  }
  if ((what & 2) && c->err.empty()) {
    // tridiagonal schedule, on the spectral right-hand side of this state
    T* P = (T*)scratch(c, S_TUNE0, (size_t)g.ld * g.ny * (nz + 1) * sizeof(T));
    T* dx = (T*)scratch(c, S_TUNE1, n3 * sizeof(T));
    T* dy = (T*)scratch(c, S_TUNE2, n3 * sizeof(T));
    op_press_stag<T>(c, u, F<T>(c, "v"), F<T>(c, "w"), F<T>(c, "divtz"), P, dx, dy);
    const long ps = (long)g.ny * g.nxh;
    cx<T>* H = (cx<T>*)scratch(c, S_PSPEC, (size_t)(3 * nl + 2) * ps * sizeof(cx<T>));
    cx<T>* PH = (cx<T>*)scratch(c, S_PHAT, (size_t)nz * ps * sizeof(cx<T>));
    char* tri = (char*)scratch(c, S_TRI, (size_t)(nz + 1) * ps * (sizeof(double) + sizeof(cx<double>)) + 256);
    int* err = (int*)scratch(c, S_TMP2, 256);
    const int cand[][2] = {{0, 64}, {0, 128}, {0, 256}, {1, 64}, {1, 128}, {2, 32},
                           {2, 64}, {3, 64}, {4, 32}, {4, 64}, {4, 128}};
    const Tune keep = c->tune;
    Tune best = keep;
    double tbest = 1e30;
    rep += (what & 1) ? ", " : "";
    rep += "\"tridiag\": {";
    bool first = true;
    for (auto& cd : cand) {
      if (tridiag_lds(cd[0], cd[1], nz) > 160 * 1024 || !c->err.empty()) continue;
      c->tune.tri_mode = cd[0];
      c->tune.tri_block = cd[1];
      const double ms = tune_time(c, [&] {
        launch_tridiag<T>(c, H, H + (size_t)3 * nl * ps, PH, (double*)tri,
                          (cx<double>*)(tri + (size_t)(nz + 1) * ps * sizeof(double)), err);
      }, 5);
      std::snprintf(buf, sizeof buf, "%s\"%d/%d\": %.4f", first ? "" : ", ", cd[0], cd[1], ms);
      rep += buf;
      first = false;
      if (ms < tbest) {
        tbest = ms;
        best = c->tune;
      }
    }
    c->tune.tri_mode = best.tri_mode;
    c->tune.tri_block = best.tri_block;
    rep += "}";
    scratch_release(c, S_TUNE0);
    scratch_release(c, S_TUNE1);
    scratch_release(c, S_TUNE2);
  }
  if ((what & 4) && c->err.empty()) {
    // padded products: level groups per workgroup or streaming level runs, on
    // the pass the step runs (momentum with both scalars, or momentum alone)
    const int p = sizeof(T) == 8 ? 1 : 0;
    const int scal = c->scal.enable ? 1 : 0;
    const int keep = c->tune.pad_mode[p];
    int best = keep;
    double tbest = 1e30;
    rep += (what & 3) ? ", " : "";
    rep += "\"pad\": {";
    // mode 2 differs from mode 1 only for the convective step with both scalars at nx2 = 384
    const int nmode = (scal && c->g.nx2 == 384) ? 3 : 2;
    for (int m = 0; m < nmode && c->err.empty(); ++m) {
      c->tune.pad_mode[p] = m;
      const double ms = tune_time(c, [&] { op_pad_products_probe<T>(c, scal); }, 5);
      std::snprintf(buf, sizeof buf, "%s\"%d\": %.4f", m ? ", " : "", m, ms);
      rep += buf;
      if (ms < tbest) {
        tbest = ms;
        best = m;
      }
    }
    c->tune.pad_mode[p] = best;
    rep += "}";
  }
  std::snprintf(buf, sizeof buf,
                ", \"chosen\": {\"col_tx\": [%d, %d], \"col_plan\": [%d, %d], \"tri_mode\": %d, \"tri_block\": %d, "
                "\"pad_mode\": [%d, %d]}}",
                c->tune.col_tx[0], c->tune.col_tx[1], c->tune.col_plan[0], c->tune.col_plan[1], c->tune.tri_mode,
                c->tune.tri_block, c->tune.pad_mode[0], c->tune.pad_mode[1]);
  rep += buf;
}
}  // extern "C++"

int les_autotune(les_ctx* c, int what, char* report, int report_len) {
  GUARD(c);
  if (!c->st) { c->err = "les_autotune: needs a state (les_state_alloc + initial condition)"; return 1; }
  if (c->gexec) {  // captured windows hold the old launch shapes
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  std::string rep;
  if (c->st->prec) tune_run<float>(c, what, rep);
  else tune_run<double>(c, what, rep);
  if (report && report_len > 0) {
    std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  DONE(c);
}
int les_get_tune(les_ctx* c, int* out8) {
  GUARD(c);
  if (!out8) return -1;
  out8[0] = c->tune.col_tx[0];
  out8[1] = c->tune.col_tx[1];
  out8[2] = c->tune.tri_mode;
  out8[3] = c->tune.tri_block;
  out8[4] = c->tune.col_plan[0];
  out8[5] = c->tune.col_plan[1];
  out8[6] = c->tune.pad_mode[0];
  out8[7] = c->tune.pad_mode[1];
  DONE(c);
}
int les_set_layout(les_ctx* c, int xs_tiled) {
  GUARD(c);
  if (xs_tiled < 0 || xs_tiled > 1 || (xs_tiled && c->g.nxh % 8)) {
    c->err = "les_set_layout: xs_tiled is 0 or 1, and 1 needs nx a multiple of 16";
    return 1;
  }
  c->tune.xs_tiled = xs_tiled;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  DONE(c);
}
int les_get_layout(les_ctx* c) { return c ? c->tune.xs_tiled : -1; }
int les_set_tune(les_ctx* c, const int* in8) {
  GUARD(c);
  if (!in8) return -1;
  const int* in6 = in8;
  const bool ok32 = in6[0] == 8 || in6[0] == 16 || in6[0] == 32, ok64 = in6[1] == 4 || in6[1] == 8 || in6[1] == 16;
  const bool okt = in6[2] >= 0 && in6[2] <= 4 && in6[3] >= 32 && in6[3] <= (in6[2] == 4 ? 256 : 512) &&
                   in6[3] % 32 == 0 &&
                   tridiag_lds(in6[2], in6[3], c->g.nz) <= 160 * 1024;
  const bool okp = (in6[4] == 0 || (in6[4] == 2 && (in6[0] == 8 || in6[0] == 16))) &&
                   (in6[5] == 0 || (in6[5] == 2 && (in6[1] == 4 || in6[1] == 8)));
  const bool okm = in8[6] >= 0 && in8[6] <= 2 && in8[7] >= 0 && in8[7] <= 2;
  if (!ok32 || !ok64 || !okt || !okp || !okm) { c->err = "les_set_tune: launch shape out of range"; return 1; }
  c->tune.col_tx[0] = in6[0];
  c->tune.col_tx[1] = in6[1];
  c->tune.tri_mode = in6[2];
  c->tune.tri_block = in6[3];
  c->tune.col_plan[0] = in6[4];
  c->tune.col_plan[1] = in6[5];
  c->tune.pad_mode[0] = in8[6];
  c->tune.pad_mode[1] = in8[7];
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  DONE(c);
}

// ---- learned closure ----
int les_mlp_nparam(void) { return les::mlp_nparam(); }
#define NEED_MLP(c)                                            \
  if (!(c)->mlp) {                                             \
    (c)->err = "learned closure: call les_mlp_init first";     \
    return 1;                                                  \
  }
int les_mlp_init(les_ctx* c, uint64_t seed) {
  GUARD(c);
  les::mlp_init(c, seed);
  DONE(c);
}
int les_mlp_set_params(les_ctx* c, const float* theta) {
  GUARD(c);
  NEED_MLP(c);
  if (!theta) return -1;
  les::mlp_set_params(c, theta);
  DONE(c);
}
int les_mlp_get_params(les_ctx* c, float* theta) {
  GUARD(c);
  NEED_MLP(c);
  if (!theta) return -1;
  les::mlp_get_params(c, theta);
  DONE(c);
}
int les_mlp_set_norm(les_ctx* c, const float* mu, const float* sd, const float* ymu, const float* ysd) {
  GUARD(c);
  NEED_MLP(c);
  if (!mu || !sd || !ymu || !ysd) return -1;
  for (int i = 0; i < 8; ++i)
    if (!(sd[i] > 0.0f)) { c->err = "les_mlp_set_norm: sd must be positive"; return 1; }
  for (int i = 0; i < 2; ++i)
    if (!(ysd[i] > 0.0f)) { c->err = "les_mlp_set_norm: ysd must be positive"; return 1; }
  les::mlp_set_norm(c, mu, sd, ymu, ysd);
  DONE(c);
}
int les_mlp_get_norm(les_ctx* c, float* mu, float* sd, float* ymu, float* ysd) {
  GUARD(c);
  NEED_MLP(c);
  if (!mu || !sd || !ymu || !ysd) return -1;
  les::mlp_get_norm(c, mu, sd, ymu, ysd);
  DONE(c);
}
int les_mlp_fit_norm(les_ctx* c, const float* X, const float* Y, int64_t P) {
  GUARD(c);
  NEED_MLP(c);
  if (!X || !Y || P < 2) return -1;
  les::mlp_fit_norm(c, X, Y, (long)P);
  DONE(c);
}
int les_mlp_forward(les_ctx* c, const float* X, int64_t P, float* Y) {
  GUARD(c);
  NEED_MLP(c);
  if (!X || !Y || P < 0) return -1;
  if (P > 0) les::op_mlp_forward(c, X, (long)P, Y);
  DONE(c);
}
int les_mlp_grad(les_ctx* c, const float* X, const float* T, int64_t P, float* grad, float* loss) {
  GUARD(c);
  NEED_MLP(c);
  if (!X || !T || P < 1) return -1;
  les::op_mlp_grad(c, X, T, (long)P, grad, loss);
  DONE(c);
}
int les_mlp_train(les_ctx* c, const float* X, const float* T, int64_t P, int batch, int steps, float lr,
                  uint64_t seed, float* loss_hist) {
  GUARD(c);
  NEED_MLP(c);
  if (!X || !T || P < 1 || batch < 1 || steps < 0 || !(lr > 0.0f)) return -1;
  les::op_mlp_train(c, X, T, (long)P, batch, steps, lr, seed, loss_hist);
  DONE(c);
}
extern "C++" {
template <typename T>
static void harvest_t(les_ctx* c, float* X, float* Y) {
  const Grid& g = c->g;
  const size_t n3 = (size_t)g.ld * g.ny * g.nz;
  const T* Fs = (const T*)scratch(c, S_LAGF, 42 * n3 * sizeof(T));
  const T* Sb = (const T*)scratch(c, S_SIJ, 6 * n3 * sizeof(T));
  const T* S[6] = {Sb, Sb + n3, Sb + 2 * n3, Sb + 3 * n3, Sb + 4 * n3, Sb + 5 * n3};
  const T* gr[9] = {F<T>(c, "dudx"), F<T>(c, "dudy"), F<T>(c, "dudz"), F<T>(c, "dvdx"), F<T>(c, "dvdy"),
                    F<T>(c, "dvdz"), F<T>(c, "dwdx"), F<T>(c, "dwdy"), F<T>(c, "dwdz")};
  op_mlp_harvest<T>(c, Fs, S, gr, F<T>(c, "Cs_opt2"), F<T>(c, "Beta"), X, Y);
}
}  // extern "C++"
int les_mlp_harvest(les_ctx* c, float* X, float* Y) {
  GUARD(c);
  if (!X) return -1;
  if (!c->st) { c->err = "les_mlp_harvest: no state"; return 1; }
  if (c->sgs_model != 5 || c->last_lag_jt != c->jt || c->jt < 1) {
    c->err = "les_mlp_harvest: the last step was not a Lagrangian dynamic update (lagrange_Sdep)";
    return 1;
  }
  if (c->st->prec) harvest_t<float>(c, X, Y);
  else harvest_t<double>(c, X, Y);
  DONE(c);
}
int les_set_closure(les_ctx* c, int mode, int period, double threshold) {
  GUARD(c);
  if (mode != 0 && mode != 1) { c->err = "les_set_closure: mode must be 0 (dynamic) or 1 (network)"; return 1; }
  if (mode == 1) {
    NEED_MLP(c);
    if (c->sgs_model != 5) { c->err = "les_set_closure: the network replaces the model-5 update"; return 1; }
    if (period < c->g.cs_count || period % c->g.cs_count != 0) {
      c->err = "les_set_closure: period must be a positive multiple of cs_count";
      return 1;
    }
    if (!(threshold > 0.0)) { c->err = "les_set_closure: threshold must be positive"; return 1; }
  }
  c->closure_mode = mode;
  c->closure_period = mode ? period : c->g.cs_count;
  c->closure_thresh = threshold;
  c->lag_cs = mode ? period : 0;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  DONE(c);
}
int les_closure_judge(les_ctx* c, double* out) {
  GUARD(c);
  NEED_MLP(c);
  if (!out) return -1;
  les::mlp_judge_read(c, out);
  DONE(c);
}

int les_set_graph(les_ctx* c, int enable) {
  GUARD(c);
  c->use_graph = enable != 0;
  if (!c->use_graph && c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  DONE(c);
}

int les_graph_prepare(les_ctx* c) {
  GUARD(c);
  if (!c->st) return -1;
  const int W = graph_window(c);
  int uc = 0;
  if (c->use_graph && window_is_regular(c, W, &uc) && !graph_matches(c, W, uc)) graph_capture(c, W, uc);
  DONE(c);
}

int les_step(les_ctx* c, int nsteps) {
  GUARD(c);
  if (!c->st) {
    c->err = "les_step: no state (call les_state_alloc)";
    return 1;
  }
  const int W = graph_window(c);
  int done = 0;
  while (done < nsteps && c->err.empty()) {
    int uc = 0;
    if (c->use_graph && !c->prof && nsteps - done >= W && window_is_regular(c, W, &uc)) {
      if (!graph_matches(c, W, uc)) graph_capture(c, W, uc);
      if (c->gexec && c->err.empty()) {
        check(c, hipGraphLaunch(c->gexec, c->stream), "hipGraphLaunch");
        advance_after_replay(c, W);
        done += W;
        continue;
      }
    }
    step_eager(c);
    ++done;
  }
  DONE(c);
}

int les_dev_alloc(les_ctx* c, uint64_t bytes, void** ptr) {
  GUARD(c);
  *ptr = nullptr;
  check(c, hipMalloc(ptr, bytes ? bytes : 16), "les_dev_alloc");
  if (*ptr) check(c, hipMemset(*ptr, 0, bytes ? bytes : 16), "les_dev_alloc zero");
  DONE(c);
}
int les_dev_free(les_ctx* c, void* ptr) {
  GUARD(c);
  check(c, hipStreamSynchronize(c->stream), "les_dev_free sync");
  if (ptr) check(c, hipFree(ptr), "les_dev_free");
  DONE(c);
}
int les_memcpy_h2d(les_ctx* c, void* dst, const void* src, uint64_t bytes) {
  GUARD(c);
  check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "h2d");
  check(c, hipStreamSynchronize(c->stream), "h2d sync");
  DONE(c);
}
int les_memcpy_d2h(les_ctx* c, void* dst, const void* src, uint64_t bytes) {
  GUARD(c);
  check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "d2h");
  check(c, hipStreamSynchronize(c->stream), "d2h sync");
  DONE(c);
}
void* les_mirror(les_ctx* c, const void* host, uint64_t bytes) {
  if (!c || !host) return nullptr;
  for (auto& m : c->mirrors)
    if (m.first == host) {
      if (m.second.second >= bytes) return m.second.first;
      check(c, hipStreamSynchronize(c->stream), "mirror sync");
      (void)hipFree(m.second.first);
      m.second.first = nullptr;
      check(c, hipMalloc(&m.second.first, bytes), "mirror grow");
      m.second.second = m.second.first ? bytes : 0;
      return m.second.first;
    }
  void* d = nullptr;
  check(c, hipMalloc(&d, bytes ? bytes : 16), "mirror alloc");
  if (!d) return nullptr;
  c->mirrors.push_back({host, {d, (size_t)bytes}});
  return d;
}
int les_mirror_push(les_ctx* c, const void* host, uint64_t bytes) {
  GUARD(c);
  void* d = les_mirror(c, host, bytes);
  if (d) check(c, hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, c->stream), "mirror push");
  DONE(c);
}
int les_mirror_pull(les_ctx* c, void* host, uint64_t bytes) {
  GUARD(c);
  void* d = les_mirror(c, host, bytes);
  if (d) {
    check(c, hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, c->stream), "mirror pull");
    check(c, hipStreamSynchronize(c->stream), "mirror pull sync");
  }
  DONE(c);
}

int les_profile(les_ctx* c, int enable) {
  GUARD(c);
  c->prof = enable != 0;
  DONE(c);
}

int les_profile_report(les_ctx* c, char* buf, int buflen) {
  GUARD(c);
  check(c, hipStreamSynchronize(c->stream), "profile sync");
  struct Acc {
    int n = 0;
    double ms = 0, bytes = 0;
  };
  std::map<std::string, Acc> m;
  for (auto& r : c->prof_recs) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    Acc& a = m[r.tag];
    a.n += 1;
    a.ms += ms;
    a.bytes += r.bytes;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  c->prof_recs.clear();
  std::string out = "{";
  bool first = true;
  for (auto& kv : m) {
    char tmp[256];
    snprintf(tmp, sizeof(tmp), "%s\"%s\": [%d, %.6f, %.1f]", first ? "" : ", ", kv.first.c_str(), kv.second.n,
             kv.second.ms, kv.second.bytes);
    out += tmp;
    first = false;
  }
  out += "}";
  if (buf && buflen > 0) {
    std::strncpy(buf, out.c_str(), (size_t)buflen - 1);
    buf[buflen - 1] = 0;
  }
  if ((int)out.size() >= buflen) c->err = "les_profile_report: buffer too small";
  DONE(c);
}

// ---- scalars ----
int les_set_scalars(les_ctx* c, const les_scalar_params* p, const double* wt_s, const double* wq_s) {
  GUARD(c);
  if (!p) return -1;
  if (p->enable && p->model == 7 && c->sgs_model != 5) {
    c->err = "les_set_scalars: the scalar Lagrangian model (model 7) runs with lagrange_Sdep; select SGS model 5";
    return 1;
  }
  if (p->enable && (p->model == 6) != (c->sgs_model == 6)) {
    c->err = "les_set_scalars: scalar model 6 (Pr_t of scalar_scaledep_dynamic) goes with SGS model 6 and only "
             "with it (sgs_stag.f90:143-149, scalars_module.f90:77-84)";
    return 1;
  }
  ScalarParams& P = c->scal;
  P.enable = p->enable; P.model = p->model; P.initsc = p->initsc; P.passive = p->passive;
  P.Pr = p->Pr; P.z_i = p->z_i; P.u_star = p->u_star; P.T_scale = p->T_scale; P.q_scale = p->q_scale;
  P.g = p->g; P.dTdz_top = p->dTdz_top; P.dqdz_top = p->dqdz_top;
  const Grid& g = c->g;
  const size_t n2 = (size_t)g.nx * g.ny;
  std::vector<double> wt(n2, 0.0), wq(n2, 0.0);
  if (wt_s) std::memcpy(wt.data(), wt_s, n2 * sizeof(double));
  if (wq_s) std::memcpy(wq.data(), wq_s, n2 * sizeof(double));
  for (double** d : {&c->wt_s, &c->wq_s})
    if (!*d) check(c, hipMalloc(d, n2 * sizeof(double)), "surface flux alloc");
  check(c, hipMemcpy(c->wt_s, wt.data(), n2 * sizeof(double), hipMemcpyHostToDevice), "wt_s upload");
  check(c, hipMemcpy(c->wq_s, wq.data(), n2 * sizeof(double), hipMemcpyHostToDevice), "wq_s upload");
  // plane constants of lagrange_scalar_Sdep.f90:110,114 in the reference's summation order
  double slog = 0.0, swt = 0.0;
  for (size_t j = 0; j < n2; ++j) slog += log(g.z0);
  for (size_t j = 0; j < n2; ++j) swt += wt[j];
  c->zo_avg = exp(slog / (double)(float)n2);
  c->wt_mean = swt / (double)n2;
  DONE(c);
}

int les_scalar_ref(les_ctx* c, double* theta_ref, double* q_ref) {
  GUARD(c);
  if (theta_ref) *theta_ref = c->scal.theta_ref;
  if (q_ref) *q_ref = c->scal.q_ref;
  DONE(c);
}

int les_set_scalar_lag_init(les_ctx* c, int glm_init, int gqn_init) {
  GUARD(c);
  c->glm_init = glm_init;
  c->gqn_init = gqn_init;
  DONE(c);
}

static les::ScalarSgsPtrs to_sptrs(const les_scalar_sgs_fields* s) {
  les::ScalarSgsPtrs p;
  p.G_LM = s->G_LM; p.G_MM = s->G_MM; p.G_QN = s->G_QN; p.G_NN = s->G_NN;
  p.u_lag_s = s->u_lag_s; p.v_lag_s = s->v_lag_s; p.w_lag_s = s->w_lag_s;
  p.Pr_t = s->Pr_t; p.Cs2_Pr = s->Cs2_Pr; p.phi_h = s->phi_h; p.psi_m = s->psi_m;
  return p;
}

int les_obukhov(les_ctx* c, const double* u, const double* v, const double* theta, const double* qmix,
                double* sgs_t3, double* sgs_q3, double* psi_m, double* psi_m0, double* psi_h, double* psi_h0,
                double* phi_m, double* phi_h, double* ustar_avg) {
  GUARD(c);
  (void)qmix;  // read by the reference only for its printed diagnostics
  if (!c->wt_s) { c->err = "les_obukhov: call les_set_scalars first"; return 1; }
  op_obukhov<double>(c, u, v, theta, sgs_t3, sgs_q3, psi_m, psi_m0, psi_h, psi_h0, phi_m, phi_h, ustar_avg);
  DONE(c);
}

int les_calcbeta(les_ctx* c, const double* theta, const double* qmix, double* beta_scal) {
  GUARD(c);
  op_calcbeta<double>(c, theta, qmix, beta_scal);
  DONE(c);
}

int les_scalar_rhs_calc(les_ctx* c, int which, const double* dsdx, const double* dsdy, double* dsdz,
                        const double* u, const double* v, const double* w, const double* Nu_t, const double* Pr_,
                        const double* ustar_avg, const double* phi_h, double* RHS, double* sgs_vert1) {
  GUARD(c);
  if (!c->wt_s) { c->err = "les_scalar_rhs_calc: call les_set_scalars first"; return 1; }
  if (which != 0 && which != 1) { c->err = "les_scalar_rhs_calc: which must be 0 (theta) or 1 (q)"; return 1; }
  ScalarJob j{nullptr, (void*)dsdx, (void*)dsdy, dsdz, RHS, nullptr, sgs_vert1, which};
  op_scalar_dz<double>(c, &j, 1, phi_h, ustar_avg, 1);
  op_scalar_rhs<double>(c, &j, 1, u, v, w, Nu_t, Pr_, 0, 0);
  DONE(c);
}

int les_step_scalar(les_ctx* c, int which, double* scalar, const double* RHS, const double* RHS_f) {
  GUARD(c);
  if (which != 0 && which != 1) { c->err = "les_step_scalar: which must be 0 (theta) or 1 (q)"; return 1; }
  op_step_scalar<double>(c, which, scalar, RHS, RHS_f);
  DONE(c);
}

int les_interpolag_scalar_sdep(les_ctx* c, const les_scalar_sgs_fields* s) {
  GUARD(c);
  if (!s) return -1;
  double* G[4] = {(double*)s->G_LM, (double*)s->G_MM, (double*)s->G_QN, (double*)s->G_NN};
  les::OpTag _t(c, "interpolag_scalar");
  op_interpolag4<double>(c, G, (double*)s->u_lag_s, (double*)s->v_lag_s, (double*)s->w_lag_s);
  DONE(c);
}

int les_lagrange_scalar_sdep(les_ctx* c, double* const S[6], const double* u, const double* v, const double* w,
                             double* theta, const les_sgs_fields* m, const les_scalar_sgs_fields* s) {
  GUARD(c);
  if (!m || !s) return -1;
  if (!c->wt_s) { c->err = "les_lagrange_scalar_sdep: call les_set_scalars first"; return 1; }
  const double* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
  op_lagrange_scalar<double>(c, Sc, u, v, w, theta, to_ptrs(m), to_sptrs(s), false);
  DONE(c);
}

int les_sgs_stag_scalar(les_ctx* c, double* const grad[9], const double* u, const double* v, const double* w,
                        double* theta, double* const tau[6], const les_sgs_fields* m,
                        const les_scalar_sgs_fields* s) {
  GUARD(c);
  if (!m || !s) return -1;
  if (!c->wt_s) { c->err = "les_sgs_stag_scalar: call les_set_scalars first"; return 1; }
  if (c->sgs_model != 5 && c->sgs_model != 6) {
    c->err = "les_sgs_stag_scalar: the scalar SGS models need SGS model 5 (scalar model 7) or 6";
    return 1;
  }
  const double* g[9];
  for (int i = 0; i < 9; ++i) g[i] = grad[i];
  const les::ScalarSgsPtrs ss = to_sptrs(s);
  op_sgs_stag<double>(c, g, u, v, w, tau, to_ptrs(m), theta, &ss);
  DONE(c);
}

int les_scalar_scaledep_dynamic(les_ctx* c, double* const S[6], const double* u, const double* v, const double* w,
                                double* theta, const double* Cs_opt2, const les_scalar_sgs_fields* s) {
  GUARD(c);
  if (!S || !s || !theta || !Cs_opt2) return -1;
  if (!c->wt_s) { c->err = "les_scalar_scaledep_dynamic: call les_set_scalars first"; return 1; }
  const double* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
  const les::ScalarSgsPtrs ss = to_sptrs(s);
  op_scalar_scaledep<double>(c, Sc, u, v, w, theta, Cs_opt2, ss, false);
  DONE(c);
}

int les_state_init_scalar_ic(les_ctx* c, uint64_t seed, double T_init, double z_inv, double inv_strength,
                             double amp_K, double q_init) {
  GUARD(c);
  if (!c->st || !c->st->scal) { c->err = "les_state_init_scalar_ic: state has no scalars"; return 1; }
  State* s = c->st;
  const Grid& g = c->g;
  const ScalarParams& P = c->scal;
  const size_t pl = (size_t)g.ld * g.ny, n3 = pl * g.nz, n2 = (size_t)g.nx * g.ny;
  // FP32-bulk stores deviations from the initial reference values
  c->scal.theta_ref = s->prec ? T_init / P.T_scale : 0.0;
  c->scal.q_ref = s->prec ? q_init / P.q_scale : 0.0;
  std::vector<double> th(n3, 0.0), q(n3, 0.0), pr(n3, P.Pr);
  for (int k = 0; k < g.nz - 1; ++k) {
    const double z = (k + 0.5) * g.dz;
    const double tk = T_init + (z > z_inv ? inv_strength * (z - z_inv) * P.z_i : 0.0);
    for (int y = 0; y < g.ny; ++y)
      for (int x = 0; x < g.nx; ++x) {
        const size_t i = pl * k + x + (size_t)g.ld * y;
        double t = tk;
        if (z < z_inv) {
          const uint64_t idx = (((uint64_t)3 * g.nz + k) * g.ny + y) * g.nx + x;
          const double U = (double)(splitmix(seed, idx) >> 11) * (1.0 / 9007199254740992.0);
          t += amp_K * (2.0 * U - 1.0);
        }
        th[i] = t / P.T_scale;
        q[i] = q_init / P.q_scale;
      }
  }
  for (size_t i = 0; i < n3; ++i)
    if ((int)(i % g.ld) < g.nx && i < pl * (g.nz - 1)) {
      th[i] -= c->scal.theta_ref;
      q[i] -= c->scal.q_ref;
    }
  const size_t top = pl * (g.nz - 1);
  const double gt = g.Lz > 1.0 ? P.dTdz_top / P.T_scale * P.z_i * g.dz : 0.0;
  const double gq = g.Lz > 1.0 ? P.dqdz_top / P.q_scale * P.z_i * g.dz : 0.0;
  for (size_t i = 0; i < pl; ++i) {
    if ((int)(i % g.ld) >= g.nx) continue;
    th[top + i] = th[top - pl + i] + gt;
    q[top + i] = q[top - pl + i] + gq;
  }
  auto up = [&](const char* name, const std::vector<double>& h) {
    void* d = state_ptr(c, name);
    if (s->prec) {
      std::vector<float> f(n3);
      for (size_t i = 0; i < n3; ++i) f[i] = (float)h[i];
      check(c, hipMemcpy(d, f.data(), n3 * 4, hipMemcpyHostToDevice), "scalar ic upload");
    } else {
      check(c, hipMemcpy(d, h.data(), n3 * 8, hipMemcpyHostToDevice), "scalar ic upload");
    }
  };
  check(c, hipStreamSynchronize(c->stream), "scalar ic sync");
  up("theta", th);
  up("qmix", q);
  up("Pr_t", pr);
  std::vector<double> one(n2, 1.0);
  check(c, hipMemcpy(state_ptr(c, "phi_h"), one.data(), n2 * 8, hipMemcpyHostToDevice), "phi_h");
  check(c, hipMemset(state_ptr(c, "psi_h"), 0, n2 * 8), "psi_h");
  check(c, hipMemset(state_ptr(c, "psi_h0"), 0, n2 * 8), "psi_h0");
  c->glm_init = 0;
  c->gqn_init = 0;
  DONE(c);
}

// ---- statistics: spectrum (io.f90:1613-1646), the current statistics of
// avg_stats (io.f90:1722-1755) and obukhov's plane averages
// (scalars_module.f90:545-561,655-661), FP64 reductions on the device ----
int les_spectrum(les_ctx* c, const double* f, int nlev, double* spec) {
  GUARD(c);
  if (nlev < 1 || nlev > c->g.nz + 1) {
    c->err = "spectrum: nlev must be 1 .. nz+1";
    return 2;
  }
  op_spectrum<double>(c, f, nlev, spec, 0);
  DONE(c);
}
int les_avg_stats(les_ctx* c, const double* u, const double* v, const double* w, const double* txz, double* stats) {
  GUARD(c);
  op_avg_stats<double>(c, u, v, w, txz, stats);
  DONE(c);
}
int les_surface_stats(les_ctx* c, const double* u, const double* v, const double* txz, const double* tyz,
                      const double* psi_m, const double* psi_m0, const double* ustar_avg, const double* theta,
                      const double* qmix, const double* sgs_t3, const double* sgs_q3, double* out8) {
  GUARD(c);
  op_surface_stats<double>(c, u, v, txz, tyz, psi_m, psi_m0, ustar_avg, theta, qmix, sgs_t3, sgs_q3, 0.0, 0.0,
                           out8);
  DONE(c);
}

}  // extern "C"

// statistics of the state into the device buffer d (stats_len + 8 doubles)
template <typename T>
static void state_stats_dev(les_ctx* c, double* d) {
  op_avg_stats<T>(c, F<T>(c, "u"), F<T>(c, "v"), F<T>(c, "w"), F<T>(c, "txz"), d);
  const bool sc = c->st->scal != 0;
  op_surface_stats<T>(c, F<T>(c, "u"), F<T>(c, "v"), F<T>(c, "txz"), F<T>(c, "tyz"),
                      (const double*)state_ptr(c, "psi_m"), (const double*)state_ptr(c, "psi_m0"),
                      (const double*)state_ptr(c, "ustar_avg"), sc ? F<T>(c, "theta") : nullptr,
                      sc ? F<T>(c, "qmix") : nullptr, sc ? (const double*)state_ptr(c, "sgs_t3") : nullptr,
                      sc ? (const double*)state_ptr(c, "sgs_q3") : nullptr, c->scal.theta_ref, c->scal.q_ref,
                      d + stats_len(c));
}
static double* state_stats_now(les_ctx* c) {
  const size_t len = stats_len(c) + 8;
  double* d = (double*)scratch(c, S_STAT, len * sizeof(double));
  if (c->st->prec)
    state_stats_dev<float>(c, d);
  else
    state_stats_dev<double>(c, d);
  return d;
}
static void stats_to_host(les_ctx* c, const double* d, double* stats, double* surface) {
  const size_t len = stats_len(c);
  if (stats) check(c, hipMemcpyAsync(stats, d, len * sizeof(double), hipMemcpyDeviceToHost, c->stream), "stats d2h");
  if (surface)
    check(c, hipMemcpyAsync(surface, d + len, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "stats d2h");
  check(c, hipStreamSynchronize(c->stream), "stats sync");
}

extern "C" {

int les_state_stats(les_ctx* c, double* stats, double* surface) {
  GUARD(c);
  if (!c->st) return -1;
  const double* d = state_stats_now(c);
  stats_to_host(c, d, stats, surface);
  DONE(c);
}
int les_state_stats_accumulate(les_ctx* c) {
  GUARD(c);
  if (!c->st) return -1;
  const size_t len = stats_len(c) + 8;
  const double* d = state_stats_now(c);
  double* avg = (double*)scratch(c, S_STATAVG, len * sizeof(double));
  if (c->stats_n == 0) op_fill<double>(c, avg, len, 0.0);
  op_running_mean(c, avg, d, len, c->stats_n);
  c->stats_n += 1;
  DONE(c);
}
int les_state_stats_average(les_ctx* c, double* stats, double* surface, int* n, int reset) {
  GUARD(c);
  if (!c->st) return -1;
  if (n) *n = (int)c->stats_n;
  if (c->stats_n > 0) {
    const double* avg = (const double*)scratch(c, S_STATAVG, (stats_len(c) + 8) * sizeof(double));
    stats_to_host(c, avg, stats, surface);
  }
  if (reset) c->stats_n = 0;
  DONE(c);
}

int les_state_diag(les_ctx* c, double* rms, double out4[4]) {
  GUARD(c);
  if (!c->st) return -1;
  if (c->st->prec) {
    if (rms) *rms = op_rmsdiv<float>(c, F<float>(c, "dudx"), F<float>(c, "dvdy"), F<float>(c, "dwdz"));
    if (out4) op_timestep_conditions<float>(c, F<float>(c, "u"), F<float>(c, "v"), F<float>(c, "w"),
                                            F<float>(c, "Nu_t"), out4);
  } else {
    if (rms) *rms = op_rmsdiv<double>(c, F<double>(c, "dudx"), F<double>(c, "dvdy"), F<double>(c, "dwdz"));
    if (out4) op_timestep_conditions<double>(c, F<double>(c, "u"), F<double>(c, "v"), F<double>(c, "w"),
                                             F<double>(c, "Nu_t"), out4);
  }
  DONE(c);
}

}  // extern "C"
