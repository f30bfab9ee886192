// les_scalars.hip -- scalar transport of the convective configuration on
// gfx950: potential temperature theta and humidity q (S_FLAG), the
// Monin-Obukhov surface layer, buoyancy and the scalar Lagrangian
// scale-dependent SGS model (model 7).
//
// In the whole step both scalars share the momentum's padded-grid pass:
//   one forward row transform of u, v, w, the vorticity and the three
//   gradients of each scalar, the 3/2 zero pad in y, and k_pad_products
//   (PolConvecScalar), which forms convec's products and scalar_RHS_calc's
//   staggered advection products in LDS (the padded fields never touch HBM).
//   The SGS flux divergence d/dx, d/dy of the scalars is a single spectral
//   divergence pass whose c2r also receives the advection spectrum (weight
//   -1); its epilogue adds the d/dz flux term and takes the AB2 step of
//   step_scalar, including the lid rule.
#include "les_xform.h"

namespace les {

// ------------------------------------------------------------------------
// deterministic, compensated plane sums in double: blockIdx.y = level, NPART
// blocks per level write partial sums, a second kernel adds them in a fixed
// order.  Every addition is an error-free TwoSum whose rounding error is
// carried in a second word, so a plane sum is the exact sum rounded once
// (up to a term of order n eps^2) whatever the order of the additions: plane
// means of nearly constant fields (theta ~ 283 K, whose deviations calcbeta
// scales by g z_i / u*^2 ~ 1e4) do not depend on the summation order.
// ------------------------------------------------------------------------
constexpr int NPART = 32;

template <class Fn>
__global__ void k_level_partial(Fn fn, int nx, int ny, double* part) {
  const int k = blockIdx.y, n = nx * ny;
  CSum a{0.0, 0.0};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / nx, x = j - y * nx;
    a = csum_add(a, fn(k, y, x));
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = a.s;
  sh[1][threadIdx.x] = a.c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const CSum m = csum_merge(CSum{sh[0][threadIdx.x], sh[1][threadIdx.x]},
                                CSum{sh[0][threadIdx.x + s], sh[1][threadIdx.x + s]});
      sh[0][threadIdx.x] = m.s;
      sh[1][threadIdx.x] = m.c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * (k * gridDim.x + blockIdx.x)] = sh[0][0];
    part[2 * (k * gridDim.x + blockIdx.x) + 1] = sh[1][0];
  }
}
__global__ void k_level_final(const double* part, int nparts, int nlev, double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nlev) return;
  CSum a{0.0, 0.0};
  for (int p = 0; p < nparts; ++p) a = csum_merge(a, CSum{part[2 * (k * nparts + p)], part[2 * (k * nparts + p) + 1]});
  out[k] = a.s + a.c;
}
template <class Fn>
static void level_sums(les_ctx* c, const Fn& fn, int nlev, double* out, double bytes) {
  const Grid& g = c->g;
  double* part = (double*)scratch(c, S_RED, (size_t)2 * NPART * nlev * sizeof(double));
  {
    ProfScope _ps(c, "level_sum", bytes);
    hipLaunchKernelGGL((k_level_partial<Fn>), dim3(NPART, nlev), dim3(256), 0, c->stream, fn, g.nx, g.ny, part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_level_final, dim3((nlev + 63) / 64), dim3(64), 0, c->stream, part, NPART, nlev, out);
  LCHK(c, hipGetLastError());
}

template <typename T>
struct SumPlane {  // f(x, y, level 0) + ref
  const T* f;
  int ld;
  double ref;
  __device__ double operator()(int, int y, int x) const { return (double)f[(long)y * ld + x] + ref; }
};
template <typename T>
struct SumThetaV {  // theta (1 + 0.61e-3 q q_scale) per level, scalars_module.f90:199-206
  const T *th, *q;
  int ld, ny;
  double qs, tref, qref;
  __device__ double operator()(int k, int y, int x) const {
    // no fused multiply-add: calcbeta scales theta_v - mean(theta_v) by
    // g z_i / u*^2 ~ 1e4, so one rounding of theta_v shows at 1e-12 of beta
#pragma clang fp contract(off)
    const long i = ((long)k * ny + y) * ld + x;
    return ((double)th[i] + tref) * (1.0 + 0.61 * 0.001 * ((double)q[i] + qref) * qs);
  }
};

// lagrange_scalar_Sdep.f90:111: |u(1)| summed over the plane
template <typename T>
struct SumSpeed {
  const T *u, *v;
  int ld;
  __device__ double operator()(int, int y, int x) const {
    const double uu = (double)u[(long)y * ld + x], vv = (double)v[(long)y * ld + x];
    return sqrt(uu * uu + vv * vv);
  }
};

// slots of the S_RED2 scratch (doubles)
constexpr int R2_TSUM = 0, R2_LAG = 2, R2_BETA = 8;

// ------------------------------------------------------------------------
// obukhov, scalars_module.f90:508-690 (patch_flag = 1), FP64 per point
// ------------------------------------------------------------------------
static __device__ double mo_psi_m_unstable(double x) {  // :621-622
  return 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + 3.1415926535897932384626433 / 2.0;
}

template <typename T>
__global__ void k_obukhov(const T* __restrict__ u1f, const T* __restrict__ v1f, const T* __restrict__ th,
                          double* t3, double* q3, const double* __restrict__ wt, const double* __restrict__ wq,
                          double* psi_m, double* psi_m0, double* psi_h, double* psi_h0, double* phi_m, double* phi_h,
                          double* ustar, const double* __restrict__ tsum, int nx, int ny, int ld, double dz, double z0,
                          double g_, double u_star, double T_scale, double q_scale, int first, int initsc,
                          double th_ref) {
  const int n = nx * ny;
  const double theta_ = tsum[0] / (double)(float)(nx * ny);  // :556
  const double z0s = z0 / 7.39;                               // bottombc.f90:55 (z_os = zo/7.39)
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / nx, x = j - y * nx;
    const long i = x + (long)ld * y;
    if (first) {  // :521-538, jt == SCAL_init
      psi_m[j] = 0.0; psi_m0[j] = 0.0; psi_h[j] = 0.0; psi_h0[j] = 0.0;
      if (!initsc) {
        t3[i] = wt[j] / (u_star * T_scale);
        q3[i] = wq[j] / (u_star * q_scale);
      }
    }
    const double u1 = (double)u1f[i], v1 = (double)v1f[i];
    const double ua = sqrt(u1 * u1 + v1 * v1);
    const double us = ua * VONK / (log(0.5 * dz / z0) - psi_m[j] + psi_m0[j]);  // :565-573
    ustar[j] = us;
    const double wtv = t3[i] + 0.61 * ((double)th[i] + th_ref) * (0.001 * q_scale * q3[i]);  // :546-547
    const double L = -(us * us * us) / (VONK * (g_ / theta_) * wtv);              // :579
    double zeta = 0.5 * dz / L, zeta0 = z0 / L, zeta0s = z0s / L;
    if (zeta < -30.0) zeta = -30.0;
    if (zeta > 2.0) zeta = 2.0;
    if (zeta0 < -30.0) zeta0 = -30.0;
    if (zeta0 > 2.0) zeta0 = 2.0;
    if (zeta0s < -30.0) zeta0s = -30.0;
    if (zeta0s > 2.0) zeta0s = 2.0;
    const double wt1 = t3[i];
    if (L < 0.0 && wt1 != 0.0) {  // :613-634
      const double xx = pow(1.0 - 16.0 * zeta, 0.25);
      psi_m[j] = mo_psi_m_unstable(xx);
      psi_h[j] = 2.0 * log((1.0 + xx * xx) / 2.0);
      phi_m[j] = 1.0 / xx;
      phi_h[j] = 1.0 / (xx * xx);
      const double x0 = pow(1.0 - 16.0 * zeta0, 0.25), x0s = pow(1.0 - 16.0 * zeta0s, 0.25);
      psi_m0[j] = mo_psi_m_unstable(x0);
      psi_h0[j] = 2.0 * log((1.0 + x0s * x0s) / 2.0);
    } else if (L > 0.0 && wt1 != 0.0) {  // :636-646
      phi_m[j] = 1.0 + 5.0 * zeta;
      phi_h[j] = 1.0 + 5.0 * zeta;
      psi_m[j] = -1.0 * 5.0 * zeta;
      psi_h[j] = -1.0 * 5.0 * zeta;
      psi_m0[j] = -1.0 * 5.0 * zeta0;
      psi_h0[j] = -1.0 * 5.0 * zeta0s;
    } else {  // :648-656
      psi_m[j] = 0.0; psi_m0[j] = 0.0; psi_h[j] = 0.0; psi_h0[j] = 0.0;
      phi_m[j] = 1.0; phi_h[j] = 1.0;
    }
  }
}

template <typename T>
void op_obukhov(les_ctx* c, const T* u, const T* v, const T* theta, double* sgs_t3, double* sgs_q3, double* psi_m,
                double* psi_m0, double* psi_h, double* psi_h0, double* phi_m, double* phi_h, double* ustar_avg,
                double th_ref) {
  OpTag _tag(c, "obukhov");
  const Grid& g = c->g;
  const ScalarParams& P = c->scal;
  const size_t pl = (size_t)g.ld * g.ny;
  T* planes = (T*)scratch(c, S_PLANE, 2 * pl * sizeof(T));
  op_test_filter_level0<T>(c, u, v, planes, tables<T>(c).G4, c->keep4);  // :560-563
  double* red = (double*)scratch(c, S_RED2, (R2_BETA + g.nz) * sizeof(double));
  level_sums(c, SumPlane<T>{theta, g.ld, th_ref}, 1, red + R2_TSUM, (double)g.nx * g.ny * sizeof(T));
  {
    ProfScope _ps(c, "obukhov", 12.0 * g.nx * g.ny * 8);
    hipLaunchKernelGGL((k_obukhov<T>), grid1((size_t)g.nx * g.ny), dim3(256), 0, c->stream, planes, planes + pl, theta,
                       sgs_t3, sgs_q3, c->wt_s, c->wq_s, psi_m, psi_m0, psi_h, psi_h0, phi_m, phi_h, ustar_avg,
                       red + R2_TSUM, g.nx, g.ny, g.ld, g.dz, g.z0, (P.g * P.z_i) / (P.u_star * P.u_star), P.u_star,
                       P.T_scale, P.q_scale, c->jt == 1 ? 1 : 0, P.initsc, th_ref);
  }
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// calcbeta, scalars_module.f90:175-228
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_beta(const T* __restrict__ th, const T* __restrict__ q, const double* __restrict__ sums, T* beta,
                       int nx, int ld, int ny, int nz, double g_hat, double qs, double tref, double qref) {
#pragma clang fp contract(off)  // the rounding of theta_v as in SumThetaV
  const long pl = (long)ld * ny, n = pl * nz;
  const double nxy = (double)(nx * ny);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    double b = 0.0;
    if (k >= 1 && k < nz - 1 && x < nx) {
      const double bk = sums[k] / nxy, bm = sums[k - 1] / nxy;
      const double tv = ((double)th[i] + tref) * (1.0 + 0.61 * 0.001 * ((double)q[i] + qref) * qs);
      const double tvm = ((double)th[i - pl] + tref) * (1.0 + 0.61 * 0.001 * ((double)q[i - pl] + qref) * qs);
      const double above = (tv - bk) / bk;
      const double below = (tvm - bm) / bm;
      b = g_hat * (above + below) / 2.0;
    }
    beta[i] = (T)b;
  }
}

template <typename T>
void op_calcbeta(les_ctx* c, const T* theta, const T* qmix, T* beta, double th_ref, double q_ref) {
  OpTag _tag(c, "calcbeta");
  const Grid& g = c->g;
  const ScalarParams& P = c->scal;
  double* red = (double*)scratch(c, S_RED2, (R2_BETA + g.nz) * sizeof(double));
  level_sums(c, SumThetaV<T>{theta, qmix, g.ld, g.ny, P.q_scale, th_ref, q_ref}, g.nz, red + R2_BETA,
             2.0 * g.nx * g.ny * g.nz * sizeof(T));
  const size_t n3 = (size_t)g.ld * g.ny * g.nz;
  {
    ProfScope _ps(c, "beta", 3.0 * n3 * sizeof(T));
    hipLaunchKernelGGL((k_beta<T>), grid1(n3), dim3(256), 0, c->stream, theta, qmix, red + R2_BETA, beta, g.nx, g.ld,
                       g.ny, g.nz, P.g * (P.z_i / (P.u_star * P.u_star)), P.q_scale, th_ref, q_ref);
  }
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// vertical scalar gradient: ddz_uv (levels 2..nz-1), the lid value
// (scalars_module.f90:90-92) and the surface gradient of scalar_RHS_calc
// (:278-283) or of lagrange_scalar_Sdep (:133-136)
// ------------------------------------------------------------------------
struct SurfaceFlux {  // lbc = 1 prescribed flux of one scalar
  const double* flux;
  int which;  // 0: wt_s/(T_scale u_star), 1: wq_s/q_scale/u_star
  double T_scale, q_scale, u_star;
  __device__ double operator()(int j) const {
    return which == 0 ? flux[j] / (T_scale * u_star) : flux[j] / q_scale / u_star;
  }
};

template <typename T>
struct DzArgs {
  const T* s[2];
  T* dz[2];
  SurfaceFlux sf[2];
  T top[2];
  int ns, surface_only, lag;
  const double *phi_h, *ustar;
  const double* lagsums;  // lag: {sum |u(1)|, sum psi_m, sum phi_h (lag 2)}
  double zlog, sflux;     // lag: ln(dz/2/zo_avg), mean surface flux
  // lag 1: lagrange_scalar_Sdep's first-level dTdz (local phi_h); lag 2:
  // scalar_scaledep_dynamic's (plane-mean phi_h, opposite sign)
};

// TPT consecutive points per thread, both scalars per point; surface_only
// launches cover the first level only
template <typename T, int TPT>
__global__ void k_scalar_dz(DzArgs<T> a, int nx, int ld, int ny, int nz, double dz, T cst, unsigned n) {
  const unsigned pl = (unsigned)ld * ny;
  const double fn = (double)(float)(nx * ny);
  const unsigned i0 = blockIdx.x * blockDim.x * TPT + threadIdx.x;
#pragma unroll
  for (int t = 0; t < TPT; ++t) {
    const unsigned i = i0 + t * blockDim.x;
    if (i >= n) break;
    const int k = (int)(i / pl);
    const int r = (int)(i - k * pl), y = r / ld, x = r - y * ld;
    for (int q = 0; q < a.ns; ++q) {
      T* d = a.dz[q];
      if (k == nz - 1) {
        d[i] = a.top[q];
        continue;
      }
      if (x >= nx) continue;
      if (k == 0) {
        const int j = x + nx * y;
        double v;
        if (a.lag == 1) {
          const double ustar_ = a.lagsums[0] / fn * VONK / (a.zlog - a.lagsums[1] / fn);
          v = -a.phi_h[j] * a.sflux / (ustar_ * VONK * dz / 2.0);
        } else if (a.lag == 2) {  // scalar_scaledep_dynamic.f90:75-79: plane-mean phi_h, no sign change
          const double ustar_ = a.lagsums[0] / fn * VONK / (a.zlog - a.lagsums[1] / fn);
          const double phi_h_avg = a.lagsums[2] / fn;
          v = phi_h_avg * a.sflux / (ustar_ * VONK * dz / 2.0);
        } else {
          v = -a.phi_h[j] * a.sf[q](j) / (a.ustar[j] * VONK * dz / 2.0);
        }
        d[i] = (T)v;
      } else {
        const T* s = a.s[q];
        d[i] = cst * (s[i] - s[i - pl]);
      }
    }
  }
}

static SurfaceFlux surface_flux(const les_ctx* c, int which) {
  const ScalarParams& P = c->scal;
  return SurfaceFlux{which == 0 ? c->wt_s : c->wq_s, which, P.T_scale, P.q_scale, P.u_star};
}
static double lid_gradient(const les_ctx* c, int which) {  // scalars_module.f90:90-92, _q
  const ScalarParams& P = c->scal;
  return which == 0 ? P.dTdz_top / P.T_scale * P.z_i : P.dqdz_top / P.q_scale * P.z_i;
}

template <typename T>
void op_scalar_dz(les_ctx* c, const ScalarJob* jobs, int ns, const double* phi_h, const double* ustar_avg,
                  int surface_only) {
  const Grid& g = c->g;
  DzArgs<T> a{};
  a.ns = ns;
  a.surface_only = surface_only;
  for (int q = 0; q < ns; ++q) {
    a.s[q] = (const T*)jobs[q].s;
    a.dz[q] = (T*)jobs[q].dsdz;
    a.sf[q] = surface_flux(c, jobs[q].which);
    a.top[q] = (T)lid_gradient(c, jobs[q].which);
  }
  a.phi_h = phi_h;
  a.ustar = ustar_avg;
  const size_t n3 = (size_t)g.ld * g.ny * g.nz, n = surface_only ? (size_t)g.ld * g.ny : n3;
  constexpr int TPT = 4;
  {
    ProfScope _ps(c, "scalar_dz", (surface_only ? 0.02 : 2.0) * ns * n3 * sizeof(T));
    hipLaunchKernelGGL((k_scalar_dz<T, TPT>), dim3((unsigned)((n + 256 * TPT - 1) / (256 * TPT))), dim3(256), 0,
                       c->stream, a, g.nx, g.ld, g.ny, g.nz, g.dz, T(1.0 / g.dz), (unsigned)n);
  }
  LCHK(c, hipGetLastError());
}

// SGS fluxes of scalars_module.f90:366-432 (times the ddx/ddy normalisation):
// field f = dir * NS + s, dir 0 -> ds/dx, 1 -> ds/dy; levels 1..nz-1
template <typename T>
struct LoadScalarTemp {
  const T* ds[2][2];  // [dir][scalar]
  const T *Nu, *Pr;
  T Prc;  // constant Pr_ when Pr == nullptr
  int ns, ppf, ny, ld;
  T cst;
  static constexpr double reads() { return 4.0; }
  __device__ T operator()(int plane, int y, int x) const {
    const int f = plane / ppf, k = plane - f * ppf;
    const int dir = f / ns, s = f - dir * ns;
    const long pl = (long)ld * ny, i = ((long)k * ny + y) * ld + x;
    const T p0 = Pr ? Pr[i] : Prc;
    const T d = ds[dir][s][i];
    T t;
    if (k == 0) {
      t = (T(1) / p0) * Nu[i] * d;
    } else {
      const T p1 = Pr ? Pr[i + pl] : Prc;
      t = (T(1) / (T(0.5) * (p0 + p1))) * T(0.5) * (Nu[i] + Nu[i + pl]) * d;
    }
    return cst * t;
  }
};

// the same 2 NS fields formed at one point (k_row_fwd_mf): Nu_t and Pr_t are
// read once for all of them; field f = dir NS + s, the same values as
// LoadScalarTemp
template <typename T, int NS>
struct LoadScalarTempAll {
  static constexpr int NF = 2 * NS;
  const T* ds[2][2];  // [dir][scalar]
  const T *Nu, *Pr;
  T Prc;
  int ny, ld;
  T cst;
  __device__ void all(int k, int y, int x, T* v) const {
    const long pl = (long)ld * ny, i = ((long)k * ny + y) * ld + x;
    const T p0 = Pr ? Pr[i] : Prc;
    T a;
    if (k == 0) {
      a = (T(1) / p0) * Nu[i];
    } else {
      const T p1 = Pr ? Pr[i + pl] : Prc;
      a = (T(1) / (T(0.5) * (p0 + p1))) * T(0.5) * (Nu[i] + Nu[i + pl]);
    }
#pragma unroll
    for (int dir = 0; dir < 2; ++dir)
#pragma unroll
      for (int s = 0; s < NS; ++s) v[dir * NS + s] = cst * (a * ds[dir][s][i]);
  }
};

// the sponge term of the scalar equations (scalars_module.f90:143-148,
// scalars_module_q.f90:116-121): field 3 = theta, 4 = q of the plane sums
template <typename T>
__device__ __forceinline__ T scalar_sponge(const ForcingArgs<T>& fo, T rhs, int field, int r, int nl, T s0) {
#pragma clang fp contract(off)
  const T m = (T)(fo.sums[field * nl + r] / fo.nxy);
  return rhs - (T)fo.coef[r] * (s0 - m);
}

// c2r epilogue: RHS = -advection + d/dx + d/dy (spectral) + d/dz of the
// vertical SGS flux (scalars_module.f90:436-452), the sponge term, then
// step_scalar (:457-504)
template <typename T>
struct StoreScalarRHS {
  T* s[2];
  T* R[2];
  T* Rf[2];
  const T* dsdz[2];
  double* sv1[2];
  SurfaceFlux sf[2];
  T gz[2];  // lid increment (dTdz_top/T_scale z_i dz when the domain is taller than z_i)
  const T *Nu, *Pr;
  T Prc;
  int ppf, ny, ld, nx, nz;
  T rdz, dt;
  int do_step, euler;
  ForcingArgs<T> fo;  // damp == 2: sponge term about the plane mean
  int sfield[2];      // plane-sum index of each scalar (3 theta, 4 q)
  static constexpr double writes() { return 7.0; }
  // load phase: every array value the update of point (plane, y, x) reads
  struct Pre {
    T p0, p1, nu0, nu1, d0, d1, rf, s0;
    double flux;
  };
  __device__ Pre pre(int plane, int y, int x) const {
    const int q = plane / ppf, k = plane - q * ppf;
    const long pl = (long)ld * ny, i = ((long)k * ny + y) * ld + x;
    Pre p{};
    p.s0 = do_step ? s[q][i] : T(0);
    if (x >= nx) return p;
    p.p0 = Pr ? Pr[i] : Prc;
    p.p1 = Pr ? Pr[i + pl] : Prc;
    if (k == 0) {
      p.flux = sf[q](x + nx * y);
    } else {
      p.nu0 = Nu[i];
      p.d0 = dsdz[q][i];
    }
    p.nu1 = Nu[i + pl];
    p.d1 = dsdz[q][i + pl];
    p.rf = (do_step && !euler) ? Rf[q][i] : T(0);
    return p;
  }
  __device__ void fin(int plane, int y, int x, T v, const Pre& p) const {
    const int q = plane / ppf, k = plane - q * ppf;
    const long pl = (long)ld * ny, i = ((long)k * ny + y) * ld + x;
    T* Rq = R[q];
    if (k == nz - 2) Rq[i + pl] = T(0);  // level nz: not defined by the reference, 0 here
    if (x >= nx) {
      Rq[i] = T(0);
      if (do_step && k == nz - 2) s[q][i + pl] = p.s0 + gz[q];
      return;
    }
    T t0;
    if (k == 0) {
      sv1[q][x + (long)ld * y] = p.flux;
      t0 = (T)(-1.0 * p.flux);
    } else {
      t0 = (T(1) / p.p0) * p.nu0 * p.d0;
    }
    const T t1 = (T(1) / p.p1) * p.nu1 * p.d1;
    T rhs = v + rdz * (t1 - t0);
    if (fo.damp == 2) rhs = scalar_sponge(fo, rhs, sfield[q], k, ppf, p.s0);
    Rq[i] = rhs;
    if (!do_step) return;
    T fo;
    if (euler) {
      fo = rhs;
      Rf[q][i] = rhs;
    } else {
      fo = p.rf;
    }
    const T sn = p.s0 + dt * (T(1.5) * rhs - T(0.5) * fo);
    s[q][i] = sn;
    if (k == nz - 2) s[q][i + pl] = sn + gz[q];
  }
  __device__ void operator()(int plane, int y, int x, T v) const { fin(plane, y, x, v, pre(plane, y, x)); }
};

// horizontal SGS flux divergence of the scalars (spectral), plus -advection
// from X3 (x-spectral, ns*(nz-1) planes), then the d/dz flux term and the AB2
// step in the c2r epilogue
template <typename T>
static void scalar_sgs_finish(les_ctx* c, const ScalarJob* jobs, int ns, const T* Nu_t, const T* Pr_t,
                              const cx<T>* X3, int do_step, int euler, const ForcingArgs<T>* fo = nullptr) {
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1;
  const long rowsT = 2L * ns * nl * g.ny;
  cx<T>* XT = (cx<T>*)scratch(c, S_XS0, (size_t)rowsT * g.nxh * sizeof(cx<T>));
  LoadScalarTemp<T> LT{};
  for (int q = 0; q < ns; ++q) {
    LT.ds[0][q] = (const T*)jobs[q].dsdx;
    LT.ds[1][q] = (const T*)jobs[q].dsdy;
  }
  LT.Nu = Nu_t; LT.Pr = Pr_t; LT.Prc = (T)c->scal.Pr;
  LT.ns = ns; LT.ppf = nl; LT.ny = g.ny; LT.ld = g.ld; LT.cst = T(1.0 / ((double)g.nx * g.ny));
  auto all_fields = [&](auto nsc) {
    constexpr int NS = decltype(nsc)::value;
    LoadScalarTempAll<T, NS> LA{};
    for (int d = 0; d < 2; ++d)
      for (int q = 0; q < NS; ++q) LA.ds[d][q] = LT.ds[d][q];
    LA.Nu = LT.Nu; LA.Pr = LT.Pr; LA.Prc = LT.Prc; LA.ny = LT.ny; LA.ld = LT.ld; LA.cst = LT.cst;
    return launch_row_fwd_mf<T>(c, LA, XT, g.nxh, (Pr_t ? 2.0 : 1.0) + 2.0 * NS, nl);
  };
  const bool mf = ns == 2 ? all_fields(std::integral_constant<int, 2>{}) : all_fields(std::integral_constant<int, 1>{});
  if (!mf) launch_row_fwd<T>(c, LT, XT, g.nxh, rowsT, g.nx, g.nxh, g.ny);
  cx<T>* XD = (cx<T>*)scratch(c, S_XS1, (size_t)ns * nl * g.ny * g.nxh * sizeof(cx<T>));
  ColArgs d = col_args(g.ny, g.ny, g.nxh, g.nxh, (long)g.ny * g.nxh);
  d.in2_plane = d.in_plane;
  col_out(d, 0, K_ID, g.nxh, g.nxh, (long)g.ny * g.nxh);
  launch_col<T>(c, XT, XT + (size_t)ns * nl * g.ny * g.nxh, XD, nullptr, nullptr, d, (long)ns * nl, g.nxh);
  StoreScalarRHS<T> S{};
  for (int q = 0; q < ns; ++q) {
    S.s[q] = (T*)jobs[q].s;
    S.R[q] = (T*)jobs[q].RHS;
    S.Rf[q] = (T*)jobs[q].RHS_f;
    S.dsdz[q] = (const T*)jobs[q].dsdz;
    S.sv1[q] = jobs[q].sgs_vert1;
    S.sf[q] = surface_flux(c, jobs[q].which);
    S.gz[q] = (T)(g.Lz > 1.0 ? lid_gradient(c, jobs[q].which) * g.dz : 0.0);
    S.sfield[q] = 3 + jobs[q].which;
  }
  if (fo && fo->damp == 2 && do_step) S.fo = *fo;
  S.Nu = Nu_t; S.Pr = Pr_t; S.Prc = (T)c->scal.Pr;
  S.ppf = nl; S.ny = g.ny; S.ld = g.ld; S.nx = g.nx; S.nz = nz;
  S.rdz = T(1.0 / g.dz); S.dt = T(g.dt);
  S.do_step = do_step; S.euler = euler;
  launch_row_inv<T>(c, XD, g.nxh, g.nxh, S, (long)ns * nl * g.ny, g.nx, g.ld, g.ny, X3, T(-1));
}

template <typename T>
void op_scalar_rhs(les_ctx* c, const ScalarJob* jobs, int ns, const T* u, const T* v, const T* w, const T* Nu_t,
                   const T* Pr_t, int do_step, int euler) {
  OpTag _tag(c, "scalar_rhs");
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1, nf = 3 + 3 * ns;
  // dealias1 of u, v, w and the gradients: x-spectral rows, padded in y
  const long rows = (long)nf * nz * g.ny;
  cx<T>* XS = (cx<T>*)scratch(c, S_XS0, (size_t)rows * g.nxh * sizeof(cx<T>));
  LoadFieldsN<T, 9> L{};
  L.p[0] = u; L.p[1] = v; L.p[2] = w;
  for (int q = 0; q < ns; ++q) {
    L.p[3 + 3 * q] = (const T*)jobs[q].dsdx;
    L.p[4 + 3 * q] = (const T*)jobs[q].dsdy;
    L.p[5 + 3 * q] = (const T*)jobs[q].dsdz;
  }
  L.ppf = nz; L.ny = g.ny; L.ld = g.ld; L.scale = T(1.0 / ((double)g.nx * g.ny));
  const int tl = xs_layout(c, g.nxh);
  launch_row_fwd<T>(c, L, XS, g.nxh, rows, g.nx, g.nxh, g.ny, tl);
  const long rowsB = (long)nf * nz * g.ny2;
  cx<T>* XB = (cx<T>*)scratch(c, S_XS1, (size_t)rowsB * g.nxh * sizeof(cx<T>));
  ColArgs a = col_args(g.ny, g.ny2, g.nxh, g.nxh, (long)g.ny * g.nxh);
  col_out(a, 0, K_ID, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
  a.in_tl = tl;
  launch_col<T>(c, XS, nullptr, XB, nullptr, nullptr, a, (long)nf * nz, g.nxh);
  // products on the padded grid (dealias2 scaling included)
  cx<T>* XP = (cx<T>*)scratch(c, S_XS2, (size_t)ns * nl * g.ny2 * g.nxh * sizeof(cx<T>));
  const T cst = T(1.0 / ((double)g.nx2 * g.ny2));
  if (ns == 2)
    launch_pad_products<T, PolScalar<T, 2>, 7, 3>(c, XB, XP, PolScalar<T, 2>{cst}, "scalar_big");
  else
    launch_pad_products<T, PolScalar<T, 1>, 8, 4>(c, XB, XP, PolScalar<T, 1>{cst}, "scalar_big");
  // truncate in y (unpadd)
  cx<T>* X3 = (cx<T>*)scratch(c, S_XS3, (size_t)ns * nl * g.ny * g.nxh * sizeof(cx<T>));
  ColArgs b = col_args(g.ny2, g.ny, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
  col_out(b, 0, K_ID, g.nxh, g.nxh, (long)g.ny * g.nxh);
  launch_col<T>(c, XP, nullptr, X3, nullptr, nullptr, b, (long)ns * nl, g.nxh);
  scalar_sgs_finish<T>(c, jobs, ns, Nu_t, Pr_t, X3, do_step, euler);
}

// convec + RHS assembly + AB2 of the momentum (main.f90:304-390) and both
// scalars' scalar_RHS_calc + step_scalar in one padded-grid pass: u, v, w and
// the vorticity are transformed once for both
template <typename T>
struct LoadConvecScalar {
  LoadConvec<T> lc;
  const T* gs[6];  // dtheta/dx,dy,dz, dq/dx,dy,dz
  static constexpr double reads() { return 15.0 / 12.0; }
  __device__ T operator()(int plane, int y, int x) const {
    const int f = plane / lc.nz;
    if (f < 6) return lc(plane, y, x);
    const int k = plane - f * lc.nz;
    return lc.cst * gs[f - 6][((long)k * lc.ny + y) * lc.ld + x];
  }
};

template <typename T>
void op_convec_scalar_ab2(les_ctx* c, T* u, T* v, T* w, const T* const gr[9], T* const R[3], T* const Rf[3],
                          const T* const divt[3], int euler_first, const T* beta, const ScalarJob* jobs,
                          const T* Nu_t, const T* Pr_t, int scal_euler, const ForcingArgs<T>* fo,
                          const SpecIn<T>* si) {
  OpTag _tag(c, "convec_scalar");
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1;
  const long rowsB = 12L * nz * g.ny2;
  cx<T>* XB = (cx<T>*)scratch(c, S_XS1, (size_t)rowsB * g.nxh * sizeof(cx<T>));
  const int tl = xs_layout(c, g.nxh);
  bool padded = false;
  if (si && si->uvw && si->s[0] && si->s[1]) {
    // the inputs from filt_da's kept spectra; wall dudz, dvdz (wallstress) and
    // the scalars' surface and lid gradients by one small forward transform
    const size_t pl = (size_t)g.ld * g.ny, top = pl * (nz - 1);
    cx<T>* W = (cx<T>*)scratch(c, S_FWALL, (size_t)6 * g.ny * g.nxh * sizeof(cx<T>));
    const T *d0 = (const T*)jobs[0].dsdz, *d1 = (const T*)jobs[1].dsdz;
    LoadPlanes6<T> LW{{gr[2], gr[5], d0, d0 + top, d1, d1 + top}, g.ld, T(1.0 / g.nx)};
    launch_row_fwd<T>(c, LW, W, g.nxh, 6L * g.ny, g.nx, g.nxh, g.ny, tl);
    ConvecSpecIn<T, 2> fn{si->uvw, {si->s[0], si->s[1]}, W, nz, g.ny, g.nxh, T(1.0 / g.ny), T(1.0 / g.dz)};
    padded = launch_col_pad_fn<T>(c, fn, XB, 12L * nz, g.nxh, g.nxh, (long)g.ny2 * g.nxh, 1.6, tl);
  }
  if (!padded) {
    const long rows = 12L * nz * g.ny;
    cx<T>* XS = (cx<T>*)scratch(c, S_XS0, (size_t)rows * g.nxh * sizeof(cx<T>));
    LoadConvecScalar<T> L{};
    L.lc = LoadConvec<T>{u, v, w, gr[1], gr[2], gr[3], gr[5], gr[6], gr[7], g.ny, nz, g.ld,
                         T(1.0 / ((double)g.nx * g.ny))};
    for (int q = 0; q < 2; ++q) {
      L.gs[3 * q] = (const T*)jobs[q].dsdx;
      L.gs[3 * q + 1] = (const T*)jobs[q].dsdy;
      L.gs[3 * q + 2] = (const T*)jobs[q].dsdz;
    }
    launch_row_fwd<T>(c, L, XS, g.nxh, rows, g.nx, g.nxh, g.ny, tl);
    ColArgs a = col_args(g.ny, g.ny2, g.nxh, g.nxh, (long)g.ny * g.nxh);
    col_out(a, 0, K_ID, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
    a.in_tl = tl;
    launch_col<T>(c, XS, nullptr, XB, nullptr, nullptr, a, 12L * nz, g.nxh);
  }
  cx<T>* XP = (cx<T>*)scratch(c, S_XS2, (size_t)5 * nl * g.ny2 * g.nxh * sizeof(cx<T>));
  launch_pad_products<T, PolConvecScalar<T>, 4, 1>(c, XB, XP, PolConvecScalar<T>{T(1.0 / ((double)g.nx2 * g.ny2))},
                                                   "convec_scalar_big");
  cx<T>* X3 = (cx<T>*)scratch(c, S_XS3, (size_t)5 * nl * g.ny * g.nxh * sizeof(cx<T>));
  ColArgs b = col_args(g.ny2, g.ny, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
  col_out(b, 0, K_ID, g.nxh, g.nxh, (long)g.ny * g.nxh);
  launch_col<T>(c, XP, nullptr, X3, nullptr, nullptr, b, 5L * nl, g.nxh);
  // scalars first: their epilogue reads no velocity, the momentum one updates it
  scalar_sgs_finish<T>(c, jobs, 2, Nu_t, Pr_t, X3 + (size_t)3 * nl * g.ny * g.nxh, 1, scal_euler, fo);
  StoreConvecAB2<T> SC{{R[0], R[1], R[2]}, {Rf[0], Rf[1], Rf[2]}, {divt[0], divt[1], divt[2]}, {u, v, w},
                       nl, g.ny, g.ld, g.nx, (T)g.dt, (T)g.force, euler_first, beta};
  if (fo) SC.fo = *fo;
  launch_row_inv<T>(c, X3, g.nxh, g.nxh, SC, 3L * nl * g.ny, g.nx, g.ld, g.ny);
  const size_t pl = (size_t)g.ld * g.ny, top = pl * nl;
  T* tops[9] = {R[0] + top, R[1] + top, R[2] + top, u + top, v + top, w + top, Rf[0] + top, Rf[1] + top,
                Rf[2] + top};
  op_fill_n<T>(c, tops, euler_first ? 9 : 6, pl, T(BOGUS));
}

// ------------------------------------------------------------------------
// Coriolis and sponge inputs of one step (main.f90:319-348): the products
// coriol*v, coriol*u of the velocities before the AB2 update, and the plane
// sums of u, v, w, theta, q on levels 1..nz-1 (compensated, as every plane
// sum here) for damping_method 2
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_coriolis_pre(const T* __restrict__ u, const T* __restrict__ v, T* cu, T* cv, T coriol, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    cv[i] = coriol * v[i];
    cu[i] = coriol * u[i];
  }
}

template <typename T>
struct SumLevels5 {  // level lev = field * nl + k of (u, v, w, theta, q); absent fields sum to 0
  const T* f[5];
  int nl, ld, ny;
  __device__ double operator()(int lev, int y, int x) const {
    const int q = lev / nl, k = lev - q * nl;
    return f[q] ? (double)f[q][((long)k * ny + y) * ld + x] : 0.0;
  }
};

template <typename T>
ForcingArgs<T> op_forcing_pre(les_ctx* c, const T* u, const T* v, const T* w, const T* theta, const T* q) {
  ForcingArgs<T> fa;
  const Forcing& F = c->frc;
  if (!F.coriolis && F.ubc != 1) return fa;
  OpTag _tag(c, "forcing");
  const Grid& g = c->g;
  const int nl = g.nz - 1;
  const long nv = (long)g.ld * g.ny * nl;
  if (F.coriolis) {
    T* cc = (T*)scratch(c, S_FORCE, 2 * (size_t)nv * sizeof(T));
    {
      ProfScope _ps(c, "coriolis_pre", 4.0 * nv * sizeof(T));
      hipLaunchKernelGGL((k_coriolis_pre<T>), grid1((size_t)nv), dim3(256), 0, c->stream, u, v, cc, cc + nv,
                         (T)F.coriol, nv);
    }
    LCHK(c, hipGetLastError());
    fa.cor[0] = cc + nv;  // coriol * v feeds RHSx
    fa.cor[1] = cc;       // coriol * u feeds RHSy
    fa.cvg = (T)F.cvg;
    fa.cug = (T)F.cug;
  }
  if (F.ubc == 1) {
    fa.damp = F.damping_method;
    fa.coef = F.dcoef;
    fa.nxy = (double)g.nx * g.ny;
    if (F.damping_method == 2) {
      double* sums = (double*)scratch(c, S_FSUM, 5 * (size_t)nl * sizeof(double));
      level_sums(c, SumLevels5<T>{{u, v, w, theta, q}, nl, g.ld, g.ny}, 5 * nl, sums,
                 (theta ? 5.0 : 3.0) * g.nx * g.ny * nl * sizeof(T));
      fa.sums = sums;
    }
  }
  return fa;
}

// step_scalar, scalars_module.f90:457-504 (routine-level entry point)
template <typename T>
__global__ void k_step_scalar(T* s, const T* R, const T* Rf, int nx, int ld, int ny, int nz, T dt, T gz) {
  const long pl = (long)ld * ny, n = pl * (nz - 1);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    T sn = s[i];
    if (x < nx) {
      sn = s[i] + dt * (T(1.5) * R[i] - T(0.5) * Rf[i]);
      s[i] = sn;
    }
    if (k == nz - 2) s[i + pl] = sn + gz;
  }
}

template <typename T>
void op_step_scalar(les_ctx* c, int which, T* s, const T* RHS, const T* RHS_f) {
  OpTag _tag(c, "step_scalar");
  const Grid& g = c->g;
  const T gz = (T)(g.Lz > 1.0 ? lid_gradient(c, which) * g.dz : 0.0);
  hipLaunchKernelGGL((k_step_scalar<T>), grid1((size_t)g.ld * g.ny * (g.nz - 1)), dim3(256), 0, c->stream, s, RHS,
                     RHS_f, g.nx, g.ld, g.ny, g.nz, T(g.dt), gz);
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// lagrange_scalar_Sdep.f90: 10 new fields, each filtered at 2 Delta and
// 4 Delta; u, v, w bars and S_ij come from the momentum model's filtered set
// ------------------------------------------------------------------------
template <typename T>
struct LoadLagScalar {  // theta_b | u_b th_b, v_b th_b, w_b th_b | |S| dT (x,y,z) | dT bars (x,y,z)
  const T *u, *v, *w, *th, *S11, *S12, *S13, *S22, *S23, *S33, *dx, *dy, *dz;
  int ny, nz, ld;
  static constexpr double reads() { return 13.0 / 10.0; }
  __device__ T operator()(int plane, int y, int x) const {
    const int f = plane / nz, k = plane - f * nz;
    const long i = ((long)k * ny + y) * ld + x, pl = (long)ld * ny;
    if (f < 4) {  // :139-150
      T tb;
      if (k == 0) tb = th[i];
      else tb = T(0.5) * (th[i] + th[i - pl]);
      if (f == 0) return tb;
      T ub;
      if (f == 1) ub = k == 0 ? u[i] : T(0.5) * (u[i] + u[i - pl]);
      else if (f == 2) ub = k == 0 ? v[i] : T(0.5) * (v[i] + v[i - pl]);
      else ub = k == 0 ? T(0.25) * w[i + pl] : w[i];
      return ub * tb;
    }
    if (f < 7) {  // :240-248
      const T S = smag(S11[i], S12[i], S13[i], S22[i], S23[i], S33[i]);
      if (k == 0) {
        if (f == 4) return S * dx[i];
        if (f == 5) return S * dy[i];
        return S * (dz[i] + dz[i + pl]) / T(2);
      }
      if (f == 4) return S * (dx[i] + dx[i - pl]) / T(2);
      if (f == 5) return S * (dy[i] + dy[i - pl]) / T(2);
      return S * dz[i];
    }
    if (k == 0) {  // :262-270
      if (f == 7) return dx[i];
      if (f == 8) return dy[i];
      return (dz[i] + dz[i + pl]) / T(2);
    }
    if (f == 7) return (dx[i] + dx[i - pl]) / T(2);
    if (f == 8) return (dy[i] + dy[i - pl]) / T(2);
    return dz[i];
  }
};

// all ten LoadLagScalar fields of a point at once (k_row_fwd_mf), in
// LoadLagScalar's evaluation order
template <typename T>
struct LoadLagScalarAll {
  static constexpr int NF = 10;
  LoadLagScalar<T> L;
  __device__ void all(int k, int y, int x, T (&v)[NF]) const {
    const long i = ((long)k * L.ny + y) * L.ld + x, pl = (long)L.ld * L.ny;
    const T tb = k == 0 ? L.th[i] : T(0.5) * (L.th[i] + L.th[i - pl]);
    const T ub = k == 0 ? L.u[i] : T(0.5) * (L.u[i] + L.u[i - pl]);
    const T vb = k == 0 ? L.v[i] : T(0.5) * (L.v[i] + L.v[i - pl]);
    const T wb = k == 0 ? T(0.25) * L.w[i + pl] : L.w[i];
    v[0] = tb;
    v[1] = ub * tb;
    v[2] = vb * tb;
    v[3] = wb * tb;
    const T S = smag(L.S11[i], L.S12[i], L.S13[i], L.S22[i], L.S23[i], L.S33[i]);
    if (k == 0) {
      v[4] = S * L.dx[i];
      v[5] = S * L.dy[i];
      v[6] = S * (L.dz[i] + L.dz[i + pl]) / T(2);
      v[7] = L.dx[i];
      v[8] = L.dy[i];
      v[9] = (L.dz[i] + L.dz[i + pl]) / T(2);
    } else {
      v[4] = S * (L.dx[i] + L.dx[i - pl]) / T(2);
      v[5] = S * (L.dy[i] + L.dy[i - pl]) / T(2);
      v[6] = S * L.dz[i];
      v[7] = (L.dx[i] + L.dx[i - pl]) / T(2);
      v[8] = (L.dy[i] + L.dy[i - pl]) / T(2);
      v[9] = L.dz[i];
    }
  }
};

template <typename T>
__global__ void k_lag_scalar_epilogue(const T* __restrict__ F, const T* __restrict__ Fs, const T* __restrict__ TnLM,
                                      const T* __restrict__ TnQN, const T* __restrict__ Cs, const T* __restrict__ Gi,
                                      T* GLM, T* GMM, T* GQN, T* GNN, T* Beta, T* Cs2_Pr, T* Pr_t, int nx, int ld,
                                      int ny, int nz, double cst, double opftdelta, double lagran_dt, double expo,
                                      int init_lm, int init_qn) {
  const long pl = (long)ld * ny, n3 = pl * nz;
  const double e24 = (double)1e-24f;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    const bool pad = x >= nx;
    auto f2m = [&](int f) { return (double)F[f * n3 + i]; };
    auto f4m = [&](int f) { return (double)F[(21 + f) * n3 + i]; };
    auto f2s = [&](int f) { return (double)Fs[f * n3 + i]; };
    auto f4s = [&](int f) { return (double)Fs[(10 + f) * n3 + i]; };
    const double ub = f2m(0), vb = f2m(1), wb = f2m(2), uh = f4m(0), vh = f4m(1), wh = f4m(2);
    const double tb = f2s(0), th = f4s(0);
    const double L1 = f2s(1) - ub * tb, L2 = f2s(2) - vb * tb, L3 = f2s(3) - wb * tb;  // :164-169
    const double Q1 = f4s(1) - uh * th, Q2 = f4s(2) - vh * th, Q3 = f4s(3) - wh * th;  // :176-181
    double sb[6], sh[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) { sb[m] = f2m(9 + m); sh[m] = f4m(9 + m); }
    const double Sb = sqrt(2.0 * (sb[0] * sb[0] + sb[3] * sb[3] + sb[5] * sb[5] +
                                  2.0 * (sb[1] * sb[1] + sb[2] * sb[2] + sb[4] * sb[4])));
    const double Sh = sqrt(2.0 * (sh[0] * sh[0] + sh[3] * sh[3] + sh[5] * sh[5] +
                                  2.0 * (sh[1] * sh[1] + sh[2] * sh[2] + sh[4] * sh[4])));
    double M[3], N[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {  // :268-274
      M[q] = cst * (f2s(4 + q) - 4.0 * Sb * f2s(7 + q));
      N[q] = cst * (f4s(4 + q) - 16.0 * Sh * f4s(7 + q));
    }
    const double LM = L1 * M[0] + L2 * M[1] + L3 * M[2];
    const double MM = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
    const double QN = Q1 * N[0] + Q2 * N[1] + Q3 * N[2];
    const double NN = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
    // the interpolated G (interpolag_scalar_Sdep's output, four fields of Gi)
    double glm = Gi[i], gmm = Gi[n3 + i], gqn = Gi[2 * n3 + i], gnn = Gi[3 * n3 + i];
    if (init_lm) { gmm = pad ? 1.0 : MM; glm = pad ? 1.0 : 0.03 * MM; }  // :289-299
    double Tn = (double)TnLM[i];                                          // :318-331
    Tn = opftdelta * pow(Tn, -0.125);
    Tn = fmax(e24, fq(Tn));
    double dumfac = lagran_dt / Tn;
    double epsi = dumfac / (1.0 + dumfac);
    glm = (epsi * LM + (1.0 - epsi) * glm);
    gmm = (epsi * MM + (1.0 - epsi) * gmm);
    glm = fmax(e24, fq(glm));
    const double c2 = fmax(e24, fq(pad ? e24 : glm / gmm));
    if (init_qn) { gnn = pad ? 1.0 : NN; gqn = pad ? 1.0 : 0.03 * NN; }  // :338-348
    Tn = (double)TnQN[i];                                                 // :350-363
    Tn = opftdelta * pow(Tn, -0.125);
    Tn = fmax(e24, fq(Tn));
    dumfac = lagran_dt / Tn;
    epsi = dumfac / (1.0 + dumfac);
    gqn = (epsi * QN + (1.0 - epsi) * gqn);
    gnn = (epsi * NN + (1.0 - epsi) * gnn);
    gqn = fmax(e24, fq(gqn));
    const double c4 = fmax(e24, fq(pad ? e24 : gqn / gnn));
    double beta = expo == 1.0 ? c4 / c2 : pow(c4 / c2, expo);  // :370; the exponent is ln2/(ln4-ln2) = 1
    if (k == nz - 1) beta = 1.0;
    const double bclip = fmax(fq(beta), fq(0.125));
    const double cpr = fmax(e24, fq(pad ? e24 : c2 / bclip));  // :384-392
    GLM[i] = (T)glm; GMM[i] = (T)gmm; GQN[i] = (T)gqn; GNN[i] = (T)gnn;
    Beta[i] = (T)beta;
    Cs2_Pr[i] = (T)cpr;
    Pr_t[i] = (T)((double)Cs[i] / cpr);  // :395
  }
}

template <typename T>
void op_lagrange_scalar(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, T* theta,
                        const SgsPtrs& m, const ScalarSgsPtrs& ss, bool reuse_mom) {
  OpTag _tag(c, "lagrange_scalar");
  const Grid& g = c->g;
  const ScalarParams& P = c->scal;
  const int nz = g.nz;
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const T* Gi;
  {  // interpolag_scalar_Sdep (:88) into scratch; the epilogue writes G
    OpTag _t2(c, "interpolag_scalar");
    const T* G[4] = {(const T*)ss.G_LM, (const T*)ss.G_MM, (const T*)ss.G_QN, (const T*)ss.G_NN};
    Gi = op_interpolag4_into<T>(c, G, (T*)ss.u_lag_s, (T*)ss.v_lag_s, (T*)ss.w_lag_s);
  }
  // theta_temp = theta (:84-86), then filt_da(theta) in place and ddz_uv (:95-99)
  T* tht = (T*)scratch(c, S_SC0, n3 * sizeof(T));
  T* dTx = (T*)scratch(c, S_SC1, n3 * sizeof(T));
  T* dTy = (T*)scratch(c, S_SC2, n3 * sizeof(T));
  T* dTz = (T*)scratch(c, S_SC3, n3 * sizeof(T));
  op_copy<T>(c, tht, theta, n3);
  op_filt_da<T>(c, theta, dTx, dTy, nz);
  double* red = (double*)scratch(c, S_RED2, (R2_BETA + nz) * sizeof(double));
  level_sums(c, SumSpeed<T>{u, v, g.ld}, 1, red + R2_LAG, 2.0 * g.nx * g.ny * sizeof(T));
  level_sums(c, SumPlane<double>{ss.psi_m, g.nx, 0.0}, 1, red + R2_LAG + 1, (double)g.nx * g.ny * 8);
  {
    DzArgs<T> a{};
    a.ns = 1;
    a.s[0] = theta;
    a.dz[0] = dTz;
    a.top[0] = (T)lid_gradient(c, 0);
    a.lag = 1;
    a.phi_h = ss.phi_h;
    a.lagsums = red + R2_LAG;
    a.zlog = log(0.5 * g.dz / c->zo_avg);
    a.sflux = c->wt_mean / P.T_scale / P.u_star;  // :114
    ProfScope _ps(c, "scalar_dz", 2.0 * n3 * sizeof(T));
    hipLaunchKernelGGL((k_scalar_dz<T, 4>), dim3((unsigned)((n3 + 1023) / 1024)), dim3(256), 0, c->stream, a, g.nx,
                       g.ld, g.ny, nz, g.dz, T(1.0 / g.dz), (unsigned)n3);
  }
  LCHK(c, hipGetLastError());
  const T* F = reuse_mom ? (const T*)scratch(c, S_LAGF, 42 * n3 * sizeof(T)) : lag_filter_mom<T>(c, S, u, v, w);
  // the ten scalar fields through both test filters
  const int k2 = c->keep2, k4 = c->keep4, kf = k2 > k4 ? k2 : k4;
  const long rows = 10L * nz * g.ny;
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, (size_t)rows * kf * sizeof(cx<T>));
  LoadLagScalar<T> L{u, v, w, tht, S[0], S[1], S[2], S[3], S[4], S[5], dTx, dTy, dTz, g.ny, nz, g.ld};
  if (!launch_row_fwd_mf<T>(c, LoadLagScalarAll<T>{L}, X0, kf, 13.0))
    launch_row_fwd<T>(c, L, X0, kf, rows, g.nx, kf, g.ny);
  cx<T>* X2 = (cx<T>*)scratch(c, S_XS1, (size_t)rows * k2 * sizeof(cx<T>));
  cx<T>* X4 = (cx<T>*)scratch(c, S_XS2, (size_t)rows * k4 * sizeof(cx<T>));
  ColArgs a = col_args(g.ny, g.ny, kf, kf, (long)g.ny * kf);
  a.lhG = g.lh;
  col_out(a, 0, K_G, k2, k2, (long)g.ny * k2);
  col_out(a, 1, K_G, k4, k4, (long)g.ny * k4);
  const Tables<T>& t = tables<T>(c);
  launch_col<T>(c, X0, nullptr, X2, X4, nullptr, a, 10L * nz, kf, t.G2, t.G4);
  T* Fs = (T*)scratch(c, S_LAGF2, 20 * n3 * sizeof(T));
  StoreFields<T> S2{{Fs, nullptr, nullptr, nullptr}, 10 * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X2, k2, k2, S2, rows, g.nx, g.ld, g.ny);
  StoreFields<T> S4{{Fs + 10 * n3, nullptr, nullptr, nullptr}, 10 * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X4, k4, k4, S4, rows, g.nx, g.ld, g.ny);
  const int init_lm = g.inilag && !c->glm_init && (c->jt == lag_count(c) || c->jt == 1);
  const int init_qn = g.inilag && !c->gqn_init && (c->jt == lag_count(c) || c->jt == 1);
  const double delta = g.delta;
  {
    ProfScope _ps(c, "lag_scalar_epilogue", (6.0 + 20 + 2 + 1 + 4 + 4 + 3) * n3 * sizeof(T));
    hipLaunchKernelGGL((k_lag_scalar_epilogue<T>), grid1(n3), dim3(256), 0, c->stream, F, Fs, (const T*)m.TnLMMM,
                       (const T*)m.TnQNNN, (const T*)m.Cs_opt2, Gi, (T*)ss.G_LM, (T*)ss.G_MM, (T*)ss.G_QN,
                       (T*)ss.G_NN,
                       (T*)m.Beta, (T*)ss.Cs2_Pr, (T*)ss.Pr_t, g.nx, g.ld, g.ny, nz, delta * delta, 1.5 * delta,
                       g.dt * (double)lag_count(c), log(2.0) / (log(4.0) - log(2.0)), init_lm, init_qn);
  }
  LCHK(c, hipGetLastError());
  if (init_lm) c->glm_init = 1;
  if (init_qn) c->gqn_init = 1;
}

// ------------------------------------------------------------------------
// scalar_scaledep_dynamic.f90 (SGS model 6 with scalars): the plane-averaged
// scale-dependent Cs^2/Pr_sgs.  Ten fields through both test filters as in
// lagrange_scalar_Sdep, but theta is read after its in-place filt_da (:60)
// and |S| dT_i, dT_i are taken on level jz as they are (:179-200, no
// staggering average); Q_i filters products of the 2-Delta bars (:123-137),
// one more 4-Delta pass of 3 fields.  Plane sums over levels 1..nz-1.
// ------------------------------------------------------------------------
constexpr int SSD_NPART = 32;  // partial-sum blocks per level

// four block sums of 256 threads in a fixed order
__device__ void block_sums4(const double (&acc)[4], double* out) {
  __shared__ double sh[4][256];
#pragma unroll
  for (int s = 0; s < 4; ++s) sh[s][threadIdx.x] = acc[s];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int s = 0; s < 4; ++s) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = sh[threadIdx.x][0];
}

template <typename T>
struct LoadSdScalarAll {  // theta_b | u_b th_b, v_b th_b, w_b th_b | |S| dT (x,y,z) | dT (x,y,z)
  static constexpr int NF = 10;
  const T *u, *v, *w, *th, *S11, *S12, *S13, *S22, *S23, *S33, *dx, *dy, *dz;
  int ny, nz, ld;
  static constexpr double reads() { return 13.0 / 10.0; }
  __device__ void all(int k, int y, int x, T (&r)[NF]) const {
    const long i = ((long)k * ny + y) * ld + x, pl = (long)ld * ny;
    const T tb = k == 0 ? th[i] : T(0.5) * (th[i] + th[i - pl]);  // :84-105
    const T ub = k == 0 ? u[i] : T(0.5) * (u[i] + u[i - pl]);
    const T vb = k == 0 ? v[i] : T(0.5) * (v[i] + v[i - pl]);
    const T wb = k == 0 ? T(0.25) * w[i + pl] : w[i];
    r[0] = tb;
    r[1] = ub * tb;
    r[2] = vb * tb;
    r[3] = wb * tb;
    const T S = smag(S11[i], S12[i], S13[i], S22[i], S23[i], S33[i]);  // :140-142
    r[4] = S * dx[i];
    r[5] = S * dy[i];
    r[6] = S * dz[i];
    r[7] = dx[i];
    r[8] = dy[i];
    r[9] = dz[i];
  }
  __device__ T operator()(int plane, int y, int x) const {
    const int f = plane / nz, k = plane - f * nz;
    T r[NF];
    all(k, y, x, r);
    return r[f];
  }
};

// u_bar_i theta_bar of the 2-Delta bars (:123-125): the input of Q_i's filter
template <typename T>
struct LoadBarTheta {
  const T *F, *Fs;  // momentum filter set (u, v, w bars: fields 0..2); scalar set (theta bar: field 0)
  long n3;
  int ny, nz, ld;
  static constexpr double reads() { return 2.0; }
  __device__ T operator()(int plane, int y, int x) const {
    const int q = plane / nz, k = plane - q * nz;
    const long i = ((long)k * ny + y) * ld + x;
    return F[q * n3 + i] * Fs[i];
  }
};

// :211-231: L_i.M_i, M_i.M_i, Q_i.N_i, N_i.N_i per level, const = Delta^2,
// M_i = const (|S| dT_i bar - 4 |S_bar| dT_i bar), N_i with 16 at 4 Delta
template <typename T>
__global__ void __launch_bounds__(256) k_ssd_sums(const T* __restrict__ F, const T* __restrict__ Fs,
                                                  const T* __restrict__ Q3, long n3, int nx, int ld, int ny,
                                                  double cst, double* part) {
  const int k = blockIdx.y, n = nx * ny;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / nx, x = j - y * nx;
    const long i = ((long)k * ny + y) * ld + x;
    auto f2m = [&](int f) { return (double)F[f * n3 + i]; };
    auto f4m = [&](int f) { return (double)F[(21 + f) * n3 + i]; };
    auto f2s = [&](int f) { return (double)Fs[f * n3 + i]; };
    auto f4s = [&](int f) { return (double)Fs[(10 + f) * n3 + i]; };
    const double tb = f2s(0), th = f4s(0);
    double sb[6], sh[6], L[3], Q[3], M[3], N[3];
#pragma unroll
    for (int m = 0; m < 6; ++m) { sb[m] = f2m(9 + m); sh[m] = f4m(9 + m); }
    const double Sb = sqrt(2.0 * (sb[0] * sb[0] + sb[3] * sb[3] + sb[5] * sb[5] +
                                  2.0 * (sb[1] * sb[1] + sb[2] * sb[2] + sb[4] * sb[4])));
    const double Sh = sqrt(2.0 * (sh[0] * sh[0] + sh[3] * sh[3] + sh[5] * sh[5] +
                                  2.0 * (sh[1] * sh[1] + sh[2] * sh[2] + sh[4] * sh[4])));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      L[q] = f2s(1 + q) - f2m(q) * tb;                    // :116-121
      Q[q] = (double)Q3[q * n3 + i] - f4m(q) * th;        // :132-137
      M[q] = cst * (f2s(4 + q) - 4.0 * Sb * f2s(7 + q));  // :212-214
      N[q] = cst * (f4s(4 + q) - 16.0 * Sh * f4s(7 + q));
    }
    a[0] += L[0] * M[0] + L[1] * M[1] + L[2] * M[2];
    a[1] += M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
    a[2] += Q[0] * N[0] + Q[1] * N[1] + Q[2] * N[2];
    a[3] += N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
  }
  block_sums4(a, part + ((long)k * gridDim.x + blockIdx.x) * 4);
}

// per level :230-240 (the clips of the momentum routine) and Pr_t = Cs_opt2 /
// Cs2_Pr (:242) over the whole (ld, ny) plane; Cs2_Pr(jz) into element jz-1
template <typename T>
__global__ void k_ssd_cs(const double* __restrict__ part, int nparts, int nl, double* cpr, T* Cs2_Pr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nl) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int p = 0; p < nparts; ++p)
    for (int q = 0; q < 4; ++q) s[q] += part[((long)k * nparts + p) * 4 + q];
  const double c2 = (double)fmaxf(1e-24f, (float)(s[0] / s[1]));
  const double c4 = (double)fmaxf(1e-24f, (float)(s[2] / s[3]));
  const double beta = c4 / c2;
  const double betaclip = (double)fmaxf((float)beta, (float)(1.0 / 8.0));
  const double v = fmax(0.0, c2 / betaclip);
  cpr[k] = v;
  Cs2_Pr[k] = (T)v;
}
// Model 6 leaves Cs_opt2 unclipped below (Cs2_2d^2 / Cs2_4d can reach
// 1e-48), so in FP32-bulk a Cs_opt2 that underflows to 0 would make Pr_t 0
// and the scalar flux Nu_t / Pr_t 0/0; there Pr_t is kept at or above the
// smallest normal float (the flux is then 0, as the FP64 ratio nearly is)
template <typename T>
__global__ void k_ssd_prt(const T* __restrict__ Cs, const double* __restrict__ cpr, T* Pr_t, unsigned pl,
                          unsigned n) {
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double r = (double)Cs[i] / cpr[i / pl];
    if constexpr (sizeof(T) == 4) r = fmax(r, 1.1754943508222875e-38);
    Pr_t[i] = (T)r;
  }
}

template <typename T>
void op_scalar_scaledep(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, T* theta,
                        const T* Cs_opt2, const ScalarSgsPtrs& ss, bool reuse_mom) {
  OpTag _tag(c, "scalar_scaledep");
  const Grid& g = c->g;
  const ScalarParams& P = c->scal;
  const int nz = g.nz, nl = nz - 1;
  const size_t n3 = (size_t)g.ld * g.ny * nz, pl = (size_t)g.ld * g.ny;
  T* dTx = (T*)scratch(c, S_SC1, n3 * sizeof(T));
  T* dTy = (T*)scratch(c, S_SC2, n3 * sizeof(T));
  T* dTz = (T*)scratch(c, S_SC3, n3 * sizeof(T));
  op_filt_da<T>(c, theta, dTx, dTy, nz);  // :60, theta filtered in place
  double* red = (double*)scratch(c, S_RED2, (R2_BETA + nz) * sizeof(double));
  level_sums(c, SumSpeed<T>{u, v, g.ld}, 1, red + R2_LAG, 2.0 * g.nx * g.ny * sizeof(T));
  level_sums(c, SumPlane<double>{ss.psi_m, g.nx, 0.0}, 1, red + R2_LAG + 1, (double)g.nx * g.ny * 8);
  level_sums(c, SumPlane<double>{ss.phi_h, g.nx, 0.0}, 1, red + R2_LAG + 2, (double)g.nx * g.ny * 8);
  {  // ddz_uv with the lid gradient and the first-level value of :61-79
    DzArgs<T> a{};
    a.ns = 1;
    a.s[0] = theta;
    a.dz[0] = dTz;
    a.top[0] = (T)lid_gradient(c, 0);
    a.lag = 2;
    a.lagsums = red + R2_LAG;
    a.zlog = log(0.5 * g.dz / c->zo_avg);
    a.sflux = c->wt_mean / P.T_scale / P.u_star;  // :78
    ProfScope _ps(c, "scalar_dz", 2.0 * n3 * sizeof(T));
    hipLaunchKernelGGL((k_scalar_dz<T, 4>), dim3((unsigned)((n3 + 1023) / 1024)), dim3(256), 0, c->stream, a, g.nx,
                       g.ld, g.ny, nz, g.dz, T(1.0 / g.dz), (unsigned)n3);
  }
  LCHK(c, hipGetLastError());
  const T* F = reuse_mom ? (const T*)scratch(c, S_LAGF, 42 * n3 * sizeof(T)) : lag_filter_mom<T>(c, S, u, v, w);
  const int k2 = c->keep2, k4 = c->keep4, kf = k2 > k4 ? k2 : k4;
  const Tables<T>& t = tables<T>(c);
  const long rows = 10L * nz * g.ny;
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, (size_t)rows * kf * sizeof(cx<T>));
  LoadSdScalarAll<T> L{u, v, w, theta, S[0], S[1], S[2], S[3], S[4], S[5], dTx, dTy, dTz, g.ny, nz, g.ld};
  if (!launch_row_fwd_mf<T>(c, L, X0, kf, 13.0)) launch_row_fwd<T>(c, L, X0, kf, rows, g.nx, kf, g.ny);
  cx<T>* X2 = (cx<T>*)scratch(c, S_XS1, (size_t)rows * k2 * sizeof(cx<T>));
  cx<T>* X4 = (cx<T>*)scratch(c, S_XS2, (size_t)rows * k4 * sizeof(cx<T>));
  ColArgs a = col_args(g.ny, g.ny, kf, kf, (long)g.ny * kf);
  a.lhG = g.lh;
  col_out(a, 0, K_G, k2, k2, (long)g.ny * k2);
  col_out(a, 1, K_G, k4, k4, (long)g.ny * k4);
  launch_col<T>(c, X0, nullptr, X2, X4, nullptr, a, 10L * nz, kf, t.G2, t.G4);
  T* Fs = (T*)scratch(c, S_LAGF2, 20 * n3 * sizeof(T));
  StoreFields<T> S2{{Fs, nullptr, nullptr, nullptr}, 10 * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X2, k2, k2, S2, rows, g.nx, g.ld, g.ny);
  StoreFields<T> S4{{Fs + 10 * n3, nullptr, nullptr, nullptr}, 10 * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X4, k4, k4, S4, rows, g.nx, g.ld, g.ny);
  // Q_i: Gtt(u_bar_i theta_bar)
  const long rq = 3L * nz * g.ny;
  cx<T>* Y0 = (cx<T>*)scratch(c, S_XS0, (size_t)rq * k4 * sizeof(cx<T>));
  cx<T>* Y1 = (cx<T>*)scratch(c, S_XS1, (size_t)rq * k4 * sizeof(cx<T>));
  launch_row_fwd<T>(c, LoadBarTheta<T>{F, Fs, (long)n3, g.ny, nz, g.ld}, Y0, k4, rq, g.nx, k4, g.ny);
  ColArgs aq = col_args(g.ny, g.ny, k4, k4, (long)g.ny * k4);
  aq.lhG = g.lh;
  col_out(aq, 0, K_G, k4, k4, (long)g.ny * k4);
  launch_col<T>(c, Y0, nullptr, Y1, nullptr, nullptr, aq, 3L * nz, k4, t.G4);
  T* Q3 = (T*)scratch(c, S_SDQ, 3 * n3 * sizeof(T));
  StoreFields<T> SQ{{Q3, nullptr, nullptr, nullptr}, 3 * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, Y1, k4, k4, SQ, rq, g.nx, g.ld, g.ny);
  double* part = (double*)scratch(c, S_RED, (size_t)SSD_NPART * nl * 4 * sizeof(double));
  double* cpr = (double*)scratch(c, S_SDLEV, 2 * (size_t)nz * sizeof(double));
  {
    ProfScope _ps(c, "ssd_sums", (42.0 + 20 + 3) * (double)g.nx * g.ny * nl * sizeof(T));
    hipLaunchKernelGGL((k_ssd_sums<T>), dim3(SSD_NPART, nl), dim3(256), 0, c->stream, F, (const T*)Fs,
                       (const T*)Q3, (long)n3, g.nx, g.ld, g.ny, g.delta * g.delta, part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL((k_ssd_cs<T>), dim3((nl + 63) / 64), dim3(64), 0, c->stream, (const double*)part, SSD_NPART,
                     nl, cpr, (T*)ss.Cs2_Pr);
  LCHK(c, hipGetLastError());
  {
    const size_t n = pl * nl;
    ProfScope _ps(c, "ssd_prt", 2.0 * n * sizeof(T));
    hipLaunchKernelGGL((k_ssd_prt<T>), grid1(n), dim3(256), 0, c->stream, Cs_opt2, (const double*)cpr,
                       (T*)ss.Pr_t, (unsigned)pl, (unsigned)n);
  }
  LCHK(c, hipGetLastError());
}

// the padded-product pass of a convective step alone, on the scratch buffers
// the step uses (the autotuner times its variants): momentum and both scalars
// (scal = 1) or momentum only
template <typename T>
void op_pad_products_probe(les_ctx* c, int scal) {
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1, nf = scal ? 12 : 6, no = scal ? 5 : 3;
  cx<T>* XB = (cx<T>*)scratch(c, S_XS1, (size_t)nf * nz * g.ny2 * g.nxh * sizeof(cx<T>));
  cx<T>* XP = (cx<T>*)scratch(c, S_XS2, (size_t)no * nl * g.ny2 * g.nxh * sizeof(cx<T>));
  if (!c->err.empty()) return;
  const T cst = T(1.0 / ((double)g.nx2 * g.ny2));
  if (scal)
    launch_pad_products<T, PolConvecScalar<T>, 4, 1>(c, XB, XP, PolConvecScalar<T>{cst}, "pad_probe");
  else
    launch_pad_products<T, PolConvec<T>, 8, 3>(c, XB, XP, PolConvec<T>{cst}, "pad_probe");
}

#define LES_SCAL_INST(T)                                                                                         \
  template void op_obukhov<T>(les_ctx*, const T*, const T*, const T*, double*, double*, double*, double*,        \
                              double*, double*, double*, double*, double*, double);                              \
  template void op_calcbeta<T>(les_ctx*, const T*, const T*, T*, double, double);                                \
  template void op_scalar_dz<T>(les_ctx*, const ScalarJob*, int, const double*, const double*, int);             \
  template void op_scalar_rhs<T>(les_ctx*, const ScalarJob*, int, const T*, const T*, const T*, const T*,        \
                                 const T*, int, int);                                                            \
  template void op_step_scalar<T>(les_ctx*, int, T*, const T*, const T*);                                        \
  template void op_convec_scalar_ab2<T>(les_ctx*, T*, T*, T*, const T* const[9], T* const[3], T* const[3],        \
                                        const T* const[3], int, const T*, const ScalarJob*, const T*, const T*,   \
                                        int, const ForcingArgs<T>*, const SpecIn<T>*);                            \
  template ForcingArgs<T> op_forcing_pre<T>(les_ctx*, const T*, const T*, const T*, const T*, const T*);          \
  template void op_lagrange_scalar<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, T*,             \
                                      const SgsPtrs&, const ScalarSgsPtrs&, bool);                               \
  template void op_scalar_scaledep<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, T*, const T*,  \
                                      const ScalarSgsPtrs&, bool);                                               \
  template void op_pad_products_probe<T>(les_ctx*, int);

LES_SCAL_INST(double)
LES_SCAL_INST(float)

}  // namespace les
