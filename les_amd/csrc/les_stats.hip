// les_stats.hip -- statistics of the flow as FP64 reductions on gfx950:
//   * the current statistics of avg_stats (io.f90:1722-1755): per level the
//     plane means of u, v, w and txz, the variances of the w-node velocity
//     fluctuations, the resolved shear stress <u'w'> and its sum with <txz>,
//     and E11(k1)/z from spectrum (io.f90:1613-1646);
//   * the plane averages obukhov reports as its time series
//     (scalars_module.f90:545-561,655-661): u* from the log law over the
//     plane-mean speed, the mean local u*, u* from the plane-mean stress, the
//     first-level theta, theta_v and q, the surface heat flux and the
//     Obukhov length.
// Every plane sum is compensated (CSum), so the results are the exact sums
// rounded once whatever the reduction order; inputs of either precision are
// widened to double before any arithmetic.  A level's sums are split over
// SPART blocks of 256 threads (levels x SPART workgroups fill the chip), and a
// second kernel merges the partials in a fixed order.
#include "les_xform.h"

namespace les {

constexpr int SPART = 32;

template <typename T>
struct StatIn {
  const T *u, *v, *w, *txz;
  int nx, ny, ld, nl;
};

// merge NQ per-thread compensated sums over the block; thread 0 writes the
// block's NQ (s, c) pairs to out
template <int NQ>
__device__ void block_csum(CSum (&a)[NQ], double* out) {
  __shared__ double sh[2 * NQ][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    sh[2 * q][t] = a[q].s;
    sh[2 * q + 1][t] = a[q].c;
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const CSum m = csum_merge(CSum{sh[2 * q][t], sh[2 * q + 1][t]}, CSum{sh[2 * q][t + s], sh[2 * q + 1][t + s]});
        sh[2 * q][t] = m.s;
        sh[2 * q + 1][t] = m.c;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < 2 * NQ; ++q) out[q] = sh[q][0];
  }
}

// pass 1: plane sums of u, v, w, txz on level blockIdx.y (io.f90:1725-1727,1749)
template <typename T>
__global__ void __launch_bounds__(256) k_stat_means(StatIn<T> a, double* part) {
  const int k = blockIdx.y, n = a.nx * a.ny;
  const long base = (long)k * a.ny * a.ld;
  CSum s[4] = {};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / a.nx, x = j - y * a.nx;
    const long i = base + (long)y * a.ld + x;
    s[0] = csum_add(s[0], (double)a.u[i]);
    s[1] = csum_add(s[1], (double)a.v[i]);
    s[2] = csum_add(s[2], (double)a.w[i]);
    s[3] = csum_add(s[3], (double)a.txz[i]);
  }
  block_csum<4>(s, part + ((long)k * gridDim.x + blockIdx.x) * 8);
}

// pass 2: the w-node fluctuations of io.f90:1729-1742 (zero on the first
// level) and their second moments and <u'w'> (:1744-1748)
template <typename T>
__global__ void __launch_bounds__(256) k_stat_moments(StatIn<T> a, const double* __restrict__ stats, double* part) {
#pragma clang fp contract(off)
  const int k = blockIdx.y, n = a.nx * a.ny, nl = a.nl;
  const long pl = (long)a.ny * a.ld, base = (long)k * pl;
  CSum s[4] = {};
  if (k > 0) {
    const double ub = stats[k], ubm = stats[k - 1], vb = stats[nl + k], vbm = stats[nl + k - 1];
    const double wb = stats[2 * nl + k];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
      const int y = j / a.nx, x = j - y * a.nx;
      const long i = base + (long)y * a.ld + x;
      const double up = 0.5 * ((double)a.u[i] - ub + (double)a.u[i - pl] - ubm);
      const double vp = 0.5 * ((double)a.v[i] - vb + (double)a.v[i - pl] - vbm);
      const double wp = (double)a.w[i] - wb;
      s[0] = csum_add(s[0], up * up);
      s[1] = csum_add(s[1], vp * vp);
      s[2] = csum_add(s[2], wp * wp);
      s[3] = csum_add(s[3], up * wp);
    }
  }
  block_csum<4>(s, part + ((long)k * gridDim.x + blockIdx.x) * 8);
}

// merge the SPART partials of quantity q on level k and divide by n; dst[q]
// is the quantity's offset in the stats array.  With `stress` set, quantity 3
// is <u'w'> and the level's <u'w'> + <txz> (io.f90:1750) is written too.
struct StatDst {
  int off[4];
};
__global__ void k_stat_final(const double* __restrict__ part, int nparts, int nl, double n, StatDst dst, int stress,
                             double* stats) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nl) return;
  for (int q = 0; q < 4; ++q) {
    CSum a{0.0, 0.0};
    for (int p = 0; p < nparts; ++p) {
      const double* e = part + ((long)k * nparts + p) * 8 + 2 * q;
      a = csum_merge(a, CSum{e[0], e[1]});
    }
    stats[dst.off[q] + k] = (a.s + a.c) / n;
  }
  if (stress) stats[8 * nl + k] = stats[6 * nl + k] + stats[7 * nl + k];
}

// spectrum (io.f90:1613-1646) from the x coefficients of f/nx: level
// blockIdx.y, wavenumber j = threadIdx + blockIdx.x * blockDim.x < nx/2;
// spec(0) = <c_0^2>/2, spec(j) = <|c_j|^2>, averaged over the ny rows; with
// eoz the level's value is divided by its height (k + 1/2) dz (io.f90:1754)
template <typename T>
__global__ void k_spec_reduce(const cx<T>* __restrict__ X, int ny, int nk, double dz, int eoz,
                              double* __restrict__ spec) {
#pragma clang fp contract(off)
  const int k = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nk) return;
  CSum a{0.0, 0.0};
  const cx<T>* row = X + (long)k * ny * nk + j;
  for (int y = 0; y < ny; ++y) {
    const double re = (double)row[(long)y * nk].x, im = (double)row[(long)y * nk].y;
    a = csum_add(a, j == 0 ? 0.5 * re * re : re * re + im * im);
  }
  double v = (a.s + a.c) / (double)ny;
  if (eoz) v = v / (((double)(k + 1) - 0.5) * dz);
  spec[(long)k * nk + j] = v;
}

size_t stats_len(const les_ctx* c) {
  const size_t nl = (size_t)c->g.nz - 1;
  return 9 * nl + nl * (size_t)(c->g.nx / 2);
}

template <typename T>
void op_spectrum(les_ctx* c, const T* f, int nlev, double* spec, int eoz) {
  OpTag _tag(c, "spectrum");
  const Grid& g = c->g;
  const int nk = g.nx / 2;
  const long rows = (long)nlev * g.ny;
  cx<T>* X = (cx<T>*)scratch(c, S_STATX, (size_t)rows * nk * sizeof(cx<T>));
  LoadFields<T> L{{f, nullptr, nullptr, nullptr}, nlev, g.ny, g.ld, T(1.0 / (double)g.nx)};
  launch_row_fwd<T>(c, L, X, nk, rows, g.nx, nk, g.ny);
  {
    ProfScope _ps(c, "spec_reduce", (double)rows * nk * sizeof(cx<T>) + (double)nlev * nk * 8);
    hipLaunchKernelGGL((k_spec_reduce<T>), dim3((unsigned)((nk + 63) / 64), (unsigned)nlev), dim3(64), 0,
                       c->stream, X, g.ny, nk, g.dz, eoz, spec);
  }
  LCHK(c, hipGetLastError());
}

template <typename T>
void op_avg_stats(les_ctx* c, const T* u, const T* v, const T* w, const T* txz, double* stats) {
  OpTag _tag(c, "avg_stats");
  const Grid& g = c->g;
  const int nl = g.nz - 1;
  double* part = (double*)scratch(c, S_STATP, (size_t)nl * SPART * 8 * sizeof(double));
  StatIn<T> a{u, v, w, txz, g.nx, g.ny, g.ld, nl};
  const double n = (double)(g.nx * g.ny);
  const unsigned fb = (unsigned)((nl + 63) / 64);
  {
    ProfScope _ps(c, "stat_means", 4.0 * nl * g.nx * g.ny * sizeof(T));
    hipLaunchKernelGGL((k_stat_means<T>), dim3(SPART, nl), dim3(256), 0, c->stream, a, part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_stat_final, dim3(fb), dim3(64), 0, c->stream, part, SPART, nl, n,
                     StatDst{{0, nl, 2 * nl, 7 * nl}}, 0, stats);
  LCHK(c, hipGetLastError());
  {
    ProfScope _ps(c, "stat_moments", 5.0 * nl * g.nx * g.ny * sizeof(T));
    hipLaunchKernelGGL((k_stat_moments<T>), dim3(SPART, nl), dim3(256), 0, c->stream, a, stats, part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_stat_final, dim3(fb), dim3(64), 0, c->stream, part, SPART, nl, n,
                     StatDst{{3 * nl, 4 * nl, 5 * nl, 6 * nl}}, 1, stats);
  LCHK(c, hipGetLastError());
  op_spectrum<T>(c, u, nl, stats + 9 * nl, 1);
}

// obukhov's plane averages: sums over the first level of |u|, psi_m, psi_m0,
// ustar_avg, txz, tyz and (with scalars) theta, theta_v, q, sgs_t3 and the
// virtual flux sgs_t3 + 0.61 theta 0.001 q_scale sgs_q3
template <typename T>
struct SurfIn {
  const T *u, *v, *txz, *tyz, *theta, *qmix;
  const double *psi_m, *psi_m0, *ustar, *t3, *q3;
  int nx, ny, ld;
  double qs, tref, qref;
};
constexpr int NSURF = 11;

template <typename T>
__global__ void __launch_bounds__(256) k_surf_sums(SurfIn<T> a, double* part) {
#pragma clang fp contract(off)
  const int n = a.nx * a.ny;
  CSum s[NSURF] = {};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / a.nx, x = j - y * a.nx;
    const long i = (long)y * a.ld + x, i2 = (long)y * a.nx + x;
    const double uu = (double)a.u[i], vv = (double)a.v[i];
    s[0] = csum_add(s[0], sqrt(uu * uu + vv * vv));
    s[1] = csum_add(s[1], a.psi_m[i2]);
    s[2] = csum_add(s[2], a.psi_m0[i2]);
    s[3] = csum_add(s[3], a.ustar[i2]);
    s[4] = csum_add(s[4], (double)a.txz[i]);
    s[5] = csum_add(s[5], (double)a.tyz[i]);
    if (a.theta) {
      const double th = (double)a.theta[i] + a.tref, q = (double)a.qmix[i] + a.qref;
      s[6] = csum_add(s[6], th);
      s[7] = csum_add(s[7], th * (1 + 0.61 * 0.001 * q * a.qs));
      s[8] = csum_add(s[8], q);
      s[9] = csum_add(s[9], a.t3[i]);
      s[10] = csum_add(s[10], a.t3[i] + 0.61 * th * (0.001 * a.qs * a.q3[i]));
    }
  }
  block_csum<NSURF>(s, part + (long)blockIdx.x * 2 * NSURF);
}

struct SurfConst {
  double dz, z0, vonk, g_hat;
  int nx, ny, scal;
};
__global__ void k_surf_final(const double* __restrict__ part, int nparts, SurfConst k, double* out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double m[NSURF];
  for (int q = 0; q < NSURF; ++q) {
    CSum a{0.0, 0.0};
    for (int p = 0; p < nparts; ++p) a = csum_merge(a, CSum{part[(long)p * 2 * NSURF + 2 * q], part[(long)p * 2 * NSURF + 2 * q + 1]});
    m[q] = a.s + a.c;
  }
  const double n = (double)(k.nx * k.ny), nf = (double)(float)(k.nx * k.ny);
  // scalars_module.f90:557-558 with zo_ = zo (uniform roughness)
  const double ustar_ = m[0] / nf * k.vonk / (log(0.5 * k.dz / k.z0) - m[1] / n + m[2] / n);
  const double mtx = m[4] / n, mty = m[5] / n;
  out[0] = ustar_;
  out[1] = m[3] / n;                            // :655
  out[2] = pow(mtx * mtx + mty * mty, 0.25);    // :656-657
  for (int q = 3; q < 8; ++q) out[q] = 0.0;
  if (k.scal) {
    const double theta_ = m[6] / n, thetav_ = m[7] / nf, wtv_ = m[10] / nf;  // :560-561, :555
    out[3] = theta_;
    out[4] = thetav_;
    out[5] = m[8] / n;                          // :660
    out[6] = m[9] / n;                          // :554
    out[7] = wtv_ != 0.0 ? -(ustar_ * ustar_ * ustar_) / (k.vonk * (k.g_hat / theta_) * wtv_) : 0.0;  // :661
  }
}

template <typename T>
void op_surface_stats(les_ctx* c, const T* u, const T* v, const T* txz, const T* tyz, const double* psi_m,
                      const double* psi_m0, const double* ustar_avg, const T* theta, const T* qmix,
                      const double* sgs_t3, const double* sgs_q3, double th_ref, double q_ref, double* out8) {
  OpTag _tag(c, "surface_stats");
  const Grid& g = c->g;
  const ScalarParams& P = c->scal;
  const bool scal = theta && qmix && sgs_t3 && sgs_q3;
  double* part = (double*)scratch(c, S_STATP, (size_t)SPART * 2 * NSURF * sizeof(double));
  SurfIn<T> a{u, v, txz, tyz, scal ? theta : nullptr, qmix, psi_m, psi_m0, ustar_avg, sgs_t3, sgs_q3,
              g.nx, g.ny, g.ld, P.q_scale, th_ref, q_ref};
  {
    ProfScope _ps(c, "surf_sums", 11.0 * g.nx * g.ny * 8);
    hipLaunchKernelGGL((k_surf_sums<T>), dim3(SPART), dim3(256), 0, c->stream, a, part);
  }
  LCHK(c, hipGetLastError());
  const double g_hat = scal ? (P.g * P.z_i) / (P.u_star * P.u_star) : 0.0;  // :545
  hipLaunchKernelGGL(k_surf_final, dim3(1), dim3(64), 0, c->stream, part, SPART,
                     SurfConst{g.dz, g.z0, 0.41, g_hat, g.nx, g.ny, scal ? 1 : 0}, out8);
  LCHK(c, hipGetLastError());
}

__global__ void k_running_mean(double* avg, const double* __restrict__ cur, size_t len, double n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x)
    avg[i] = (n * avg[i] + cur[i]) / (n + 1.0);  // io.f90:1800-1813
}
void op_running_mean(les_ctx* c, double* avg, const double* cur, size_t len, long n) {
  const unsigned blocks = (unsigned)((len + 255) / 256 < 1024 ? (len + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_running_mean, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, c->stream, avg, cur, len,
                     (double)n);
  LCHK(c, hipGetLastError());
}

#define LES_STATS_INST(T)                                                                                        \
  template void op_spectrum<T>(les_ctx*, const T*, int, double*, int);                                           \
  template void op_avg_stats<T>(les_ctx*, const T*, const T*, const T*, const T*, double*);                     \
  template void op_surface_stats<T>(les_ctx*, const T*, const T*, const T*, const T*, const double*,           \
                                    const double*, const double*, const T*, const T*, const double*,            \
                                    const double*, double, double, double*);
LES_STATS_INST(double)
LES_STATS_INST(float)
#undef LES_STATS_INST

}  // namespace les
