// les_dynamic.hip -- the plane-averaged dynamic SGS models on gfx950: model 2
// (std_dynamic.f90), model 3 (scaledep_dynamic.f90 with rtnewt of
// sgsmodule.f90:37-59, as sgs_stag.f90:128-140 calls it) and model 6
// (new_scaledep_dynamic.f90, sgs_stag.f90:143-148).
//
// Per level they need the 21 fields that lagrange_Sdep filters at 2 Delta
// and 4 Delta -- lag_filter_mom serves all three -- and models 3 and 6 also
// Q_ij, whose products are of the already 2-Delta-filtered velocities
// (scaledep_dynamic.f90:94-115): one more batched transform of 6 fields.
// Model 3's ten plane sums of the quintic (:183-215), then L.M, M.M
// (:233-244), and model 6's four sums are deterministic two-pass reductions
// in double (NPART blocks per level, fixed-order final add); the quintic's
// Newton solve runs on the device, one thread per level, so every model
// stays inside a captured step.
#include "les_xform.h"

// Evaluated as written (no multiply-add contraction), like the CPU
// restatement: rtnewt's Newton path from x = 7.5 can be long, and its
// rounding should differ from the reference's only through the order of
// the plane sums.
#pragma clang fp contract(off)

namespace les {

constexpr int SD_NPART = 32;
constexpr int SD_NS = 10;  // a1 b1 c1 d1 e1 a2 b2 c2 d2 e2 plane sums

// index pairs of the symmetric tensor components 11 12 13 22 23 33
__device__ __forceinline__ int sym_a(int f) { return f < 3 ? 0 : (f < 5 ? 1 : 2); }
__device__ __forceinline__ int sym_b(int f) { return f < 3 ? f : (f < 5 ? f - 2 : 2); }

// a:b with the off-diagonal components counted twice
__device__ __forceinline__ double sym_dot(const double* a, const double* b) {
  return a[0] * b[0] + a[3] * b[3] + a[5] * b[5] + 2.0 * (a[1] * b[1] + a[2] * b[2] + a[4] * b[4]);
}
__device__ __forceinline__ double sym_norm(const double* s) { return sqrt(2.0 * sym_dot(s, s)); }

// products of the 2-Delta filtered velocities (fields 0..2 of F2), the input
// of the 4-Delta filter that gives Q_ij (scaledep_dynamic.f90:94-99)
template <typename T>
struct LoadBarProducts {
  const T* F2;
  long n3;
  int ny, nz, ld;
  static constexpr double reads() { return 2.0; }
  __device__ T operator()(int plane, int y, int x) const {
    const int f = plane / nz, k = plane - f * nz;
    const long i = ((long)k * ny + y) * ld + x;
    return F2[sym_a(f) * n3 + i] * F2[sym_b(f) * n3 + i];
  }
};

// the filtered quantities of one plane point (scaledep_dynamic.f90:78-181):
// L, Q, S_bar_ij, S_hat_ij, |S|S_ij bar and hat, |S_bar|, |S_hat|
template <typename T>
struct SdPoint {
  double L[6], Q[6], sb[6], sh[6], ssb[6], ssh[6], Sb, Sh;
  __device__ void load(const T* F2, const T* F4, const T* Q6, long n3, long i) {
    auto f2 = [&](int f) { return (double)F2[f * n3 + i]; };
    auto f4 = [&](int f) { return (double)F4[f * n3 + i]; };
    const double ub[3] = {f2(0), f2(1), f2(2)}, uh[3] = {f4(0), f4(1), f4(2)};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      L[q] = f2(3 + q) - ub[sym_a(q)] * ub[sym_b(q)];
      Q[q] = (double)Q6[q * n3 + i] - uh[sym_a(q)] * uh[sym_b(q)];
      sb[q] = f2(9 + q);
      sh[q] = f4(9 + q);
      ssb[q] = f2(15 + q);
      ssh[q] = f4(15 + q);
    }
    Sb = sym_norm(sb);
    Sh = sym_norm(sh);
  }
};

template <int NS>
__device__ void block_sum_out(double* acc, double* out) {
  __shared__ double sh[NS][256];
#pragma unroll
  for (int s = 0; s < NS; ++s) sh[s][threadIdx.x] = acc[s];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int s = 0; s < NS; ++s) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < NS) out[threadIdx.x] = sh[threadIdx.x][0];
}

// the ten sums of :183-215 without their constant factors; pads hold zeros
// in every filtered field, so the plane loop stops at nx
template <typename T>
__global__ void __launch_bounds__(256) k_sd_sums(const T* __restrict__ F2, const T* __restrict__ F4,
                                                 const T* __restrict__ Q6, long n3, int nx, int ld, int ny,
                                                 double* part) {
  const int k = blockIdx.y, n = nx * ny;
  double a[SD_NS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / nx, x = j - y * nx;
    SdPoint<T> p;
    p.load(F2, F4, Q6, n3, ((long)k * ny + y) * ld + x);
    a[0] += p.Sb * sym_dot(p.sb, p.L);
    a[1] += sym_dot(p.ssb, p.L);
    a[2] += sym_dot(p.ssb, p.ssb);
    a[3] += 0.5 * ((p.Sb * p.Sb) * (p.Sb * p.Sb));  // S_bar**4, an integer power
    a[4] += p.Sb * sym_dot(p.sb, p.ssb);
    a[5] += p.Sh * sym_dot(p.sh, p.Q);
    a[6] += sym_dot(p.ssh, p.Q);
    a[7] += sym_dot(p.ssh, p.ssh);
    a[8] += 0.5 * ((p.Sh * p.Sh) * (p.Sh * p.Sh));
    a[9] += p.Sh * sym_dot(p.sh, p.ssh);
  }
  block_sum_out<SD_NS>(a, part + ((long)k * gridDim.x + blockIdx.x) * SD_NS);
}

// rtnewt, sgsmodule.f90:37-59: Newton from the midpoint of (0, 15), stop at
// |dx| < 1e-3, beta = 1 after 100 iterations (also when f/df is not finite)
__device__ double sd_rtnewt(const double* A) {
  double x = 0.5 * (0.0 + 15.0);
  for (int j = 1; j <= 100; ++j) {
    const double f = A[0] + x * (A[1] + x * (A[2] + x * (A[3] + x * (A[4] + x * A[5]))));
    const double df = A[1] + x * (2.0 * A[2] + x * (3.0 * A[3] + x * (4.0 * A[4] + x * (5.0 * A[5]))));
    const double dx = f / df;
    x = x - dx;
    if (fabs(dx) < 0.001) return x;
  }
  return 1.0;
}

// one thread per level: the quintic of :187-222 and its root (:223)
__global__ void k_sd_beta(const double* __restrict__ part, int nparts, int nz, double delta, double nxy,
                          double* beta) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  double s[SD_NS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = 0; p < nparts; ++p)
    for (int q = 0; q < SD_NS; ++q) s[q] += part[((long)k * nparts + p) * SD_NS + q];
  const double d2 = delta * delta, cc = (2.0 * d2) * (2.0 * d2);
  const double a1 = -(2.0 * d2 * 4.0 * s[0] / nxy);
  const double b1 = -(2.0 * d2 * s[1] / nxy);
  const double c1 = cc * s[2] / nxy;
  const double d1 = cc * 16.0 * s[3] / nxy;
  const double e1 = 2.0 * cc * 4.0 * s[4] / nxy;
  const double a2 = -(2.0 * d2 * 16.0 * s[5] / nxy);
  const double b2 = -(2.0 * d2 * s[6] / nxy);
  const double c2 = cc * s[7] / nxy;
  const double dd2 = cc * 256.0 * s[8] / nxy;
  const double e2 = 2.0 * cc * 16.0 * s[9] / nxy;
  const double A[6] = {b2 * c1 - b1 * c2,       a1 * c2 - b2 * e1,  b2 * d1 + b1 * e2 - a2 * c1,
                       a2 * e1 - a1 * e2,       -a2 * d1 - b1 * dd2, a1 * dd2};
  beta[k] = sd_rtnewt(A);
}

// L.M and M.M with M_ij = 2 Delta^2 (|S|S_ij bar - 4 beta |S_bar| S_ij bar), :225-242
template <typename T>
__global__ void __launch_bounds__(256) k_sd_lm(const T* __restrict__ F2, const T* __restrict__ F4,
                                               const T* __restrict__ Q6, long n3, int nx, int ld, int ny,
                                               const double* __restrict__ beta, double cst, double* part) {
  const int k = blockIdx.y, n = nx * ny;
  const double b = beta[k];
  double a[2] = {0.0, 0.0};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / nx, x = j - y * nx;
    const long i = ((long)k * ny + y) * ld + x;
    double L[6], M[6];
    const double ub[3] = {(double)F2[i], (double)F2[n3 + i], (double)F2[2 * n3 + i]};
    double sb[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      L[q] = (double)F2[(3 + q) * n3 + i] - ub[sym_a(q)] * ub[sym_b(q)];
      sb[q] = (double)F2[(9 + q) * n3 + i];
    }
    const double Sb = sym_norm(sb);
#pragma unroll
    for (int q = 0; q < 6; ++q) M[q] = cst * ((double)F2[(15 + q) * n3 + i] - 4.0 * b * Sb * sb[q]);
    a[0] += sym_dot(L, M);
    a[1] += sym_dot(M, M);
  }
  (void)F4;
  (void)Q6;
  block_sum_out<2>(a, part + ((long)k * gridDim.x + blockIdx.x) * 2);
}

// Cs_opt2(jz) = max(0, sum L.M / sum M.M), :243-244
__global__ void k_sd_cs(const double* __restrict__ part, int nparts, int nz, double* cs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  double num = 0.0, den = 0.0;
  for (int p = 0; p < nparts; ++p) {
    num += part[((long)k * nparts + p) * 2];
    den += part[((long)k * nparts + p) * 2 + 1];
  }
  cs[k] = fmax(0.0, num / den);
}

// sgs_stag.f90:133-135: every point of level jz takes Cs_opt2(jz)
template <typename T>
__global__ void k_sd_fill(const double* __restrict__ cs, T* Cs, long pl, long n3) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x)
    Cs[i] = (T)cs[(unsigned)i / (unsigned)pl];
}

// Q_ij input of models 3 and 6: the 4-Delta filter of products of the
// 2-Delta filtered velocities (scaledep_dynamic.f90:94-115,
// new_scaledep_dynamic.f90:94-115) into scratch, 6 fields (ld,ny,nz)
template <typename T>
static const T* sd_q6(les_ctx* c, const T* F2) {
  const Grid& g = c->g;
  const int nz = g.nz, k4 = c->keep4;
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const long rows = 6L * nz * g.ny;
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, (size_t)rows * k4 * sizeof(cx<T>));
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, (size_t)rows * k4 * sizeof(cx<T>));
  launch_row_fwd<T>(c, LoadBarProducts<T>{F2, (long)n3, g.ny, nz, g.ld}, X0, k4, rows, g.nx, k4, g.ny);
  ColArgs a = col_args(g.ny, g.ny, k4, k4, (long)g.ny * k4);
  a.lhG = g.lh;
  col_out(a, 0, K_G, k4, k4, (long)g.ny * k4);
  launch_col<T>(c, X0, nullptr, X1, nullptr, nullptr, a, 6L * nz, k4, tables<T>(c).G4);
  T* Q6 = (T*)scratch(c, S_SDQ, 6 * n3 * sizeof(T));
  StoreFields<T> SQ{{Q6, nullptr, nullptr, nullptr}, 6 * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X1, k4, k4, SQ, rows, g.nx, g.ld, g.ny);
  return Q6;
}

// SGS model 6, new_scaledep_dynamic.f90:183-205: per level L.M, M.M, Q.N and
// N.N with M_ij = 2 Delta^2 (|S|S_ij bar - 4 |S_bar| S_ij bar) and N_ij the
// same at 4 Delta with 16
template <typename T>
__global__ void __launch_bounds__(256) k_nsd_sums(const T* __restrict__ F2, const T* __restrict__ F4,
                                                  const T* __restrict__ Q6, long n3, int nx, int ld, int ny,
                                                  double cst, double* part) {
  const int k = blockIdx.y, n = nx * ny;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / nx, x = j - y * nx;
    SdPoint<T> p;
    p.load(F2, F4, Q6, n3, ((long)k * ny + y) * ld + x);
    double M[6], N[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      M[q] = cst * (p.ssb[q] - 4.0 * p.Sb * p.sb[q]);
      N[q] = cst * (p.ssh[q] - 16.0 * p.Sh * p.sh[q]);
    }
    a[0] += sym_dot(p.L, M);
    a[1] += sym_dot(M, M);
    a[2] += sym_dot(p.Q, N);
    a[3] += sym_dot(N, N);
  }
  block_sum_out<4>(a, part + ((long)k * gridDim.x + blockIdx.x) * 4);
}

// :210-220: LM/MM and QN/NN clipped at 1E-24 in single precision, Beta their
// ratio clipped below at 1/8 in single precision, Cs_opt2 = max(0, LM/MM /
// betaclip); the same rule gives Cs2_Pr of scalar_scaledep_dynamic.f90:230-240
__device__ __forceinline__ double sd_clipped_ratio(double lm, double mm, double qn, double nn) {
  const double c2 = (double)fmaxf(1e-24f, (float)(lm / mm));
  const double c4 = (double)fmaxf(1e-24f, (float)(qn / nn));
  const double beta = c4 / c2;
  const double betaclip = (double)fmaxf((float)beta, (float)(1.0 / 8.0));
  return fmax(0.0, c2 / betaclip);
}

__global__ void k_nsd_cs(const double* __restrict__ part, int nparts, int nlev, double* cs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nlev) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int p = 0; p < nparts; ++p)
    for (int q = 0; q < 4; ++q) s[q] += part[((long)k * nparts + p) * 4 + q];
  cs[k] = sd_clipped_ratio(s[0], s[1], s[2], s[3]);
}

template <typename T>
void op_scaledep_dynamic(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, double* cs_lev,
                         double* beta_lev, T* Cs_planes) {
  OpTag _tag(c, "scaledep");
  const Grid& g = c->g;
  const int nz = g.nz;
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const T* F = lag_filter_mom<T>(c, S, u, v, w);  // F2 = F, F4 = F + 21 n3
  const T* F2 = F;
  const T* F4 = F + 21 * n3;
  const T* Q6 = sd_q6<T>(c, F2);  // :94-115
  double* part = (double*)scratch(c, S_RED, (size_t)SD_NPART * nz * SD_NS * sizeof(double));
  double* lev = (double*)scratch(c, S_SDLEV, 2 * (size_t)nz * sizeof(double));
  double* beta = beta_lev ? beta_lev : lev;
  double* cs = cs_lev ? cs_lev : lev + nz;
  {
    ProfScope _ps(c, "sd_sums", (42.0 + 6) * (double)g.nx * g.ny * nz * sizeof(T));
    hipLaunchKernelGGL((k_sd_sums<T>), dim3(SD_NPART, nz), dim3(256), 0, c->stream, F2, F4, (const T*)Q6, (long)n3,
                       g.nx, g.ld, g.ny, part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_sd_beta, dim3((nz + 63) / 64), dim3(64), 0, c->stream, (const double*)part, SD_NPART, nz,
                     g.delta, (double)g.nx * g.ny, beta);
  LCHK(c, hipGetLastError());
  {
    ProfScope _ps(c, "sd_lm", 21.0 * (double)g.nx * g.ny * nz * sizeof(T));
    hipLaunchKernelGGL((k_sd_lm<T>), dim3(SD_NPART, nz), dim3(256), 0, c->stream, F2, F4, (const T*)Q6, (long)n3,
                       g.nx, g.ld, g.ny, (const double*)beta, 2.0 * g.delta * g.delta, part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_sd_cs, dim3((nz + 63) / 64), dim3(64), 0, c->stream, (const double*)part, SD_NPART, nz, cs);
  LCHK(c, hipGetLastError());
  if (Cs_planes) {
    ProfScope _ps(c, "sd_fill", (double)n3 * sizeof(T));
    hipLaunchKernelGGL((k_sd_fill<T>), grid1(n3), dim3(256), 0, c->stream, (const double*)cs, Cs_planes,
                       (long)g.ld * g.ny, (long)n3);
  }
  LCHK(c, hipGetLastError());
}

// SGS model 2, std_dynamic.f90:4-140: the L.M / M.M sums of the
// scale-dependent model with beta = 1 (M_ij = 2 Delta^2 (|S|S_ij bar - 4
// |S_bar| S_ij bar), :124-132), Delta = (dx dy dz)^(1/3) (:28), over the
// same 2-Delta filter set
template <typename T>
void op_std_dynamic(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, double* cs_lev,
                    T* Cs_planes) {
  OpTag _tag(c, "std_dynamic");
  const Grid& g = c->g;
  const int nz = g.nz;
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const T* F2 = lag_filter_mom<T>(c, S, u, v, w);
  double* part = (double*)scratch(c, S_RED, (size_t)SD_NPART * nz * 2 * sizeof(double));
  double* lev = (double*)scratch(c, S_SDLEV, 2 * (size_t)nz * sizeof(double));
  double* beta = lev;
  double* cs = cs_lev ? cs_lev : lev + nz;
  op_fill<double>(c, beta, nz, 1.0);
  const double delta = pow(g.dx * g.dy * g.dz, 1.0 / 3.0);
  {
    ProfScope _ps(c, "sd_lm", 21.0 * (double)g.nx * g.ny * nz * sizeof(T));
    hipLaunchKernelGGL((k_sd_lm<T>), dim3(SD_NPART, nz), dim3(256), 0, c->stream, F2, (const T*)nullptr,
                       (const T*)nullptr, (long)n3, g.nx, g.ld, g.ny, (const double*)beta, 2.0 * (delta * delta),
                       part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_sd_cs, dim3((nz + 63) / 64), dim3(64), 0, c->stream, (const double*)part, SD_NPART, nz, cs);
  LCHK(c, hipGetLastError());
  if (Cs_planes) {
    ProfScope _ps(c, "sd_fill", (double)n3 * sizeof(T));
    hipLaunchKernelGGL((k_sd_fill<T>), grid1(n3), dim3(256), 0, c->stream, (const double*)cs, Cs_planes,
                       (long)g.ld * g.ny, (long)n3);
  }
  LCHK(c, hipGetLastError());
}

// SGS model 6, new_scaledep_dynamic.f90 (sgs_stag.f90:143-148): the model-3
// filter set (lag_filter_mom on calc_Sij's S_ij, then Q_ij), four plane sums
// per level and the clipped ratios, one Cs_opt2 per level
template <typename T>
void op_new_scaledep_dynamic(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, double* cs_lev,
                             T* Cs_planes) {
  OpTag _tag(c, "new_scaledep");
  const Grid& g = c->g;
  const int nz = g.nz;
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const T* F = lag_filter_mom<T>(c, S, u, v, w);
  const T* F2 = F;
  const T* F4 = F + 21 * n3;
  const T* Q6 = sd_q6<T>(c, F2);
  double* part = (double*)scratch(c, S_RED, (size_t)SD_NPART * nz * 4 * sizeof(double));
  double* lev = (double*)scratch(c, S_SDLEV, 2 * (size_t)nz * sizeof(double));
  double* cs = cs_lev ? cs_lev : lev + nz;
  {
    ProfScope _ps(c, "sd_sums", (42.0 + 6) * (double)g.nx * g.ny * nz * sizeof(T));
    hipLaunchKernelGGL((k_nsd_sums<T>), dim3(SD_NPART, nz), dim3(256), 0, c->stream, F2, F4, Q6, (long)n3, g.nx,
                       g.ld, g.ny, 2.0 * (g.delta * g.delta), part);
  }
  LCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_nsd_cs, dim3((nz + 63) / 64), dim3(64), 0, c->stream, (const double*)part, SD_NPART, nz, cs);
  LCHK(c, hipGetLastError());
  if (Cs_planes) {
    ProfScope _ps(c, "sd_fill", (double)n3 * sizeof(T));
    hipLaunchKernelGGL((k_sd_fill<T>), grid1(n3), dim3(256), 0, c->stream, (const double*)cs, Cs_planes,
                       (long)g.ld * g.ny, (long)n3);
  }
  LCHK(c, hipGetLastError());
}

#define LES_SD_INST(T)                                                                                      \
  template void op_scaledep_dynamic<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, double*, \
                                       double*, T*);                                                        \
  template void op_std_dynamic<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, double*, T*);      \
  template void op_new_scaledep_dynamic<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, double*,  \
                                           T*);
LES_SD_INST(double)
LES_SD_INST(float)
#undef LES_SD_INST

}  // namespace les
