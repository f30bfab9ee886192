// les_kernels.hip -- gfx950 kernels and host launchers for the LES hot path.
//
// Horizontal operators are built from three transform kernels:
//   k_row_fwd : real rows (x) -> first `kxo` Fourier coefficients, two real
//               rows packed into one complex FFT (x-spectral rows)
//   k_col     : LDS tile of TX adjacent kx columns x all ky of one plane;
//               forward y-FFT, a fused spectral operation (i kx, i ky, test
//               filter G, 3/2 zero-pad / truncate, divergence), inverse y-FFT
//   k_row_inv : x-spectral rows -> real rows (c2r), with a fused epilogue
// Every spectral operation of the reference zeroes the kx = nx/2 column and
// the ky = ny/2 row, so the x-spectral intermediates keep only kx < nx/2
// (and fewer for the sharp test filters, whose cutoff disc prunes columns).
// Loaders / storers are functors so that pointwise work before and after a
// transform (vorticity, staggered products, d/dz terms) is fused into the
// transform kernels instead of making extra passes over HBM.
#include "les_xform.h"

namespace les {

void check(les_ctx* c, hipError_t e, const char* what) {
  if (e != hipSuccess && c->err.empty()) {
    c->err = std::string(what) + ": " + hipGetErrorString(e);
  }
}

template <>
const Tables<double>& tables<double>(const les_ctx* c) { return c->td; }
template <>
const Tables<float>& tables<float>(const les_ctx* c) { return c->tf; }


// ------------------------------------------------------------------------
// filt_da.f90 / ddx.f90 / ddy.f90 / ddxy.f90
// ------------------------------------------------------------------------
// nf fields (up to 3) -> for each field the outputs kinds[0..nout): one
// row_fwd, one column pass and one row_inv over all fields.
template <typename T>
static const cx<T>* spectral_deriv(les_ctx* c, const T* const* fs, int nf, int nlev, T* const* outs,
                                   const int kinds[3], int nout, int keep_slot = -1) {
  const Grid& g = c->g;
  const long rows = (long)nlev * g.ny;
  const size_t fo = (size_t)nf * rows * g.nxh;
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, fo * sizeof(cx<T>));
  cx<T>* X1 = (cx<T>*)scratch(c, keep_slot >= 0 ? keep_slot : S_XS1, fo * nout * sizeof(cx<T>));
  LoadFields<T> L{{fs[0], nf > 1 ? fs[1] : nullptr, nf > 2 ? fs[2] : nullptr, nullptr}, nlev, g.ny, g.ld,
                  T(1.0 / ((double)g.nx * g.ny))};
  const int tl = xs_layout(c, g.nxh);
  launch_row_fwd<T>(c, L, X0, g.nxh, rows * nf, g.nx, g.nxh, g.ny, tl);
  ColArgs a = col_args(g.ny, g.ny, g.nxh, g.nxh, (long)g.ny * g.nxh);
  a.in_tl = tl;
  for (int j = 0; j < nout; ++j) {
    col_out(a, j, kinds[j], g.nxh, g.nxh, (long)g.ny * g.nxh);
    a.out_tl[j] = tl;
  }
  launch_col<T>(c, X0, nullptr, X1, X1 + fo, X1 + 2 * fo, a, (long)nlev * nf, g.nxh);
  // output rows are ordered (kind j, field f): outs[j*nf + f]
  StoreFields<T> S{};
  for (int q = 0; q < nf * nout; ++q) S.p[q] = outs[q];
  S.ppf = nlev;
  S.ny = g.ny;
  S.ld = g.ld;
  S.nx = g.nx;
  launch_row_inv<T>(c, X1, g.nxh, g.nxh, S, rows * nf * nout, g.nx, g.ld, g.ny, nullptr, T(1), tl);
  return X1;
}

template <typename T>
void op_filt_da(les_ctx* c, T* f, T* dfdx, T* dfdy, int nlev, int keep_slot, const cx<T>** spec) {
  OpTag _tag(c, "filt_da");
  const T* fs[1] = {f};
  T* outs[3] = {f, dfdx, dfdy};
  const int kinds[3] = {K_ID, K_IKX, K_IKY};
  const cx<T>* X1 = spectral_deriv<T>(c, fs, 1, nlev, outs, kinds, 3, keep_slot);
  if (spec) *spec = keep_slot >= 0 ? X1 : nullptr;
}

// filt_da on u, v, w in one pipeline (main.f90:231-233)
template <typename T>
void op_filt_da3(les_ctx* c, T* u, T* v, T* w, T* const gr[9], int nlev, int keep_slot, const cx<T>** spec) {
  OpTag _tag(c, "filt_da");
  const T* fs[3] = {u, v, w};
  T* outs[9] = {u, v, w, gr[0], gr[3], gr[6], gr[1], gr[4], gr[7]};
  const int kinds[3] = {K_ID, K_IKX, K_IKY};
  const cx<T>* X1 = spectral_deriv<T>(c, fs, 3, nlev, outs, kinds, 3, keep_slot);
  if (spec) *spec = keep_slot >= 0 ? X1 : nullptr;
}

template <typename T>
void op_ddxy(les_ctx* c, T* dfdx, T* dfdy, const T* f, int nlev) {
  OpTag _tag(c, "ddxy");
  T* outs[3] = {nullptr, nullptr, nullptr};
  int kinds[3] = {0, 0, 0};
  int n = 0;
  if (dfdx) { outs[n] = dfdx; kinds[n++] = K_IKX; }
  if (dfdy) { outs[n] = dfdy; kinds[n++] = K_IKY; }
  const T* fs[1] = {f};
  if (n) spectral_deriv<T>(c, fs, 1, nlev, outs, kinds, n);
}

// test_filtermodule.f90:13-30 applied to `nplanes` (ld,ny) planes in place
template <typename T>
void op_test_filter(les_ctx* c, T* planes, int nplanes, const T* G, int keep) {
  OpTag _tag(c, "test_filter");
  const Grid& g = c->g;
  if (keep <= 0 || keep > g.nxh) keep = g.nxh;
  const long rows = (long)nplanes * g.ny;
  const size_t xs = (size_t)rows * keep * sizeof(cx<T>);
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, xs);
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, xs);
  LoadFields<T> L{{planes, nullptr, nullptr, nullptr}, nplanes, g.ny, g.ld, T(1)};
  launch_row_fwd<T>(c, L, X0, keep, rows, g.nx, keep, g.ny);
  ColArgs a = col_args(g.ny, g.ny, keep, keep, (long)g.ny * keep);
  a.lhG = g.lh;
  col_out(a, 0, K_G, keep, keep, (long)g.ny * keep);
  launch_col<T>(c, X0, nullptr, X1, nullptr, nullptr, a, nplanes, keep, G);
  StoreFields<T> S{{planes, nullptr, nullptr, nullptr}, nplanes, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X1, keep, keep, S, rows, g.nx, g.ld, g.ny);
}

// test_filtermodule.f90:13-30 on the first level of two fields, filtered
// into two (ld,ny) planes of dst: the same transforms as op_test_filter on
// copies of those levels, without the copies
template <typename T>
void op_test_filter_level0(les_ctx* c, const T* a, const T* b, T* dst, const T* G, int keep) {
  OpTag _tag(c, "test_filter");
  const Grid& g = c->g;
  if (keep <= 0 || keep > g.nxh) keep = g.nxh;
  const long rows = 2L * g.ny;
  const size_t xs = (size_t)rows * keep * sizeof(cx<T>);
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, xs);
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, xs);
  LoadFields<T> L{{a, b, nullptr, nullptr}, 1, g.ny, g.ld, T(1)};
  launch_row_fwd<T>(c, L, X0, keep, rows, g.nx, keep, g.ny);
  ColArgs ca = col_args(g.ny, g.ny, keep, keep, (long)g.ny * keep);
  ca.lhG = g.lh;
  col_out(ca, 0, K_G, keep, keep, (long)g.ny * keep);
  launch_col<T>(c, X0, nullptr, X1, nullptr, nullptr, ca, 2, keep, G);
  StoreFields<T> S{{dst, nullptr, nullptr, nullptr}, 2, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X1, keep, keep, S, rows, g.nx, g.ld, g.ny);
}

// dealias1.f90: u (ld,ny,nlev) -> u_big (ld_big,ny2,nlev); u untouched
template <typename T>
void op_dealias1(les_ctx* c, const T* u, T* u_big, int nlev) {
  OpTag _tag(c, "dealias1");
  const Grid& g = c->g;
  const long rows = (long)nlev * g.ny, rows2 = (long)nlev * g.ny2;
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, (size_t)rows * g.nxh * sizeof(cx<T>));
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, (size_t)rows2 * g.nxh * sizeof(cx<T>));
  LoadFields<T> L{{u, nullptr, nullptr, nullptr}, nlev, g.ny, g.ld, T(1.0 / ((double)g.nx * g.ny))};
  launch_row_fwd<T>(c, L, X0, g.nxh, rows, g.nx, g.nxh, g.ny);
  ColArgs a = col_args(g.ny, g.ny2, g.nxh, g.nxh, (long)g.ny * g.nxh);
  col_out(a, 0, K_ID, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
  launch_col<T>(c, X0, nullptr, X1, nullptr, nullptr, a, nlev, g.nxh);
  const int ldb = 2 * (g.nx2 / 2 + 1);
  StoreFields<T> S{{u_big, nullptr, nullptr, nullptr}, nlev, g.ny2, ldb, g.nx2};
  launch_row_inv<T>(c, X1, g.nxh, g.nxh, S, rows2, g.nx2, ldb, g.ny2);
}

// dealias2.f90: u_big (ld_big,ny2,nlev) -> u (ld,ny,nlev).  The reference
// leaves the scaled spectrum of u_big behind in u_big; here u_big is read only.
template <typename T>
void op_dealias2(les_ctx* c, T* u, T* u_big, int nlev) {
  OpTag _tag(c, "dealias2");
  const Grid& g = c->g;
  const long rows = (long)nlev * g.ny, rows2 = (long)nlev * g.ny2;
  const int ldb = 2 * (g.nx2 / 2 + 1);
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, (size_t)rows2 * g.nxh * sizeof(cx<T>));
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, (size_t)rows * g.nxh * sizeof(cx<T>));
  LoadFields<T> L{{u_big, nullptr, nullptr, nullptr}, nlev, g.ny2, ldb, T(1.0 / ((double)g.nx2 * g.ny2))};
  launch_row_fwd<T>(c, L, X0, g.nxh, rows2, g.nx2, g.nxh, g.ny2);
  ColArgs a = col_args(g.ny2, g.ny, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
  col_out(a, 0, K_ID, g.nxh, g.nxh, (long)g.ny * g.nxh);
  launch_col<T>(c, X0, nullptr, X1, nullptr, nullptr, a, nlev, g.nxh);
  StoreFields<T> S{{u, nullptr, nullptr, nullptr}, nlev, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X1, g.nxh, g.nxh, S, rows, g.nx, g.ld, g.ny);
}

// padd.f90 / unpadd.f90 on complex (lh,ny) <-> (lh_big,ny2) planes
template <typename T>
__global__ void k_padd(cx<T>* big, const cx<T>* sm, int nx, int ny, int ny2, int lh, int lhb, int nplanes) {
  const long n = (long)lhb * ny2 * nplanes;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long pl = i / ((long)lhb * ny2);
    const int r = (int)(i - pl * lhb * ny2), yb = r / lhb, x = r - yb * lhb;
    const int ys = map_row(yb, ny, ny2);
    cx<T> v = {T(0), T(0)};
    if (ys >= 0 && x < nx / 2) v = sm[pl * lh * ny + (long)ys * lh + x];
    big[i] = v;
  }
}
template <typename T>
__global__ void k_unpadd(cx<T>* sm, const cx<T>* big, int nx, int ny, int ny2, int lh, int lhb, int nplanes) {
  const long n = (long)lh * ny * nplanes;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long pl = i / ((long)lh * ny);
    const int r = (int)(i - pl * lh * ny), ys = r / lh, x = r - ys * lh;
    const int yb = map_row(ys, ny2, ny);
    cx<T> v = {T(0), T(0)};
    if (yb >= 0 && x < nx / 2) v = big[pl * lhb * ny2 + (long)yb * lhb + x];
    sm[i] = v;
  }
}
template <typename T>
void op_padd(les_ctx* c, T* big_c, const T* small_c, int nplanes) {
  const Grid& g = c->g;
  const int lhb = g.nx2 / 2 + 1;
  hipLaunchKernelGGL((k_padd<T>), grid1((size_t)lhb * g.ny2 * nplanes), dim3(256), 0, c->stream, (cx<T>*)big_c,
                     (const cx<T>*)small_c, g.nx, g.ny, g.ny2, g.lh, lhb, nplanes);
  LCHK(c, hipGetLastError());
}
template <typename T>
void op_unpadd(les_ctx* c, T* small_c, const T* big_c, int nplanes) {
  const Grid& g = c->g;
  const int lhb = g.nx2 / 2 + 1;
  hipLaunchKernelGGL((k_unpadd<T>), grid1((size_t)g.lh * g.ny * nplanes), dim3(256), 0, c->stream, (cx<T>*)small_c,
                     (const cx<T>*)big_c, g.nx, g.ny, g.ny2, g.lh, lhb, nplanes);
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// ddz_uv.f90 / ddz_w.f90
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_ddz(T* __restrict__ dfdz, const T* __restrict__ f, int nx, int ld, int ny, int nz, T cst,
                      int wnode) {
  const long pl = (long)ld * ny, n = pl * nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    if (k == nz - 1) {
      dfdz[i] = T(0);
    } else if (x < nx) {
      if (wnode)
        dfdz[i] = cst * (f[i + pl] - f[i]);  // ddz_w: k = 1..nz-1
      else if (k >= 1)
        dfdz[i] = cst * (f[i] - f[i - pl]);  // ddz_uv: k = 2..nz-1
    }
  }
}
// the three vertical derivatives of a step in one pass: du/dz and dv/dz at
// w nodes (ddz_uv), dw/dz at uvp nodes (ddz_w); TPT consecutive points per
// thread so that each wave keeps several loads of every field in flight
template <typename T, int TPT>
__global__ void k_ddz3(T* __restrict__ dudz, T* __restrict__ dvdz, T* __restrict__ dwdz, const T* __restrict__ u,
                       const T* __restrict__ v, const T* __restrict__ w, int nx, int ld, int ny, int nz, T cst) {
  const unsigned pl = (unsigned)ld * ny, n = pl * nz;
  const unsigned i0 = (blockIdx.x * blockDim.x) * TPT + threadIdx.x;
#pragma unroll
  for (int t = 0; t < TPT; ++t) {
    const unsigned i = i0 + t * blockDim.x;
    if (i >= n) break;
    const unsigned k = i / pl;
    const int x = (int)((i - k * pl) % (unsigned)ld);
    if (k == (unsigned)nz - 1) {
      dudz[i] = T(0);
      dvdz[i] = T(0);
      dwdz[i] = T(0);
    } else if (x < nx) {
      dwdz[i] = cst * (w[i + pl] - w[i]);
      if (k >= 1) {
        dudz[i] = cst * (u[i] - u[i - pl]);
        dvdz[i] = cst * (v[i] - v[i - pl]);
      }
    }
  }
}
template <typename T>
void op_ddz3(les_ctx* c, T* dudz, T* dvdz, T* dwdz, const T* u, const T* v, const T* w) {
  const Grid& g = c->g;
  constexpr int TPT = 4;
  const size_t n = (size_t)g.ld * g.ny * g.nz;
  {
    ProfScope _ps(c, "ddz", 6.0 * n * sizeof(T));
    hipLaunchKernelGGL((k_ddz3<T, TPT>), dim3((unsigned)((n + 256 * TPT - 1) / (256 * TPT))), dim3(256), 0,
                       c->stream, dudz, dvdz, dwdz, u, v, w, g.nx, g.ld, g.ny, g.nz, T(1.0 / g.dz));
  }
  LCHK(c, hipGetLastError());
}
template <typename T>
void op_ddz_uv(les_ctx* c, T* dfdz, const T* f) {
  const Grid& g = c->g;
  {
    ProfScope _ps(c, "ddz", 2.0 * g.ld * g.ny * g.nz * sizeof(T));
    hipLaunchKernelGGL((k_ddz<T>), grid1((size_t)g.ld * g.ny * g.nz), dim3(256), 0, c->stream, dfdz, f, g.nx, g.ld,
                       g.ny, g.nz, T(1.0 / g.dz), 0);
  }
  LCHK(c, hipGetLastError());
}
template <typename T>
void op_ddz_w(les_ctx* c, T* dfdz, const T* f) {
  const Grid& g = c->g;
  {
    ProfScope _ps(c, "ddz", 2.0 * g.ld * g.ny * g.nz * sizeof(T));
    hipLaunchKernelGGL((k_ddz<T>), grid1((size_t)g.ld * g.ny * g.nz), dim3(256), 0, c->stream, dfdz, f, g.nx, g.ld,
                       g.ny, g.nz, T(1.0 / g.dz), 1);
  }
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// divstress_uv.f90 / divstress_w.f90:  d/dx tx + d/dy ty (one spectral
// pass) + d/dz tz fused into the c2r epilogue
// ------------------------------------------------------------------------
template <typename T>
struct StoreDiv {
  T* divt;
  const T* tz;
  int ny, ld, nx, nz, wnode;
  T rdz;
  static constexpr double writes() { return 2.0; }  // divt + tz read
  struct Pre {
    T a, b;  // the two tz values of the vertical difference
  };
  __device__ Pre pre(int plane, int y, int x) const {
    const int k = plane;
    const long i = ((long)plane * ny + y) * ld + x;
    const long pl = (long)ld * ny;
    Pre p{T(0), T(0)};
    if (x >= nx || (!wnode && k == nz - 1)) return p;
    if (wnode) {
      if (k > 0 && k < nz - 1) p = {tz[i], tz[i - pl]};
    } else {
      p = {tz[i + pl], tz[i]};
    }
    return p;
  }
  __device__ void fin(int plane, int y, int x, T v, const Pre& p) const {
    const int k = plane;
    const long i = ((long)plane * ny + y) * ld + x;
    if (!wnode && k == nz - 1) { divt[i] = T(BOGUS); return; }
    if (x >= nx) { divt[i] = T(0); return; }
    T d;
    if (wnode) {
      if (k == 0) d = v;                                            // divstress_w.f90:107-115
      else if (k < nz - 1) d = v + rdz * (p.a - p.b);               // ddz_uv
      else d = v + T(0);                                            // ddz_uv: dfdz(nz) = 0
    } else {
      d = v + rdz * (p.a - p.b);                                    // ddz_w
    }
    divt[i] = d;
  }
  __device__ void operator()(int plane, int y, int x, T v) const { fin(plane, y, x, v, pre(plane, y, x)); }
};

// the three divergences of the step in one pipeline (main.f90:294-297):
// plane f*nz + k of output f, with f's vertical stress and node type
template <typename T>
struct StoreDiv3 {
  StoreDiv<T> d[3];
  int nz;
  static constexpr double writes() { return 2.0; }
  using Pre = typename StoreDiv<T>::Pre;
  __device__ Pre pre(int plane, int y, int x) const {
    const int f = plane / nz;
    return d[f].pre(plane - f * nz, y, x);
  }
  __device__ void fin(int plane, int y, int x, T v, const Pre& p) const {
    const int f = plane / nz;
    d[f].fin(plane - f * nz, y, x, v, p);
  }
  __device__ void operator()(int plane, int y, int x, T v) const { fin(plane, y, x, v, pre(plane, y, x)); }
};

// divstress_uv (divtx, divty) and divstress_w (divtz) of the step: one
// forward row pass over the six horizontal-derivative inputs, one column
// pass and one inverse row pass for the three outputs
template <typename T>
void op_divstress3(les_ctx* c, T* const divt[3], T* const tau[6]) {
  OpTag _tag(c, "divstress");
  const Grid& g = c->g;
  const int nlev = g.nz;
  const long rows = (long)nlev * g.ny;
  const size_t xs = (size_t)rows * g.nxh * sizeof(cx<T>);
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, 6 * xs);
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, 3 * xs);
  // tau: txx, txy, txz, tyy, tyz, tzz; d/dx of (txx, txy, txz), d/dy of (txy, tyy, tyz)
  LoadFieldsN<T, 9> L{};
  const T* in[6] = {tau[0], tau[1], tau[2], tau[1], tau[3], tau[4]};
  for (int q = 0; q < 6; ++q) L.p[q] = in[q];
  L.ppf = nlev; L.ny = g.ny; L.ld = g.ld; L.scale = T(1.0 / ((double)g.nx * g.ny));
  const int tl = xs_layout(c, g.nxh);
  launch_row_fwd<T>(c, L, X0, g.nxh, 6 * rows, g.nx, g.nxh, g.ny, tl);
  ColArgs a = col_args(g.ny, g.ny, g.nxh, g.nxh, (long)g.ny * g.nxh);
  a.in2_plane = a.in_plane;
  col_out(a, 0, K_ID, g.nxh, g.nxh, (long)g.ny * g.nxh);
  a.in_tl = a.out_tl[0] = tl;
  launch_col<T>(c, X0, X0 + (size_t)3 * rows * g.nxh, X1, nullptr, nullptr, a, 3L * nlev, g.nxh);
  StoreDiv3<T> S{{StoreDiv<T>{divt[0], tau[2], g.ny, g.ld, g.nx, g.nz, 0, T(1.0 / g.dz)},
                  StoreDiv<T>{divt[1], tau[4], g.ny, g.ld, g.nx, g.nz, 0, T(1.0 / g.dz)},
                  StoreDiv<T>{divt[2], tau[5], g.ny, g.ld, g.nx, g.nz, 1, T(1.0 / g.dz)}},
                 nlev};
  launch_row_inv<T>(c, X1, g.nxh, g.nxh, S, 3 * rows, g.nx, g.ld, g.ny, nullptr, T(1), tl);
}

template <typename T>
void op_divstress(les_ctx* c, T* divt, const T* tx, const T* ty, const T* tz, int wnode) {
  OpTag _tag(c, "divstress");
  const Grid& g = c->g;
  const int nlev = g.nz;
  const long rows = (long)nlev * g.ny;
  const size_t xs = (size_t)rows * g.nxh * sizeof(cx<T>);
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, 2 * xs);
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, xs);
  LoadFields<T> L{{tx, ty, nullptr, nullptr}, nlev, g.ny, g.ld, T(1.0 / ((double)g.nx * g.ny))};
  const int tl = xs_layout(c, g.nxh);
  launch_row_fwd<T>(c, L, X0, g.nxh, 2 * rows, g.nx, g.nxh, g.ny, tl);
  ColArgs a = col_args(g.ny, g.ny, g.nxh, g.nxh, (long)g.ny * g.nxh);
  a.in2_plane = a.in_plane;
  col_out(a, 0, K_ID, g.nxh, g.nxh, (long)g.ny * g.nxh);
  a.in_tl = a.out_tl[0] = tl;
  launch_col<T>(c, X0, X0 + (size_t)rows * g.nxh, X1, nullptr, nullptr, a, nlev, g.nxh);
  StoreDiv<T> S{divt, tz, g.ny, g.ld, g.nx, g.nz, wnode, T(1.0 / g.dz)};
  launch_row_inv<T>(c, X1, g.nxh, g.nxh, S, rows, g.nx, g.ld, g.ny, nullptr, T(1), tl);
}

// ------------------------------------------------------------------------
// convec.f90: -(u x omega) with 3/2 dealiasing.  Fields never exist on the
// padded grid in physical form except inside `U6` (one pass); vorticity is
// formed in the forward-transform loader, the staggered products in the
// big-grid forward-transform loader.
// ------------------------------------------------------------------------
template <typename T>
struct StoreConvec {  // 3 outputs, levels 0..nz-2 ; level nz-1 BOGUS handled apart
  T *ox, *oy, *oz;
  int ppf, ny, ld, nx;
  static constexpr double writes() { return 1.0; }
  __device__ void operator()(int plane, int y, int x, T v) const {
    const int f = plane / ppf, r = plane - f * ppf;
    T* p = f == 0 ? ox : (f == 1 ? oy : oz);
    p[((long)r * ny + y) * ld + x] = x < nx ? v : T(0);
  }
};

// convec.f90:43-233 up to the small-grid x-spectral products (3 fields,
// levels 1..nz-1); returns the scratch holding them.
template <typename T>
static cx<T>* convec_spectral(les_ctx* c, const T* u, const T* v, const T* w, const T* dudy, const T* dudz,
                              const T* dvdx, const T* dvdz, const T* dwdx, const T* dwdy,
                              const SpecIn<T>* si = nullptr) {
  const Grid& g = c->g;
  const int nz = g.nz;
  // pad in y: 6*nz planes
  const long rowsB = 6L * nz * g.ny2;
  cx<T>* XB = (cx<T>*)scratch(c, S_XS1, (size_t)rowsB * g.nxh * sizeof(cx<T>));
  const int tl = xs_layout(c, g.nxh);
  bool padded = false;
  if (si && si->uvw) {
    // the inputs from filt_da's kept spectra; the wall planes of dudz, dvdz
    // (wallstress) by one small forward transform
    cx<T>* W = (cx<T>*)scratch(c, S_FWALL, (size_t)6 * g.ny * g.nxh * sizeof(cx<T>));
    LoadPlanes6<T> LW{{dudz, dvdz, dudz, dudz, dudz, dudz}, g.ld, T(1.0 / g.nx)};
    launch_row_fwd<T>(c, LW, W, g.nxh, 2L * g.ny, g.nx, g.nxh, g.ny, tl);
    ConvecSpecIn<T, 0> fn{si->uvw, {nullptr, nullptr}, W, nz, g.ny, g.nxh, T(1.0 / g.ny), T(1.0 / g.dz)};
    padded = launch_col_pad_fn<T>(c, fn, XB, 6L * nz, g.nxh, g.nxh, (long)g.ny2 * g.nxh, 1.5, tl);
  }
  if (!padded) {
    const long rows6 = 6L * nz * g.ny;
    cx<T>* XS = (cx<T>*)scratch(c, S_XS0, (size_t)rows6 * g.nxh * sizeof(cx<T>));
    LoadConvec<T> L{u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, g.ny, nz, g.ld, T(1.0 / ((double)g.nx * g.ny))};
    launch_row_fwd<T>(c, L, XS, g.nxh, rows6, g.nx, g.nxh, g.ny, tl);
    ColArgs a = col_args(g.ny, g.ny2, g.nxh, g.nxh, (long)g.ny * g.nxh);
    col_out(a, 0, K_ID, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
    a.in_tl = tl;
    launch_col<T>(c, XS, nullptr, XB, nullptr, nullptr, a, 6L * nz, g.nxh);
  }
  // padded-grid products without materialising the padded fields
  const long rowsP = 3L * (nz - 1) * g.ny2;
  cx<T>* XP = (cx<T>*)scratch(c, S_XS2, (size_t)rowsP * g.nxh * sizeof(cx<T>));
  launch_pad_products<T, PolConvec<T>, 8, 3>(c, XB, XP, PolConvec<T>{T(1.0 / ((double)g.nx2 * g.ny2))},
                                             "convec_big");
  // truncate in y
  const long rows3 = 3L * (nz - 1) * g.ny;
  cx<T>* X3 = (cx<T>*)scratch(c, S_XS3, (size_t)rows3 * g.nxh * sizeof(cx<T>));
  ColArgs b = col_args(g.ny2, g.ny, g.nxh, g.nxh, (long)g.ny2 * g.nxh);
  col_out(b, 0, K_ID, g.nxh, g.nxh, (long)g.ny * g.nxh);
  launch_col<T>(c, XP, nullptr, X3, nullptr, nullptr, b, 3L * (nz - 1), g.nxh);
  return X3;
}

template <typename T>
void op_convec(les_ctx* c, const T* u, const T* v, const T* w, const T* dudy, const T* dudz, const T* dvdx,
               const T* dvdz, const T* dwdx, const T* dwdy, T* ox, T* oy, T* oz) {
  OpTag _tag(c, "convec");
  const Grid& g = c->g;
  const int nz = g.nz;
  cx<T>* X3 = convec_spectral<T>(c, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy);
  StoreConvec<T> SC{ox, oy, oz, nz - 1, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X3, g.nxh, g.nxh, SC, 3L * (nz - 1) * g.ny, g.nx, g.ld, g.ny);
  const size_t pl = (size_t)g.ld * g.ny, top = pl * (nz - 1);
  op_fill<T>(c, ox + top, pl, T(BOGUS));
  op_fill<T>(c, oy + top, pl, T(BOGUS));
  op_fill<T>(c, oz + top, pl, T(BOGUS));
}

template <typename T>
void op_convec_ab2(les_ctx* c, T* u, T* v, T* w, const T* dudy, const T* dudz, const T* dvdx, const T* dvdz,
                   const T* dwdx, const T* dwdy, T* RHSx, T* RHSy, T* RHSz, T* RHSx_f, T* RHSy_f, T* RHSz_f,
                   const T* divtx, const T* divty, const T* divtz, int euler_first, const T* beta,
                   const ForcingArgs<T>* fo, const SpecIn<T>* si) {
  OpTag _tag(c, "convec");
  const Grid& g = c->g;
  const int nz = g.nz;
  cx<T>* X3 = convec_spectral<T>(c, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, si);
  StoreConvecAB2<T> SC{{RHSx, RHSy, RHSz}, {RHSx_f, RHSy_f, RHSz_f}, {divtx, divty, divtz}, {u, v, w},
                       nz - 1, g.ny, g.ld, g.nx, (T)g.dt, (T)g.force, euler_first, beta};
  if (fo) SC.fo = *fo;
  launch_row_inv<T>(c, X3, g.nxh, g.nxh, SC, 3L * (nz - 1) * g.ny, g.nx, g.ld, g.ny);
  const size_t pl = (size_t)g.ld * g.ny, top = pl * (nz - 1);
  T* tops[9] = {RHSx + top, RHSy + top, RHSz + top, u + top, v + top, w + top, RHSx_f + top, RHSy_f + top,
                RHSz_f + top};
  op_fill_n<T>(c, tops, euler_first ? 9 : 6, pl, T(BOGUS));
}

// ------------------------------------------------------------------------
// wallstress.f90 ('wall'): 4-Delta test filter of u, v at k=1 and the log
// law, FP64 throughout
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_wall(const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ u1f,
                       const T* __restrict__ v1f, T* dudz, T* dvdz, T* txz, T* tyz, const double* psi_m,
                       const double* psi_m0, const double* phi_m, double* ustar_avg, int nx, int ny, int ld,
                       double dz, double z0) {
  const int n = nx * ny;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int y = j / nx, x = j - y * nx;
    const long i = x + (long)ld * y;
    const double u1 = (double)u1f[i], v1 = (double)v1f[i];
    const double denom = log(0.5 * dz / z0) - psi_m[j] + psi_m0[j];
    const double u_avg = sqrt(u1 * u1 + v1 * v1);
    const double us = u_avg * VONK / denom;
    ustar_avg[j] = us;
    const double cst = -(us * us) / u_avg;
    txz[i] = (T)(cst * u1);
    tyz[i] = (T)(cst * v1);
    const double du = us / (0.5 * dz * VONK) * u1 / u_avg * phi_m[j];
    const double dv = us / (0.5 * dz * VONK) * v1 / u_avg * phi_m[j];
    dudz[i] = u[i] == T(0) ? T(0) : (T)du;
    dvdz[i] = v[i] == T(0) ? T(0) : (T)dv;
  }
}

template <typename T>
void op_wallstress(les_ctx* c, const T* u, const T* v, T* dudz, T* dvdz, T* txz, T* tyz, const double* psi_m,
                   const double* psi_m0, const double* phi_m, double* ustar_avg) {
  OpTag _tag(c, "wallstress");
  const Grid& g = c->g;
  const size_t pl = (size_t)g.ld * g.ny;
  T* planes = (T*)scratch(c, S_PLANE, 2 * pl * sizeof(T));
  op_test_filter_level0<T>(c, u, v, planes, tables<T>(c).G4, c->keep4);
  {
    ProfScope _ps(c, "wall", 8.0 * g.nx * g.ny * sizeof(T));
    hipLaunchKernelGGL((k_wall<T>), grid1((size_t)g.nx * g.ny), dim3(256), 0, c->stream, u, v, planes, planes + pl,
                       dudz, dvdz, txz, tyz, psi_m, psi_m0, phi_m, ustar_avg, g.nx, g.ny, g.ld, g.dz, g.z0);
  }
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// sgs_stag.f90: calc_Sij (:371-498), Nu_t and tau (:176-364)
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_calc_sij(const T* __restrict__ ux_, const T* __restrict__ uy_, const T* __restrict__ uz_,
                           const T* __restrict__ vx_, const T* __restrict__ vy_, const T* __restrict__ vz_,
                           const T* __restrict__ wx_, const T* __restrict__ wy_, const T* __restrict__ wz_,
                           T* S11, T* S12, T* S13, T* S22, T* S23, T* S33, int nx, int ld, int ny, int nz,
                           const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w, T* u_lag,
                           T* v_lag, T* w_lag, T* u_lag_s, T* v_lag_s, T* w_lag_s) {
  const long pl = (long)ld * ny, n = pl * nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    if (u_lag) {  // sgs_stag.f90:62-64, whole arrays
      u_lag[i] = u_lag[i] + u[i];
      v_lag[i] = v_lag[i] + v[i];
      w_lag[i] = w_lag[i] + w[i];
    }
    if (u_lag_s) {  // sgs_stag.f90:66-68
      u_lag_s[i] = u_lag_s[i] + u[i];
      v_lag_s[i] = v_lag_s[i] + v[i];
      w_lag_s[i] = w_lag_s[i] + w[i];
    }
    T s11 = T(0), s12 = T(0), s13 = T(0), s22 = T(0), s23 = T(0), s33 = T(0);
    if (x < nx) {
      T ux, uy, uz, vx, vy, vz, wx, wy, wz;
      if (k == 0) {
        ux = ux_[i]; uy = uy_[i]; uz = uz_[i];
        vx = vx_[i]; vy = vy_[i]; vz = vz_[i];
        wx = T(0.5) * (wx_[i] + wx_[i + pl]);
        wy = T(0.5) * (wy_[i] + wy_[i + pl]);
        wz = wz_[i];
      } else {
        const long m = i - pl;
        ux = T(0.5) * (ux_[i] + ux_[m]);
        uy = T(0.5) * (uy_[i] + uy_[m]);
        uz = uz_[i];
        vx = T(0.5) * (vx_[i] + vx_[m]);
        vy = T(0.5) * (vy_[i] + vy_[m]);
        vz = vz_[i];
        wx = wx_[i];
        wy = wy_[i];
        wz = T(0.5) * (wz_[i] + wz_[m]);
      }
      s11 = ux; s12 = T(0.5) * (uy + vx); s13 = T(0.5) * (uz + wx);
      s22 = vy; s23 = T(0.5) * (vz + wy); s33 = wz;
    }
    S11[i] = s11; S12[i] = s12; S13[i] = s13; S22[i] = s22; S23[i] = s23; S33[i] = s33;
  }
}

template <typename T>
void op_calc_sij(les_ctx* c, const T* const gr[9], T* const S[6]) {
  const Grid& g = c->g;
  hipLaunchKernelGGL((k_calc_sij<T>), grid1((size_t)g.ld * g.ny * g.nz), dim3(256), 0, c->stream, gr[0], gr[1],
                     gr[2], gr[3], gr[4], gr[5], gr[6], gr[7], gr[8], S[0], S[1], S[2], S[3], S[4], S[5], g.nx,
                     g.ld, g.ny, g.nz, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (T*)nullptr,
                     (T*)nullptr, (T*)nullptr, (T*)nullptr, (T*)nullptr, (T*)nullptr);
  LCHK(c, hipGetLastError());
}


template <typename T>
__global__ void k_sgs_tau(const T* __restrict__ S11, const T* __restrict__ S12, const T* __restrict__ S13,
                          const T* __restrict__ S22, const T* __restrict__ S23, const T* __restrict__ S33,
                          const T* __restrict__ Cs, T cs_const, int use_const, T* Cs_out, T* Nu_t, T* txx, T* txy,
                          T* txz, T* tyy, T* tyz, T* tzz, int nx, int ld, int ny, int nz, T l2_const,
                          const double* __restrict__ l2z, T* Pr_t, T Pr) {
  const long pl = (long)ld * ny, n = pl * nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    if (use_const) Cs_out[i] = cs_const;  // sgs_stag.f90:166-173, whole arrays
    if (use_const && Pr_t) Pr_t[i] = Pr;
    if (k == nz - 1) {                    // sgs_stag.f90:356-364
      txx[i] = T(BOGUS); txy[i] = T(BOGUS); tyy[i] = T(BOGUS); tzz[i] = T(BOGUS);
      txz[i] = T(0); tyz[i] = T(0);
    }
    if (x >= nx) continue;
    const T l2 = l2z ? (T)l2z[k] : l2_const;  // model 1: l(jz)^2 (sgs_stag.f90:84-94), else delta^2
    const T c0 = use_const ? cs_const : Cs[i];
    const T nu0 = smag(S11[i], S12[i], S13[i], S22[i], S23[i], S33[i]) * c0 * l2;
    Nu_t[i] = nu0;
    if (k == nz - 1) continue;
    if (k == 0) {  // wall level, sgs_stag.f90:223-241 (txz, tyz(1) kept from wallstress)
      const T cst = nu0;
      txx[i] = T(-2) * (cst + T(0)) * S11[i];
      txy[i] = T(-2) * (cst + T(0)) * S12[i];
      tyy[i] = T(-2) * (cst + T(0)) * S22[i];
      tzz[i] = T(-2) * (cst + T(0)) * S33[i];
    } else {
      const long p = i + pl;
      const T c1 = use_const ? cs_const : Cs[p];
      const T l2p = l2z ? (T)l2z[k + 1] : l2_const;
      const T nu1 = smag(S11[p], S12[p], S13[p], S22[p], S23[p], S33[p]) * c1 * l2p;
      const T cst = T(0.5) * (nu0 + nu1);
      txx[i] = T(-2) * (cst + T(0)) * T(0.5) * (S11[i] + S11[p]);
      txy[i] = T(-2) * (cst + T(0)) * T(0.5) * (S12[i] + S12[p]);
      tyy[i] = T(-2) * (cst + T(0)) * T(0.5) * (S22[i] + S22[p]);
      tzz[i] = T(-2) * (cst + T(0)) * T(0.5) * (S33[i] + S33[p]);
      txz[i] = T(-2) * (nu0 + T(0)) * S13[i];
      tyz[i] = T(-2) * (nu0 + T(0)) * S23[i];
    }
  }
}

// k_calc_sij and k_sgs_tau in one pass (k_sij_tau_col), for steps in which no
// dynamic model reads S_ij: S_ij at a level and the level above from the
// gradients (the expressions of k_calc_sij), then Nu_t and tau as k_sgs_tau,
// so S_ij never goes to memory
template <typename T>
__device__ __forceinline__ void sij_at(const T* const* gr, long i, long pl, int k, T (&s)[6]) {
  T ux, uy, uz, vx, vy, vz, wx, wy, wz;
  if (k == 0) {
    ux = gr[0][i]; uy = gr[1][i]; uz = gr[2][i];
    vx = gr[3][i]; vy = gr[4][i]; vz = gr[5][i];
    wx = T(0.5) * (gr[6][i] + gr[6][i + pl]);
    wy = T(0.5) * (gr[7][i] + gr[7][i + pl]);
    wz = gr[8][i];
  } else {
    const long m = i - pl;
    ux = T(0.5) * (gr[0][i] + gr[0][m]);
    uy = T(0.5) * (gr[1][i] + gr[1][m]);
    uz = gr[2][i];
    vx = T(0.5) * (gr[3][i] + gr[3][m]);
    vy = T(0.5) * (gr[4][i] + gr[4][m]);
    vz = gr[5][i];
    wx = gr[6][i];
    wy = gr[7][i];
    wz = T(0.5) * (gr[8][i] + gr[8][m]);
  }
  s[0] = ux; s[1] = T(0.5) * (uy + vx); s[2] = T(0.5) * (uz + wx);
  s[3] = vy; s[4] = T(0.5) * (vz + wy); s[5] = wz;
  // opaque from here on, as if read back from memory: the Nu_t and tau
  // arithmetic then contracts exactly as in k_sgs_tau (identical bits)
#pragma unroll
  for (int q = 0; q < 6; ++q) asm volatile("" : "+v"(s[q]));
}
template <typename T>
struct Grads9 {
  const T* p[9];
};
// One thread per column and run of L levels: S_ij of the level above is
// formed once and carried to the next level, and the five gradients the next
// level's averages need stay in registers (the expressions of sij_at, so the
// bits of k_calc_sij followed by k_sgs_tau)
template <typename T, int L>
__global__ void __launch_bounds__(256)
    k_sij_tau_col(Grads9<T> G, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
                  T* u_lag, T* v_lag, T* w_lag, T* u_lag_s, T* v_lag_s, T* w_lag_s, const T* __restrict__ Cs,
                  T cs_const, int use_const, T* Cs_out, T* Nu_t, T* txx, T* txy, T* txz, T* tyy, T* tyz, T* tzz,
                  int nx, int ld, int ny, int nz, T l2_const, const double* __restrict__ l2z, T* Pr_t, T Pr) {
  const long pl = (long)ld * ny;
  const long col = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (col >= pl) return;
  const int k0 = blockIdx.y * L, k1 = min(k0 + L, nz);
  const int x = (int)(col % ld);
  T sc[6], sn[6], g5[5];  // S_ij at k and k + 1; ux, uy, vx, vy, wz at k
  if (x < nx) {
    const long i0 = (long)k0 * pl + col;
    sij_at<T>(G.p, i0, pl, k0, sc);
    g5[0] = G.p[0][i0]; g5[1] = G.p[1][i0]; g5[2] = G.p[3][i0]; g5[3] = G.p[4][i0]; g5[4] = G.p[8][i0];
  }
#pragma unroll
  for (int kk = 0; kk < L; ++kk) {
    const int k = k0 + kk;
    if (k >= k1) break;
    const long i = (long)k * pl + col;
    if (u_lag) {  // sgs_stag.f90:62-64, whole arrays
      u_lag[i] = u_lag[i] + u[i];
      v_lag[i] = v_lag[i] + v[i];
      w_lag[i] = w_lag[i] + w[i];
    }
    if (u_lag_s) {  // sgs_stag.f90:66-68
      u_lag_s[i] = u_lag_s[i] + u[i];
      v_lag_s[i] = v_lag_s[i] + v[i];
      w_lag_s[i] = w_lag_s[i] + w[i];
    }
    if (use_const) Cs_out[i] = cs_const;  // sgs_stag.f90:166-173, whole arrays
    if (use_const && Pr_t) Pr_t[i] = Pr;
    if (k == nz - 1) {  // sgs_stag.f90:356-364
      txx[i] = T(BOGUS); txy[i] = T(BOGUS); tyy[i] = T(BOGUS); tzz[i] = T(BOGUS);
      txz[i] = T(0); tyz[i] = T(0);
    }
    if (x >= nx) continue;
    if (k + 1 < nz) {  // S_ij one level up, from level k + 1 and the level-k gradients held in g5
      const long p = i + pl;
      const T ux1 = G.p[0][p], uy1 = G.p[1][p], vx1 = G.p[3][p], vy1 = G.p[4][p], wz1 = G.p[8][p];
      const T ux = T(0.5) * (ux1 + g5[0]);
      const T uy = T(0.5) * (uy1 + g5[1]);
      const T uz = G.p[2][p];
      const T vx = T(0.5) * (vx1 + g5[2]);
      const T vy = T(0.5) * (vy1 + g5[3]);
      const T vz = G.p[5][p];
      const T wx = G.p[6][p];
      const T wy = G.p[7][p];
      const T wz = T(0.5) * (wz1 + g5[4]);
      sn[0] = ux; sn[1] = T(0.5) * (uy + vx); sn[2] = T(0.5) * (uz + wx);
      sn[3] = vy; sn[4] = T(0.5) * (vz + wy); sn[5] = wz;
#pragma unroll
      for (int q = 0; q < 6; ++q) asm volatile("" : "+v"(sn[q]));
      g5[0] = ux1; g5[1] = uy1; g5[2] = vx1; g5[3] = vy1; g5[4] = wz1;
    }
    const T l2 = l2z ? (T)l2z[k] : l2_const;
    const T c0 = use_const ? cs_const : Cs[i];
    const T nu0 = smag(sc[0], sc[1], sc[2], sc[3], sc[4], sc[5]) * c0 * l2;
    Nu_t[i] = nu0;
    if (k == nz - 1) continue;
    if (k == 0) {  // wall level, sgs_stag.f90:223-241 (txz, tyz(1) kept from wallstress)
      const T cst = nu0;
      txx[i] = T(-2) * (cst + T(0)) * sc[0];
      txy[i] = T(-2) * (cst + T(0)) * sc[1];
      tyy[i] = T(-2) * (cst + T(0)) * sc[3];
      tzz[i] = T(-2) * (cst + T(0)) * sc[5];
    } else {
      const long p = i + pl;
      const T c1 = use_const ? cs_const : Cs[p];
      const T l2p = l2z ? (T)l2z[k + 1] : l2_const;
      const T nu1 = smag(sn[0], sn[1], sn[2], sn[3], sn[4], sn[5]) * c1 * l2p;
      const T cst = T(0.5) * (nu0 + nu1);
      txx[i] = T(-2) * (cst + T(0)) * T(0.5) * (sc[0] + sn[0]);
      txy[i] = T(-2) * (cst + T(0)) * T(0.5) * (sc[1] + sn[1]);
      tyy[i] = T(-2) * (cst + T(0)) * T(0.5) * (sc[3] + sn[3]);
      tzz[i] = T(-2) * (cst + T(0)) * T(0.5) * (sc[5] + sn[5]);
      txz[i] = T(-2) * (nu0 + T(0)) * sc[2];
      tyz[i] = T(-2) * (nu0 + T(0)) * sc[4];
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) sc[q] = sn[q];
  }
}

// ------------------------------------------------------------------------
// interpolag_Sdep.f90: Lagrangian back-trajectory + trilinear interpolation
// of NF fields.  SIM: interpolag_Ssim.f90, whose two lowest levels keep the
// doubled displacement without the clamp to [-1, 1] (and level 1 its own
// 0.25 w_lag(2) value) where interpolag_Sdep.f90:264-269 clamps both to
// the clamped level-2 value.
// ------------------------------------------------------------------------
template <typename T>
struct In4 {
  const T* p[4];
};
template <typename T>
struct Out4 {
  T* p[4];
};

// in: the fields before the step (never the arrays written); out: the
// interpolated fields on the nx interior columns and, with pad_copy, the
// input's pad columns as they are (the reference leaves them untouched)
template <typename T, int NF = 4, bool SIM = false>
__global__ void k_interpolag(const T* __restrict__ ul, const T* __restrict__ vl, const T* __restrict__ wl, In4<T> in,
                             Out4<T> out, int pad_copy, int nx, int ld, int ny, int nz, T cs, T dt, T dx, T dy,
                             T dz) {
  const long pl = (long)ld * ny, n3 = pl * nz;
  const long n = (long)nx * ny * nz;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const unsigned pn = (unsigned)(nx * ny), qq = (unsigned)q;  // nx*ny*nz < 2^31
    const int k = (int)(qq / pn);
    const int r = (int)(qq - (unsigned)k * pn), y = r / nx, x = r - y * nx;
    const long i = (long)k * pl + (long)y * ld + x;
    // u_lag, v_lag to w-nodes (interpolag_Sdep.f90:228-241)
    T uu = ul[i] / cs, vv = vl[i] / cs;
    if (k >= 1) {
      uu = T(0.5) * (uu + ul[i - pl] / cs);
      vv = T(0.5) * (vv + vl[i - pl] / cs);
    }
    const T xp = -uu * dt * cs / dx;
    const T yp = -vv * dt * cs / dy;
    T zp;
    if (SIM && k <= 1) {  // interpolag_Ssim.f90: doubling only
      const long i2 = (long)pl + (long)y * ld + x;
      if (k == 1) {
        const T z2 = -(wl[i2] / cs) * dt * cs / dz;
        zp = z2 + (z2 < T(0) ? z2 : T(0));
      } else {
        const T z1 = -(T(0.25) * (wl[i2] / cs)) * dt * cs / dz;
        zp = z1 + (z1 > T(0) ? z1 : T(0));
      }
    } else if (k <= 1) {  // :243-295: level 1 = 0.25 w_lag(2); then wall doubling + clamps
      const long i2 = (long)pl + (long)y * ld + x;
      T z2 = -(wl[i2] / cs) * dt * cs / dz;
      z2 = z2 + (z2 < T(0) ? z2 : T(0));
      z2 = z2 < T(1) ? z2 : T(1);
      z2 = z2 > T(-1) ? z2 : T(-1);
      zp = z2;
    } else {
      zp = -(wl[i] / cs) * dt * cs / dz;
    }
    const int ax = signbit(xp) ? -1 : 1, ay = signbit(yp) ? -1 : 1, az = signbit(zp) ? -1 : 1;
    const int xa = (x + ax + nx) % nx, ya = (y + ay + ny) % ny;
    int za = k + az;
    za = za < 0 ? 0 : (za > nz - 1 ? nz - 1 : za);
    const T cxv = fabs(xp), cyv = fabs(yp), czv = fabs(zp);
    const T fx = T(1) - cxv, fy = T(1) - cyv, fz = T(1) - czv;
    const long i000 = i, i00z = (long)za * pl + (long)y * ld + x;
    const long i0y0 = (long)k * pl + (long)ya * ld + x, i0yz = (long)za * pl + (long)ya * ld + x;
    const long ix00 = (long)k * pl + (long)y * ld + xa, ix0z = (long)za * pl + (long)y * ld + xa;
    const long ixy0 = (long)k * pl + (long)ya * ld + xa, ixyz = (long)za * pl + (long)ya * ld + xa;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const T* G = in.p[f];
      out.p[f][i] = fx * fy * (G[i000] * fz + G[i00z] * czv) + fx * cyv * (G[i0y0] * fz + G[i0yz] * czv) +
                    cxv * fy * (G[ix00] * fz + G[ix0z] * czv) + cxv * cyv * (G[ixy0] * fz + G[ixyz] * czv);
    }
  }
  if (pad_copy) {
    const unsigned npad = (unsigned)(ld - nx), m = npad * (unsigned)ny * (unsigned)nz;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x) {
      const unsigned row = q / npad, xo = q - row * npad;  // row = k ny + y
      const long i = (long)row * ld + nx + xo;
#pragma unroll
      for (int f = 0; f < NF; ++f) out.p[f][i] = in.p[f][i];
    }
  }
  (void)n3;
}

// The same interpolation walking up the columns of a TX x TY tile through a
// run of KC levels (blockIdx.z): the fields of levels k - 1, k, k + 1 sit in
// an LDS ring of three level tiles with a one-point periodic halo in x and y,
// so every neighbour value comes from LDS and each field is read from memory
// about once (plus the halo and two levels per run) instead of once per
// neighbour; u_lag, v_lag of the level below come from registers.  The
// arithmetic and its order are those of k_interpolag (identical results).
template <typename T, int NF, bool SIM, int TX>
__global__ void __launch_bounds__(256)
    k_interpolag_walk(const T* __restrict__ ul, const T* __restrict__ vl, const T* __restrict__ wl, In4<T> in,
                      Out4<T> out, int pad_copy, int nx, int ld, int ny, int nz, int kc, T cs, T dt, T dx, T dy,
                      T dz) {
  constexpr int TY = 256 / TX, HX = TX + 2, HP = HX * (TY + 2);
  const int kb = blockIdx.z * kc, ke = min(kb + kc, nz);
  __shared__ T tile[3][NF][HP];
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, x = x0 + lx, y = y0 + ly;
  const long pl = (long)ld * ny;
  auto load_level = [&](int kk) {
    const int sl = kk % 3;
    for (int e = threadIdx.x; e < HP; e += 256) {
      const int hy = e / HX, hx = e - hy * HX;
      const int gx = (x0 + hx - 1 + nx) % nx, gy = (y0 + hy - 1 + ny) % ny;
      const long gi = (long)kk * pl + (long)gy * ld + gx;
#pragma unroll
      for (int f = 0; f < NF; ++f) tile[sl][f][e] = in.p[f][gi];
    }
  };
  if (kb >= 1) load_level(kb - 1);
  load_level(kb);
  if (kb + 1 < nz) load_level(kb + 1);
  const int c0 = (ly + 1) * HX + lx + 1;
  T ulm = T(0), vlm = T(0);  // u_lag, v_lag of this column one level down
  if (kb >= 1) {
    const long im = (long)(kb - 1) * pl + (long)y * ld + x;
    ulm = ul[im];
    vlm = vl[im];
  }
  for (int k = kb; k < ke; ++k) {
    __syncthreads();  // levels k - 1, k and k + 1 are in the ring
    const long i = (long)k * pl + (long)y * ld + x;
    const T ulk = ul[i], vlk = vl[i];
    T uu = ulk / cs, vv = vlk / cs;
    if (k >= 1) {
      uu = T(0.5) * (uu + ulm / cs);
      vv = T(0.5) * (vv + vlm / cs);
    }
    ulm = ulk;
    vlm = vlk;
    const T xp = -uu * dt * cs / dx;
    const T yp = -vv * dt * cs / dy;
    T zp;
    if (SIM && k <= 1) {
      const long i2 = (long)pl + (long)y * ld + x;
      if (k == 1) {
        const T z2 = -(wl[i2] / cs) * dt * cs / dz;
        zp = z2 + (z2 < T(0) ? z2 : T(0));
      } else {
        const T z1 = -(T(0.25) * (wl[i2] / cs)) * dt * cs / dz;
        zp = z1 + (z1 > T(0) ? z1 : T(0));
      }
    } else if (k <= 1) {
      const long i2 = (long)pl + (long)y * ld + x;
      T z2 = -(wl[i2] / cs) * dt * cs / dz;
      z2 = z2 + (z2 < T(0) ? z2 : T(0));
      z2 = z2 < T(1) ? z2 : T(1);
      z2 = z2 > T(-1) ? z2 : T(-1);
      zp = z2;
    } else {
      zp = -(wl[i] / cs) * dt * cs / dz;
    }
    const int ax = signbit(xp) ? -1 : 1, ay = signbit(yp) ? -1 : 1, az = signbit(zp) ? -1 : 1;
    int za = k + az;
    za = za < 0 ? 0 : (za > nz - 1 ? nz - 1 : za);
    const T cxv = fabs(xp), cyv = fabs(yp), czv = fabs(zp);
    const T fx = T(1) - cxv, fy = T(1) - cyv, fz = T(1) - czv;
    const int cy = c0 + ay * HX, cx = c0 + ax, cxy = cy + ax;
    const int sk = k % 3, sz = za % 3;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const T* G = tile[sk][f];
      const T* Z = tile[sz][f];
      out.p[f][i] = fx * fy * (G[c0] * fz + Z[c0] * czv) + fx * cyv * (G[cy] * fz + Z[cy] * czv) +
                    cxv * fy * (G[cx] * fz + Z[cx] * czv) + cxv * cyv * (G[cxy] * fz + Z[cxy] * czv);
    }
    if (pad_copy && x0 + TX == nx) {  // the pad columns of this tile's rows, as they are
      const int npad = ld - nx;
      for (int e = threadIdx.x; e < npad * TY; e += 256) {
        const int r = e / npad, xo = e - r * npad;
        const long ip = (long)k * pl + (long)(y0 + r) * ld + nx + xo;
#pragma unroll
        for (int f = 0; f < NF; ++f) out.p[f][ip] = in.p[f][ip];
      }
    }
    __syncthreads();  // level k - 1's slot is free
    if (k + 2 < nz && k + 2 <= ke) load_level(k + 2);
  }
}

// k_interpolag over the grid: the column walk where the tile divides the
// grid, the pointwise kernel elsewhere
template <typename T, int NF, bool SIM>
static void launch_interpolag(les_ctx* c, const T* ul, const T* vl, const T* wl, const In4<T>& in,
                              const Out4<T>& out, int pad_copy) {
  const Grid& g = c->g;
  const T cs = T(lag_count(c)), dt = T(g.dt), dx = T(g.dx), dy = T(g.dy), dz = T(g.dz);
  const int kc = 16, gz = (g.nz + kc - 1) / kc;  // runs of 16 levels: enough workgroups to fill the CUs
  if (g.nx % 64 == 0 && g.ny % 4 == 0)
    hipLaunchKernelGGL((k_interpolag_walk<T, NF, SIM, 64>), dim3(g.nx / 64, g.ny / 4, gz), dim3(256), 0, c->stream,
                       ul, vl, wl, in, out, pad_copy, g.nx, g.ld, g.ny, g.nz, kc, cs, dt, dx, dy, dz);
  else if (g.nx % 32 == 0 && g.ny % 8 == 0)
    hipLaunchKernelGGL((k_interpolag_walk<T, NF, SIM, 32>), dim3(g.nx / 32, g.ny / 8, gz), dim3(256), 0, c->stream,
                       ul, vl, wl, in, out, pad_copy, g.nx, g.ld, g.ny, g.nz, kc, cs, dt, dx, dy, dz);
  else
    hipLaunchKernelGGL((k_interpolag<T, NF, SIM>), grid1((size_t)g.nx * g.ny * g.nz), dim3(256), 0, c->stream, ul,
                       vl, wl, in, out, pad_copy, g.nx, g.ld, g.ny, g.nz, cs, dt, dx, dy, dz);
}

// interpolag_Sdep.f90:2-380 on (F_LM, F_MM, F_QN, F_NN) with u_lag; the same
// scheme serves interpolag_scalar_Sdep.f90:1-316 on G_* with u_lag_s
template <typename T>
void op_interpolag4(les_ctx* c, T* const F[4], T* ul, T* vl, T* wl) {
  const Grid& g = c->g;
  const size_t n3 = (size_t)g.ld * g.ny * g.nz;
  T* FF = (T*)scratch(c, S_FF, 4 * n3 * sizeof(T));
  for (int q = 0; q < 4; ++q) op_copy<T>(c, FF + q * n3, F[q], n3);
  {
    ProfScope _ps(c, "interpolag", (4.0 + 3 + 4) * n3 * sizeof(T));
    const In4<T> in{{FF, FF + n3, FF + 2 * n3, FF + 3 * n3}};
    const Out4<T> out{{F[0], F[1], F[2], F[3]}};
    launch_interpolag<T, 4, false>(c, ul, vl, wl, in, out, 0);
  }
  LCHK(c, hipGetLastError());
  op_fill<T>(c, ul, n3, T(0));
  op_fill<T>(c, vl, n3, T(0));
  op_fill<T>(c, wl, n3, T(0));
}

// the same interpolation from the fields F into four fields of scratch (pads
// copied), for an update whose epilogue reads the interpolated values and
// writes F: no copy of F beforehand.  Returns the scratch fields (4 n3).
template <typename T>
T* op_interpolag4_into(les_ctx* c, const T* const F[4], T* ul, T* vl, T* wl) {
  const Grid& g = c->g;
  const size_t n3 = (size_t)g.ld * g.ny * g.nz;
  T* Fo = (T*)scratch(c, S_FF, 4 * n3 * sizeof(T));
  {
    ProfScope _ps(c, "interpolag", (4.0 + 3 + 4) * n3 * sizeof(T));
    const In4<T> in{{F[0], F[1], F[2], F[3]}};
    const Out4<T> out{{Fo, Fo + n3, Fo + 2 * n3, Fo + 3 * n3}};
    launch_interpolag<T, 4, false>(c, ul, vl, wl, in, out, 1);
  }
  LCHK(c, hipGetLastError());
  op_fill<T>(c, ul, n3, T(0));
  op_fill<T>(c, vl, n3, T(0));
  op_fill<T>(c, wl, n3, T(0));
  return Fo;
}

template <typename T>
void op_interpolag(les_ctx* c, const SgsPtrs& s) {
  OpTag _tag(c, "interpolag");
  T* F[4] = {(T*)s.F_LM, (T*)s.F_MM, (T*)s.F_QN, (T*)s.F_NN};
  op_interpolag4<T>(c, F, (T*)s.u_lag, (T*)s.v_lag, (T*)s.w_lag);
}

// interpolag_Ssim.f90 on (F_LM, F_MM)
template <typename T>
void op_interpolag_ssim(les_ctx* c, const SgsPtrs& s) {
  OpTag _tag(c, "interpolag_ssim");
  const Grid& g = c->g;
  const size_t n3 = (size_t)g.ld * g.ny * g.nz;
  T* FF = (T*)scratch(c, S_FF, 2 * n3 * sizeof(T));
  op_copy<T>(c, FF, (T*)s.F_LM, n3);
  op_copy<T>(c, FF + n3, (T*)s.F_MM, n3);
  {
    ProfScope _ps(c, "interpolag", (2.0 + 3 + 2) * n3 * sizeof(T));
    const In4<T> in{{FF, FF + n3, nullptr, nullptr}};
    const Out4<T> out{{(T*)s.F_LM, (T*)s.F_MM, nullptr, nullptr}};
    launch_interpolag<T, 2, true>(c, (const T*)s.u_lag, (const T*)s.v_lag, (const T*)s.w_lag, in, out, 0);
  }
  LCHK(c, hipGetLastError());
  op_fill<T>(c, (T*)s.u_lag, n3, T(0));
  op_fill<T>(c, (T*)s.v_lag, n3, T(0));
  op_fill<T>(c, (T*)s.w_lag, n3, T(0));
}

// ------------------------------------------------------------------------
// lagrange_Sdep.f90: 21 fields, each filtered at 2 Delta and 4 Delta
// ------------------------------------------------------------------------
template <typename T>
struct LoadLag {  // field order: ub vb wb | uu uv uw vv vw ww | S11 S12 S13 S22 S23 S33 | |S|S_ij
  const T *u, *v, *w, *S11, *S12, *S13, *S22, *S23, *S33;
  int ny, nz, ld;
  int f0;  // first field transformed
  static constexpr double reads() { return 9.0 / 21.0; }  // u,v,w,S_ij feed 21 rows
  __device__ T operator()(int plane, int y, int x) const {
    const int fr = plane / nz, k = plane - fr * nz, f = fr + f0;
    const long i = ((long)k * ny + y) * ld + x, pl = (long)ld * ny;
    if (f < 9) {
      T ub, vb, wb;
      if (k == 0) { ub = u[i]; vb = v[i]; wb = T(0.25) * w[i + pl]; }
      else { ub = T(0.5) * (u[i] + u[i - pl]); vb = T(0.5) * (v[i] + v[i - pl]); wb = w[i]; }
      switch (f) {
        case 0: return ub;
        case 1: return vb;
        case 2: return wb;
        case 3: return ub * ub;
        case 4: return ub * vb;
        case 5: return ub * wb;
        case 6: return vb * vb;
        case 7: return vb * wb;
        default: return wb * wb;
      }
    }
    const T* Sp[6] = {S11, S12, S13, S22, S23, S33};
    if (f < 15) return Sp[f - 9][i];
    const T s = smag(S11[i], S12[i], S13[i], S22[i], S23[i], S33[i]);
    return s * Sp[f - 15][i];
  }
};


// all LoadLag fields F0 .. 20 of a point at once (k_row_fwd_mf), in LoadLag's
// evaluation order
template <typename T, int F0>
struct LoadLagAll {
  static constexpr int NF = 21 - F0;
  LoadLag<T> L;
  __device__ void all(int k, int y, int x, T (&v)[NF]) const {
    const long i = ((long)k * L.ny + y) * L.ld + x, pl = (long)L.ld * L.ny;
    if constexpr (F0 == 0) {
      T ub, vb, wb;
      if (k == 0) { ub = L.u[i]; vb = L.v[i]; wb = T(0.25) * L.w[i + pl]; }
      else { ub = T(0.5) * (L.u[i] + L.u[i - pl]); vb = T(0.5) * (L.v[i] + L.v[i - pl]); wb = L.w[i]; }
      v[0] = ub; v[1] = vb; v[2] = wb;
      v[3] = ub * ub; v[4] = ub * vb; v[5] = ub * wb;
      v[6] = vb * vb; v[7] = vb * wb; v[8] = wb * wb;
    }
    const T s11 = L.S11[i], s12 = L.S12[i], s13 = L.S13[i], s22 = L.S22[i], s23 = L.S23[i], s33 = L.S33[i];
    const T Sv[6] = {s11, s12, s13, s22, s23, s33};
    const T sm = smag(s11, s12, s13, s22, s23, s33);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      v[9 - F0 + q] = Sv[q];
      v[15 - F0 + q] = sm * Sv[q];
    }
  }
};

template <typename T>
__global__ void k_lag_epilogue(const T* __restrict__ F2, const T* __restrict__ F4, const T* __restrict__ S11,
                               const T* __restrict__ S12, const T* __restrict__ S13, const T* __restrict__ S22,
                               const T* __restrict__ S23, const T* __restrict__ S33, const T* __restrict__ Fi,
                               T* FLM, T* FMM, T* FQN, T* FNN, T* TnLM, T* TnQN, T* Beta, T* Cs, T* L11t, T* L22t,
                               T* L33t, T* Q11t, T* Q22t, T* Q33t, double* red, int nx, int ld, int ny, int nz,
                               double cst, double opftdelta, double lagran_dt, double expo, int init_lm,
                               int init_qn) {
  const long pl = (long)ld * ny, n3 = pl * nz;
  const int n2 = ld * ny;  // pad columns too: the reference loops over ld there
  const double e24 = (double)1e-24f;
  const int k = blockIdx.y;  // one level per grid row: plane sums reduce per block
  double acc2 = 0.0, acc4 = 0.0, accc = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n2; r += gridDim.x * blockDim.x) {
    const int y = r / ld, x = r - y * ld;
    const bool pad = x >= nx;  // the filtered fields are 0 there; lagrange_Sdep.f90:295-296,344-420
    const long i = (long)k * pl + (long)y * ld + x;
    auto f2 = [&](int f) { return (double)F2[f * n3 + i]; };
    auto f4 = [&](int f) { return (double)F4[f * n3 + i]; };
    const double ub = f2(0), vb = f2(1), wb = f2(2), uh = f4(0), vh = f4(1), wh = f4(2);
    double Lv[6], Qv[6];
    Lv[0] = f2(3) - ub * ub; Lv[1] = f2(4) - ub * vb; Lv[2] = f2(5) - ub * wb;
    Lv[3] = f2(6) - vb * vb; Lv[4] = f2(7) - vb * wb; Lv[5] = f2(8) - wb * wb;
    Qv[0] = f4(3) - uh * uh; Qv[1] = f4(4) - uh * vh; Qv[2] = f4(5) - uh * wh;
    Qv[3] = f4(6) - vh * vh; Qv[4] = f4(7) - vh * wh; Qv[5] = f4(8) - wh * wh;
    L11t[i] = (T)Lv[0]; L22t[i] = (T)Lv[3]; L33t[i] = (T)Lv[5];
    Q11t[i] = (T)Qv[0]; Q22t[i] = (T)Qv[3]; Q33t[i] = (T)Qv[5];
    double sb[6], sh[6], ssb[6], ssh[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) { sb[m] = f2(9 + m); sh[m] = f4(9 + m); ssb[m] = f2(15 + m); ssh[m] = f4(15 + m); }
    const double Sb = sqrt(2.0 * (sb[0] * sb[0] + sb[3] * sb[3] + sb[5] * sb[5] +
                                  2.0 * (sb[1] * sb[1] + sb[2] * sb[2] + sb[4] * sb[4])));
    const double Sh = sqrt(2.0 * (sh[0] * sh[0] + sh[3] * sh[3] + sh[5] * sh[5] +
                                  2.0 * (sh[1] * sh[1] + sh[2] * sh[2] + sh[4] * sh[4])));
    double M[6], N[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      M[m] = cst * (ssb[m] - 4.0 * Sb * sb[m]);
      N[m] = cst * (ssh[m] - 16.0 * Sh * sh[m]);
    }
    const double LM = Lv[0] * M[0] + Lv[3] * M[3] + Lv[5] * M[5] + 2.0 * (Lv[1] * M[1] + Lv[2] * M[2] + Lv[4] * M[4]);
    const double MM = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
    const double QN = Qv[0] * N[0] + Qv[3] * N[3] + Qv[5] * N[5] + 2.0 * (Qv[1] * N[1] + Qv[2] * N[2] + Qv[4] * N[4]);
    const double NN = N[0] * N[0] + N[3] * N[3] + N[5] * N[5] + 2.0 * (N[1] * N[1] + N[2] * N[2] + N[4] * N[4]);
    // the interpolated F (interpolag_Sdep's output, four fields of Fi)
    double flm = Fi[i], fmm = Fi[n3 + i], fqn = Fi[2 * n3 + i], fnn = Fi[3 * n3 + i];
    if (init_lm) { fmm = pad ? 1.0 : MM; flm = pad ? 1.0 : 0.03 * MM; }
    double Tn = fmax(fq(flm * fmm), e24);
    TnLM[i] = (T)Tn;
    Tn = opftdelta * pow(Tn, -0.125);
    Tn = fmax(e24, fq(Tn));
    double dumfac = lagran_dt / Tn;
    double epsi = dumfac / (1.0 + dumfac);
    flm = (epsi * LM + (1.0 - epsi) * flm);
    fmm = (epsi * MM + (1.0 - epsi) * fmm);
    flm = fmax(e24, fq(flm));
    const double cs2d = fmax(e24, fq(pad ? e24 : flm / fmm));
    if (init_qn) { fnn = pad ? 1.0 : NN; fqn = pad ? 1.0 : 0.03 * NN; }
    Tn = fmax(fq(fqn * fnn), e24);
    TnQN[i] = (T)Tn;
    Tn = opftdelta * pow(Tn, -0.125);
    Tn = fmax(e24, fq(Tn));
    dumfac = lagran_dt / Tn;
    epsi = dumfac / (1.0 + dumfac);
    fqn = (epsi * QN + (1.0 - epsi) * fqn);
    fnn = (epsi * NN + (1.0 - epsi) * fnn);
    fqn = fmax(e24, fq(fqn));
    const double cs4d = fmax(e24, fq(pad ? e24 : fqn / fnn));
    double beta = expo == 1.0 ? cs4d / cs2d : pow(cs4d / cs2d, expo);  // exponent ln2/(ln4-ln2) is 1
    if (k == nz - 1) beta = 1.0;
    const double bclip = fmax(fq(beta), fq(0.125));
    const double cso = fmax(e24, fq(pad ? e24 : cs2d / bclip));
    FLM[i] = (T)flm; FMM[i] = (T)fmm; FQN[i] = (T)fqn; FNN[i] = (T)fnn;
    Beta[i] = (T)beta;
    Cs[i] = (T)cso;
    if (!pad) {
      acc2 += cs2d;
      acc4 += cs4d;
      accc += cso;
    }
  }
  __shared__ double sh[3][256];
  sh[0][threadIdx.x] = acc2;
  sh[1][threadIdx.x] = acc4;
  sh[2][threadIdx.x] = accc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int j = 0; j < 3; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + s];
    __syncthreads();
  }
  // one partial per block; k_lag_avg adds them in block order, so the plane
  // sums have the same bits on every run (no atomics)
  if (threadIdx.x < 3) red[(3 * (long)k + threadIdx.x) * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// lagrange_Sdep.f90:425-426: Beta_avg = sum Cs2_2d / sum Cs_opt2,
// Betaclip_avg = sum Cs2_4d / sum Cs2_2d over the nx*ny points of each level
__global__ void k_lag_avg(const double* red, double* Beta_avg, double* Betaclip_avg, int nz, int nb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nz) {
    double s[3];
    for (int j = 0; j < 3; ++j) {
      const double* p = red + (3 * (long)k + j) * nb;
      double a = 0.0;
      for (int b = 0; b < nb; ++b) a += p[b];
      s[j] = a;
    }
    Beta_avg[k] = s[0] / s[2];
    Betaclip_avg[k] = s[1] / s[0];
  }
}

// The 21 lagrange_Sdep fields (lagrange_Sdep.f90:131-266) filtered at 2 Delta
// and 4 Delta into the S_LAGF scratch: F2 = fields 0..20, F4 = 21..41, each
// (ld,ny,nz).  Field order: ub vb wb | uu uv uw vv vw ww | S11..S33 | |S|S_ij.
template <typename T>
T* lag_filter_mom(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, int f0) {
  const Grid& g = c->g;
  const int nz = g.nz, k2 = c->keep2, k4 = c->keep4, kf = k2 > k4 ? k2 : k4, nf = 21 - f0;
  const long rows = (long)nf * nz * g.ny;
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, (size_t)rows * kf * sizeof(cx<T>));
  LoadLag<T> L{u, v, w, S[0], S[1], S[2], S[3], S[4], S[5], g.ny, nz, g.ld, f0};
  const bool mf = f0 == 0 ? launch_row_fwd_mf<T>(c, LoadLagAll<T, 0>{L}, X0, kf, 9.0)
                          : (f0 == 9 ? launch_row_fwd_mf<T>(c, LoadLagAll<T, 9>{L}, X0, kf, 6.0) : false);
  if (!mf) launch_row_fwd<T>(c, L, X0, kf, rows, g.nx, kf, g.ny);
  cx<T>* X2 = (cx<T>*)scratch(c, S_XS1, (size_t)rows * k2 * sizeof(cx<T>));
  cx<T>* X4 = (cx<T>*)scratch(c, S_XS2, (size_t)rows * k4 * sizeof(cx<T>));
  ColArgs a = col_args(g.ny, g.ny, kf, kf, (long)g.ny * kf);
  a.lhG = g.lh;
  col_out(a, 0, K_G, k2, k2, (long)g.ny * k2);
  col_out(a, 1, K_G, k4, k4, (long)g.ny * k4);
  const Tables<T>& t = tables<T>(c);
  launch_col<T>(c, X0, nullptr, X2, X4, nullptr, a, (long)nf * nz, kf, t.G2, t.G4);
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  T* F = (T*)scratch(c, S_LAGF, 42 * n3 * sizeof(T));
  // row_inv of fields f0..20: one "field" spanning them (rows are field-major)
  StoreFields<T> S2{{F + f0 * n3, nullptr, nullptr, nullptr}, nf * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X2, k2, k2, S2, rows, g.nx, g.ld, g.ny);
  StoreFields<T> S4{{F + (21 + f0) * n3, nullptr, nullptr, nullptr}, nf * nz, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X4, k4, k4, S4, rows, g.nx, g.ld, g.ny);
  return F;
}

template <typename T>
void op_lagrange(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, const SgsPtrs& s) {
  OpTag _tag(c, "lagrange");
  const Grid& g = c->g;
  const int nz = g.nz;
  c->last_lag_jt = c->jt;
  const T* Fi;
  {  // interpolag_Sdep (lagrange_Sdep.f90:128) into scratch; the epilogue writes F
    OpTag _t2(c, "interpolag");
    const T* F4s[4] = {(const T*)s.F_LM, (const T*)s.F_MM, (const T*)s.F_QN, (const T*)s.F_NN};
    Fi = op_interpolag4_into<T>(c, F4s, (T*)s.u_lag, (T*)s.v_lag, (T*)s.w_lag);
  }
  T* F = lag_filter_mom<T>(c, S, u, v, w);
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const unsigned bpl = (unsigned)((g.ld * g.ny + 255) / 256 < 64 ? (g.ld * g.ny + 255) / 256 : 64);
  double* red = (double*)scratch(c, S_RED, 3 * (size_t)nz * bpl * sizeof(double));
  const int init_lm = g.inilag && !c->flm_init && (c->jt == lag_count(c) || c->jt == 1);
  const int init_qn = g.inilag && !c->fqn_init && (c->jt == lag_count(c) || c->jt == 1);
  const double delta = g.delta;
  {
    ProfScope _ps(c, "lag_epilogue", (42.0 + 4 + 14) * n3 * sizeof(T));
    hipLaunchKernelGGL((k_lag_epilogue<T>), dim3(bpl, nz), dim3(256), 0, c->stream, F, F + 21 * n3,
                       S[0], S[1], S[2], S[3], S[4], S[5], Fi, (T*)s.F_LM, (T*)s.F_MM, (T*)s.F_QN, (T*)s.F_NN,
                       (T*)s.TnLMMM, (T*)s.TnQNNN, (T*)s.Beta, (T*)s.Cs_opt2, (T*)s.L11t, (T*)s.L22t, (T*)s.L33t,
                       (T*)s.Q11t, (T*)s.Q22t, (T*)s.Q33t, red, g.nx, g.ld, g.ny, nz, 2.0 * (delta * delta),
                       1.5 * delta, g.dt * (double)lag_count(c), log(2.0) / (log(4.0) - log(2.0)), init_lm, init_qn);
  }
  LCHK(c, hipGetLastError());
  if (s.Beta_avg && s.Betaclip_avg) {
    hipLaunchKernelGGL(k_lag_avg, dim3((nz + 63) / 64), dim3(64), 0, c->stream, red, (double*)s.Beta_avg,
                       (double*)s.Betaclip_avg, nz, (int)bpl);
    LCHK(c, hipGetLastError());
  }
  if (init_lm) c->flm_init = 1;
  if (init_qn) c->fqn_init = 1;
}

// lagrange_Ssim.f90:142-271 from the 2-Delta filter set: Beta = 1, no
// single-precision rounding, eps = 1e-32, F_LM initialised to 0.025 MM.
// Every point of the (ld, ny) planes, pads included, like the reference.
template <typename T>
__global__ void k_lag_ssim_epilogue(const T* __restrict__ F2, T* FLM, T* FMM, T* Beta, T* Cs, int nx, int ld, int ny,
                                    int nz, double cst, double opftdelta, double lagran_dt, int init) {
#pragma clang fp contract(off)
  const long pl = (long)ld * ny, n3 = pl * nz;
  const double eps = 1.e-32;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const bool pad = kx_.x >= nx;
    auto f2 = [&](int f) { return (double)F2[f * n3 + i]; };
    const double ub = f2(0), vb = f2(1), wb = f2(2);
    double Lv[6];
    Lv[0] = f2(3) - ub * ub; Lv[1] = f2(4) - ub * vb; Lv[2] = f2(5) - ub * wb;
    Lv[3] = f2(6) - vb * vb; Lv[4] = f2(7) - vb * wb; Lv[5] = f2(8) - wb * wb;
    double sb[6], ssb[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) { sb[m] = f2(9 + m); ssb[m] = f2(15 + m); }
    const double Sb = sqrt(2.0 * (sb[0] * sb[0] + sb[3] * sb[3] + sb[5] * sb[5] +
                                  2.0 * (sb[1] * sb[1] + sb[2] * sb[2] + sb[4] * sb[4])));
    double M[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) M[m] = cst * (ssb[m] - 4.0 * Sb * sb[m]);  // fourbeta = 4 Beta = 4
    const double LM = Lv[0] * M[0] + Lv[3] * M[3] + Lv[5] * M[5] + 2.0 * (Lv[1] * M[1] + Lv[2] * M[2] + Lv[4] * M[4]);
    const double MM = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
    double flm = FLM[i], fmm = FMM[i];
    if (init) { fmm = pad ? 1.0 : MM; flm = pad ? 1.0 : 0.025 * MM; }  // :235-244
    double Tn = fmax(flm * fmm, eps);                                     // :256-261
    Tn = opftdelta * pow(Tn, -0.125);
    const double dumfac = lagran_dt / Tn;
    const double epsi = dumfac / (1.0 + dumfac);
    flm = (epsi * LM + (1.0 - epsi) * flm);
    fmm = (epsi * MM + (1.0 - epsi) * fmm);
    flm = fmax(eps, flm);
    const double cs = fmax(eps, pad ? 0.0 : flm / fmm);                   // :267-269
    FLM[i] = (T)flm;
    FMM[i] = (T)fmm;
    Beta[i] = T(1);
    Cs[i] = (T)cs;
  }
}

template <typename T>
void op_lagrange_ssim(les_ctx* c, const T* const S[6], const T* u, const T* v, const T* w, const SgsPtrs& s) {
  OpTag _tag(c, "lagrange_ssim");
  const Grid& g = c->g;
  const int nz = g.nz;
  c->last_lag_jt = c->jt;
  op_interpolag_ssim<T>(c, s);
  T* F = lag_filter_mom<T>(c, S, u, v, w);
  const size_t n3 = (size_t)g.ld * g.ny * nz;
  const int init = g.inilag && !c->flm_init && (c->jt == lag_count(c) || c->jt == 1);
  const double delta = g.delta;
  {
    ProfScope _ps(c, "lag_epilogue", (21.0 + 2 + 4) * n3 * sizeof(T));
    hipLaunchKernelGGL((k_lag_ssim_epilogue<T>), grid1(n3), dim3(256), 0, c->stream, (const T*)F, (T*)s.F_LM,
                       (T*)s.F_MM, (T*)s.Beta, (T*)s.Cs_opt2, g.nx, g.ld, g.ny, nz, 2.0 * (delta * delta),
                       1.5 * delta, g.dt * (double)lag_count(c), init);
  }
  LCHK(c, hipGetLastError());
  // :270-271: plane sums of Beta = 1
  if (s.Beta_avg) op_fill<double>(c, (double*)s.Beta_avg, nz, (double)g.nx * g.ny);
  if (s.Betaclip_avg) op_fill<double>(c, (double*)s.Betaclip_avg, nz, (double)g.nx * g.ny);
  if (init) c->flm_init = 1;
}

template <typename T>
void op_sgs_stag(les_ctx* c, const T* const gr[9], const T* u, const T* v, const T* w, T* const tau[6],
                 const SgsPtrs& s, T* theta, const ScalarSgsPtrs* ss) {
  OpTag _tag(c, "sgs_stag");
  const Grid& g = c->g;
  const size_t n3 = (size_t)g.ld * g.ny * g.nz;
  const int model = c->sgs_model;
  const bool lagr = model == 4 || model == 5;  // u_lag accumulates under models 4, 5, 7 (sgs_stag.f90:59-69)
  const bool lag_s = ss && model == 5;          // u_lag_s too, with the scalar Lagrangian model
  const int jt = c->jt;
  // :165-174 sets Cs_opt2 = 0.16**2 (default real) under models 5-7 while
  // jt_total < 500; model 1 sets Co**2 = 0.16_rprec**2 every step
  const int use_const = model == 1 || ((model == 5 || model == 6) && c->jt_total < 500);
  const T cs0 = model == 1 ? (T)(0.16 * 0.16) : (T)(double)(0.16f * 0.16f);
  const T l2 = (T)(g.delta * g.delta);
  // a dynamic update this step reads S_ij; otherwise S_ij and tau are one pass
  const bool dyn = model != 1 && !(jt == 1 && g.inilag) && (jt >= 1 || g.initu) && jt % g.cs_count == 0;
  if (!dyn) {
    if (model != 1 && jt == 1 && g.inilag) op_fill<T>(c, (T*)s.Cs_opt2, n3, T(0.03));
    Grads9<T> G;
    for (int q = 0; q < 9; ++q) G.p[q] = gr[q];
    ProfScope _ps(c, "sij_tau", (9.0 + (lagr ? 6 : 0) + (lag_s ? 6 : 0) + 3 + 1 + 1 + 6) * n3 * sizeof(T));
    constexpr int LR = 8;  // levels per thread
    const dim3 gcol((unsigned)(((size_t)g.ld * g.ny + 255) / 256), (unsigned)((g.nz + LR - 1) / LR));
    hipLaunchKernelGGL((k_sij_tau_col<T, LR>), gcol, dim3(256), 0, c->stream, G, u, v, w,
                       lagr ? (T*)s.u_lag : (T*)nullptr, (T*)s.v_lag, (T*)s.w_lag,
                       lag_s ? (T*)ss->u_lag_s : (T*)nullptr, lag_s ? (T*)ss->v_lag_s : (T*)nullptr,
                       lag_s ? (T*)ss->w_lag_s : (T*)nullptr, (const T*)s.Cs_opt2, cs0, use_const, (T*)s.Cs_opt2,
                       (T*)s.Nu_t, tau[0], tau[1], tau[2], tau[3], tau[4], tau[5], g.nx, g.ld, g.ny, g.nz, l2,
                       model == 1 ? (const double*)c->l2z : nullptr, ss ? (T*)ss->Pr_t : (T*)nullptr, (T)c->scal.Pr);
    LCHK(c, hipGetLastError());
    return;
  }
  T* Sb = (T*)scratch(c, S_SIJ, 6 * n3 * sizeof(T));
  T* S[6] = {Sb, Sb + n3, Sb + 2 * n3, Sb + 3 * n3, Sb + 4 * n3, Sb + 5 * n3};
  {
    ProfScope _ps(c, "calc_sij", (9.0 + 6 + (lagr ? 6 : 0) + (lag_s ? 6 : 0) + 3) * n3 * sizeof(T));
    hipLaunchKernelGGL((k_calc_sij<T>), grid1(n3), dim3(256), 0, c->stream, gr[0], gr[1], gr[2], gr[3], gr[4], gr[5],
                       gr[6], gr[7], gr[8], S[0], S[1], S[2], S[3], S[4], S[5], g.nx, g.ld, g.ny, g.nz, u, v, w,
                       lagr ? (T*)s.u_lag : (T*)nullptr, (T*)s.v_lag, (T*)s.w_lag,
                       lag_s ? (T*)ss->u_lag_s : (T*)nullptr, lag_s ? (T*)ss->v_lag_s : (T*)nullptr,
                       lag_s ? (T*)ss->w_lag_s : (T*)nullptr);
  }
  LCHK(c, hipGetLastError());
  {  // the dynamic update (:56-163)
    if (model == 2) {
      const T* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
      op_std_dynamic<T>(c, Sc, u, v, w, nullptr, (T*)s.Cs_opt2);
    } else if (model == 4) {
      const T* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
      op_lagrange_ssim<T>(c, Sc, u, v, w, s);
    } else if (model == 3) {
      // sgs_stag.f90:132 passes ten arguments to the seven-argument routine,
      // so its S11..S33 are bound to u, v, w, S11, S12, S13
      const T* Sd[6] = {u, v, w, S[0], S[1], S[2]};
      op_scaledep_dynamic<T>(c, Sd, u, v, w, nullptr, nullptr, (T*)s.Cs_opt2);
    } else if (model == 6) {  // :143-149
      const T* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
      op_new_scaledep_dynamic<T>(c, Sc, u, v, w, nullptr, (T*)s.Cs_opt2);
      if (ss) op_scalar_scaledep<T>(c, Sc, u, v, w, theta, (const T*)s.Cs_opt2, *ss, true);
    } else if (c->closure_mode && c->mlp && jt % c->closure_period != 0) {
      // learned closure between dynamic updates: only the filtered strain
      // rates are needed (fields 9..20 of the lagrange_Sdep filter set)
      const T* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
      const T* F = lag_filter_mom<T>(c, Sc, u, v, w, 9);
      op_mlp_closure<T>(c, F, Sc, gr, (T*)s.Cs_opt2, (T*)s.Beta, 0, c->closure_thresh);
    } else {
      const T* Sc[6] = {S[0], S[1], S[2], S[3], S[4], S[5]};
      op_lagrange<T>(c, Sc, u, v, w, s);
      if (ss) op_lagrange_scalar<T>(c, Sc, u, v, w, theta, s, *ss, true);  // sgs_stag.f90:150-153
      if (c->closure_mode && c->mlp) {  // the dynamic step judges the network
        const T* F = (const T*)scratch(c, S_LAGF, 42 * n3 * sizeof(T));
        op_mlp_closure<T>(c, F, Sc, gr, (T*)s.Cs_opt2, (T*)s.Beta, 1, c->closure_thresh);
      }
    }
  }
  {
    ProfScope _ps2(c, "sgs_tau", (6.0 + 1 + 1 + 6) * n3 * sizeof(T));
    hipLaunchKernelGGL((k_sgs_tau<T>), grid1(n3), dim3(256), 0, c->stream, S[0], S[1], S[2], S[3], S[4], S[5],
                       (const T*)s.Cs_opt2, cs0, use_const, (T*)s.Cs_opt2, (T*)s.Nu_t, tau[0], tau[1], tau[2], tau[3],
                       tau[4], tau[5], g.nx, g.ld, g.ny, g.nz, l2, model == 1 ? (const double*)c->l2z : nullptr,
                       ss ? (T*)ss->Pr_t : (T*)nullptr, (T)c->scal.Pr);
  }
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// press_stag_array.f90 + tridag_array.f90
// ------------------------------------------------------------------------
template <typename T>
struct LoadPress {  // fields Hx, Hy, Hz on levels 0..nz-2, scaled (const/tadv1)*(u/dt)
  const T *u, *v, *w;
  int ny, nl, ld;
  T cst, dt;
  int v2 = 0;  // rows aligned to two elements (set by the launcher)
  static constexpr double reads() { return 1.0; }
  __device__ T operator()(int plane, int y, int x) const { return xf(row_ptr(plane, y)[x]); }
  __device__ T xf(T a) const { return cst * (a / dt); }
  __device__ const T* row_ptr(int plane, int y) const {
    const int f = plane / nl, r = plane - f * nl;
    const T* p = f == 0 ? u : (f == 1 ? v : w);
    return p + ((long)r * ny + y) * ld;
  }
  bool rows_aligned() const {
    for (const T* q : {u, v, w})
      if ((uintptr_t)q % (2 * sizeof(T))) return false;
    return ld % 2 == 0;
  }
};
template <typename T>
struct LoadPlanes2 {  // two planes of one field: level 0 and level kt
  const T* f;
  int ny, ld, kt;
  T cst;
  static constexpr double reads() { return 1.0; }
  __device__ T operator()(int plane, int y, int x) const {
    const int k = plane == 0 ? 0 : kt;
    return cst * f[((long)k * ny + y) * ld + x];
  }
};

// Two threads per horizontal mode (kx < nx/2, all ky), one per component of
// the complex right-hand side (the matrix is real); Thomas algorithm in FP64
// along z with the boundary rows of press_stag_array.f90:142-165 and the
// zero-mode integration of :238-247.  H: 3 fields x (nz-1) planes of
// (ny, nxh) complex; BT: bottom/top planes; P: (nz) planes of p_hat.  The
// real part of i (kx Hx + ky Hy) takes the imaginary parts of Hx, Hy and vice
// versa.  Both threads of a mode form the same elimination factors and each
// writes and re-reads its own copy of gamma's slot (same address, same value).
// MODE (the schedule les_autotune picks): where the sweep keeps gamma and
// the eliminated right-hand side U -- 0 both in HBM scratch, 1 gamma in LDS,
// 2 both in LDS (one column of (nz+1) values per mode / component).
template <typename T, int MODE>
__global__ void k_tridiag(const cx<T>* __restrict__ H, const cx<T>* __restrict__ BT, cx<T>* __restrict__ P,
                          double* __restrict__ gam_g, cx<double>* __restrict__ U_g, const T* __restrict__ kxt,
                          const T* __restrict__ kyt, const double* __restrict__ kxd, const double* __restrict__ kyd,
                          int nxh, int ny, int nz, double dz, int* err) {
  extern __shared__ __align__(16) unsigned char tri_smem[];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = t >> 1, comp = t & 1;
  const int nm = nxh * ny;
  if (m >= nm) return;  // blockDim is even: both threads of a mode leave together
  const int tid = threadIdx.x, bd = blockDim.x, bm = bd >> 1, tm = tid >> 1;
  double* gl = reinterpret_cast<double*>(tri_smem);
  double* ul = reinterpret_cast<double*>(tri_smem + (size_t)(nz + 1) * bm * sizeof(double));
  double* Ug = reinterpret_cast<double*>(U_g);
  auto gam = [&](int j) -> double& { return MODE >= 1 ? gl[j * bm + tm] : gam_g[(long)j * nm + m]; };
  auto U = [&](int j) -> double& { return MODE >= 2 ? ul[j * bd + tid] : Ug[2 * ((long)j * nm + m) + comp]; };
  const int y = m / nxh, x = m - y * nxh;
  const long ps = (long)nm;               // plane stride (complex)
  const long fs = ps * (nz - 1);          // field stride in H
  const T* Hx = reinterpret_cast<const T*>(H);
  const T* Hy = reinterpret_cast<const T*>(H + fs);
  const T* Hz = reinterpret_cast<const T*>(H + 2 * fs);
  const T* B2 = reinterpret_cast<const T*>(BT);
  T* P2 = reinterpret_cast<T*>(P);
  auto at = [&](long j) { return 2 * (j * ps + m); };  // element of mode m at plane j
  const double bot = (double)B2[at(0) + comp], top = (double)B2[at(1) + comp];
  if (y == ny / 2) {
    for (int j = 0; j < nz; ++j) P2[at(j) + comp] = T(0);
    return;
  }
  if (m == 0) {
    double pv = 0.0;
    P2[at(0) + comp] = T(0);
    pv = pv - dz * bot;
    P2[at(1) + comp] = (T)pv;
    for (int j = 2; j < nz; ++j) {
      pv = pv + (double)Hz[at(j - 1) + comp] * dz;
      P2[at(j) + comp] = (T)pv;
    }
    return;
  }
  const double kx = kxd[x], ky = kyd[y];
  const double a_in = 1.0 / (dz * dz), c_in = 1.0 / (dz * dz);
  const double b_in = -(kx * kx + ky * ky + 2.0 / (dz * dz));
  const int oc = 1 - comp;                 // component of Hx, Hy this one takes
  const double sg = comp == 0 ? -1.0 : 1.0;
  // row 0: b=-1, c=1, r=-dz*bottom
  double bet = -1.0;
  double u0 = (-dz * bot) / bet;
  U(0) = u0;
  double cprev = 1.0;
  double uprev = u0;
  // level L's (Hx, Hy, Hz) components (zero above the last H plane, nz - 2),
  // loaded eight levels ahead of the recurrence through a shift register
  auto fetch = [&](int L, T& a, T& b, T& cz) {
    a = b = cz = T(0);
    if (L <= nz - 2) {
      a = Hx[at(L) + oc];
      b = Hy[at(L) + oc];
      cz = Hz[at(L) + comp];
    }
  };
  T x1, y1, z1, x2, y2, z2, x3, y3, z3, x4, y4, z4, x5, y5, z5, x6, y6, z6, x7, y7, z7, x8, y8, z8;
  fetch(1, x1, y1, z1);
  fetch(2, x2, y2, z2);
  fetch(3, x3, y3, z3);
  fetch(4, x4, y4, z4);
  fetch(5, x5, y5, z5);
  fetch(6, x6, y6, z6);
  fetch(7, x7, y7, z7);
  fetch(8, x8, y8, z8);
  double hxm = (double)Hx[at(0) + oc], hym = (double)Hy[at(0) + oc], hzm = (double)Hz[at(0) + comp];
  for (int j = 1; j <= nz; ++j) {
    const double hxj = (double)x1, hyj = (double)y1, hzj = (double)z1;  // level j
    x1 = x2; y1 = y2; z1 = z2;
    x2 = x3; y2 = y3; z2 = z3;
    x3 = x4; y3 = y4; z3 = z4;
    x4 = x5; y4 = y5; z4 = z5;
    x5 = x6; y5 = y6; z5 = z6;
    x6 = x7; y6 = y7; z6 = z7;
    x7 = x8; y7 = y8; z7 = z8;
    fetch(j + 8, x8, y8, z8);
    double aj, bj, rj;
    if (j < nz) {
      aj = a_in;
      bj = b_in;
      const double sv = kx * hxm + ky * hym;
      rj = sg * sv + (hzj - hzm) / dz;
      hxm = hxj;
      hym = hyj;
      hzm = hzj;
    } else {
      aj = -1.0;
      bj = 1.0;
      rj = -top * dz;
    }
    const double gj = cprev / bet;
    bet = bj - aj * gj;
    if (bet == 0.0) atomicExch(err, 2);
    gam(j) = gj;
    uprev = (rj - aj * uprev) / bet;
    U(j) = uprev;
    cprev = j < nz ? c_in : 0.0;
  }
  // back substitution; p_hat levels 0..nz-1 are kept (level nz is BOGUS later)
  double un = uprev;
  for (int j = nz - 1; j >= 0; --j) {
    double uj = U(j);
    const double gn = gam(j + 1);
    uj = uj - gn * un;
    P2[at(j) + comp] = (T)uj;
    un = uj;
  }
}

template <typename T>
struct StorePress {  // field 0: p levels 1..nz-1 ; 1: dpdx levels 1..nz-1 ; 2: dpdy
  T *p, *dpdx, *dpdy;
  int ppf, ld, nx, ny;
  static constexpr double writes() { return 1.0; }
  __device__ void operator()(int plane, int y, int x, T v) const {
    const int f = plane / ppf, r = plane - f * ppf;
    const T val = x < nx ? v : T(0);
    const long row = (long)r * ny + y;
    if (f == 0) p[(row + ny) * ld + x] = val;  // p(:,:,1:nz-1) of p(ld,ny,0:nz)
    else if (f == 1) dpdx[row * ld + x] = val;
    else dpdy[row * ld + x] = val;
  }
};

// Two-ended Thomas sweep (schedule 4): four threads per mode, (side, component).
// Side 0 eliminates rows 0 .. mid-1 downward as k_tridiag does; side 1
// eliminates rows nz .. mid+1 upward (x_j = v_j - delta_j x_{j-1}, delta_j =
// a_j / beta'_j, beta'_j = b_j - c_j delta_{j+1}, v_j = (r_j - c_j v_{j+1}) /
// beta'_j).  The two sides meet at row mid, exchange their partial results
// through a lane swizzle, and both form x_mid from the same four numbers;
// each side then substitutes back over its own half.  Half the sequential
// chain and twice the waves of the one-ended sweep; another elimination
// order, so the same solution to rounding.  gamma / delta in LDS (one column
// of nz + 1 values per mode), u / v in HBM scratch; Nyquist row and zero mode
// as k_tridiag.
template <typename T>
__global__ void k_tridiag_tw(const cx<T>* __restrict__ H, const cx<T>* __restrict__ BT, cx<T>* __restrict__ P,
                             cx<double>* __restrict__ U_g, const double* __restrict__ kxd,
                             const double* __restrict__ kyd, int nxh, int ny, int nz, double dz, int* err) {
  extern __shared__ __align__(16) unsigned char tri_smem[];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = t >> 2, side = (t >> 1) & 1, comp = t & 1;
  const int nm = nxh * ny;
  if (m >= nm) return;  // blockDim is a multiple of 4: the four threads of a mode leave together
  const int tm = threadIdx.x >> 2, bm = blockDim.x >> 2;
  double* gl = reinterpret_cast<double*>(tri_smem);
  auto gam = [&](int j) -> double& { return gl[j * bm + tm]; };
  double* Ug = reinterpret_cast<double*>(U_g);
  auto U = [&](int j) -> double& { return Ug[2 * ((long)j * nm + m) + comp]; };
  const int y = m / nxh, x = m - y * nxh;
  const long ps = (long)nm, fs = ps * (nz - 1);
  const T* Hx = reinterpret_cast<const T*>(H);
  const T* Hy = reinterpret_cast<const T*>(H + fs);
  const T* Hz = reinterpret_cast<const T*>(H + 2 * fs);
  const T* B2 = reinterpret_cast<const T*>(BT);
  T* P2 = reinterpret_cast<T*>(P);
  auto at = [&](long j) { return 2 * (j * ps + m); };
  if (y == ny / 2) {
    if (side == 0)
      for (int j = 0; j < nz; ++j) P2[at(j) + comp] = T(0);
    return;
  }
  if (m == 0) {  // zero mode: the integration of k_tridiag, on side 0
    if (side) return;
    const double bot = (double)B2[at(0) + comp];
    double pv = 0.0;
    P2[at(0) + comp] = T(0);
    pv = pv - dz * bot;
    P2[at(1) + comp] = (T)pv;
    for (int j = 2; j < nz; ++j) {
      pv = pv + (double)Hz[at(j - 1) + comp] * dz;
      P2[at(j) + comp] = (T)pv;
    }
    return;
  }
  const double kx = kxd[x], ky = kyd[y];
  const double a_in = 1.0 / (dz * dz), c_in = 1.0 / (dz * dz);
  const double b_in = -(kx * kx + ky * ky + 2.0 / (dz * dz));
  const int oc = 1 - comp;
  const double sg = comp == 0 ? -1.0 : 1.0;
  const int mid = nz / 2;  // 1 <= mid <= nz - 1
  // level L's (Hx, Hy, Hz) components, zero above the last H plane (nz - 2)
  auto fetch = [&](int L, T& a, T& b, T& cz) {
    a = b = cz = T(0);
    if (L >= 0 && L <= nz - 2) {
      a = Hx[at(L) + oc];
      b = Hy[at(L) + oc];
      cz = Hz[at(L) + comp];
    }
  };
  // interior row j: r_j from levels j - 1 and j (press_stag_array.f90:142-165)
  auto rhs = [&](double hxm, double hym, double hzm, double hzj) { return sg * (kx * hxm + ky * hym) + (hzj - hzm) / dz; };
  constexpr int D = 8;  // levels loaded ahead of the recurrence
  T xs[D], ys[D], zs[D];
  double pnum = 0.0, pden = 0.0;  // this side's share of row mid's equation
  if (side == 0) {
    const double bot = (double)B2[at(0) + comp];
    double bet = -1.0, u = (-dz * bot) / bet, cprev = 1.0;
    U(0) = u;
#pragma unroll
    for (int q = 0; q < D; ++q) fetch(q + 1, xs[q], ys[q], zs[q]);
    double hxm = (double)Hx[at(0) + oc], hym = (double)Hy[at(0) + oc], hzm = (double)Hz[at(0) + comp];
    for (int j = 1; j <= mid; ++j) {
      const double hxj = (double)xs[0], hyj = (double)ys[0], hzj = (double)zs[0];  // level j
#pragma unroll
      for (int q = 0; q < D - 1; ++q) {
        xs[q] = xs[q + 1];
        ys[q] = ys[q + 1];
        zs[q] = zs[q + 1];
      }
      fetch(j + D, xs[D - 1], ys[D - 1], zs[D - 1]);
      const double rj = rhs(hxm, hym, hzm, hzj);
      hxm = hxj;
      hym = hyj;
      hzm = hzj;
      const double gj = cprev / bet;
      if (j == mid) {  // row mid: a (u_{mid-1} - gamma_mid x_mid) + b x_mid + c x_{mid+1} = r
        pnum = rj - a_in * u;
        pden = b_in - a_in * gj;
        gam(j) = gj;
        break;
      }
      bet = b_in - a_in * gj;
      if (bet == 0.0) atomicExch(err, 2);
      gam(j) = gj;
      u = (rj - a_in * u) / bet;
      U(j) = u;
      cprev = c_in;
    }
  } else {
    const double top = (double)B2[at(1) + comp];
    double bet = 1.0, v = (-top * dz) / bet, aprev = -1.0;  // row nz: a = -1, b = 1
    U(nz) = v;
    gam(nz) = aprev / bet;
#pragma unroll
    for (int q = 0; q < D; ++q) fetch(nz - 2 - q, xs[q], ys[q], zs[q]);
    double hzj = 0.0;  // Hz at level nz - 1 (above the last H plane)
    for (int j = nz - 1; j >= mid + 1; --j) {
      const double hxm = (double)xs[0], hym = (double)ys[0], hzm = (double)zs[0];  // level j - 1
#pragma unroll
      for (int q = 0; q < D - 1; ++q) {
        xs[q] = xs[q + 1];
        ys[q] = ys[q + 1];
        zs[q] = zs[q + 1];
      }
      fetch(j - 1 - D, xs[D - 1], ys[D - 1], zs[D - 1]);
      const double rj = rhs(hxm, hym, hzm, hzj);
      hzj = hzm;
      const double dj = aprev / bet;  // delta_{j+1}
      bet = b_in - c_in * dj;
      if (bet == 0.0) atomicExch(err, 2);
      v = (rj - c_in * v) / bet;
      U(j) = v;
      aprev = a_in;
      gam(j) = aprev / bet;  // delta_j
    }
    pnum = c_in * v;             // c x_{mid+1} = c v_{mid+1} - c delta_{mid+1} x_mid
    pden = c_in * gam(mid + 1);
  }
  // both sides form x_mid from the same four numbers
  const double onum = __shfl_xor(pnum, 2), oden = __shfl_xor(pden, 2);
  const double num0 = side == 0 ? pnum : onum, den0 = side == 0 ? pden : oden;
  const double num1 = side == 0 ? onum : pnum, den1 = side == 0 ? oden : pden;
  const double den = den0 - den1;
  if (den == 0.0) atomicExch(err, 2);
  const double xm = (num0 - num1) / den;
  if (side == 0) {
    P2[at(mid) + comp] = (T)xm;
    double un = xm;
    for (int j = mid - 1; j >= 0; --j) {
      const double uj = U(j) - gam(j + 1) * un;
      P2[at(j) + comp] = (T)uj;
      un = uj;
    }
  } else {
    double xn = xm;
    for (int j = mid + 1; j <= nz - 1; ++j) {
      const double xj = U(j) - gam(j) * xn;
      P2[at(j) + comp] = (T)xj;
      xn = xj;
    }
  }
}

// Parallel cyclic reduction: one wavefront per mode (kx, ky), four modes per
// workgroup; the (nz+1)-row system of press_stag_array.f90:142-165 (the rows
// and right-hand side k_tridiag forms) sits in LDS, FP64, and each step
// eliminates the neighbours at distance s = 1, 2, 4, ... until every row
// stands alone.  Nyquist row and zero mode as k_tridiag.  Another elimination
// order than the Thomas sweep: the same solution to rounding.
template <typename T>
__global__ void __launch_bounds__(256)
    k_tridiag_pcr(const cx<T>* __restrict__ H, const cx<T>* __restrict__ BT, cx<T>* __restrict__ P,
                  const double* __restrict__ kxd, const double* __restrict__ kyd, int nxh, int ny, int nz,
                  double dz) {
  extern __shared__ __align__(16) unsigned char tri_smem[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = nz + 1, nm = nxh * ny;
  const int m = blockIdx.x * 4 + wv;
  const bool act = m < nm;
  const int y = act ? m / nxh : 0, x = act ? m - y * nxh : 0;
  const long ps = (long)nm, fs = ps * (nz - 1);
  const bool solve = act && y != ny / 2 && m != 0;
  // per wave two buffers of the rows: a, b, c, Re r, Im r (n values each)
  double* wbase = reinterpret_cast<double*>(tri_smem) + (size_t)wv * 10 * n;
  if (act && y == ny / 2)
    for (int j = lane; j < nz; j += 64) P[j * ps + m] = {T(0), T(0)};
  if (act && m == 0 && lane == 0) {  // zero mode: integration (press_stag_array.f90:238-247)
    const cx<double> bot = {(double)BT[m].x, (double)BT[m].y};
    cx<double> pv = {0.0, 0.0};
    P[0] = {T(0), T(0)};
    pv = {pv.x - dz * bot.x, pv.y - dz * bot.y};
    P[ps] = {(T)pv.x, (T)pv.y};
    for (int j = 2; j < nz; ++j) {
      const cx<T> h = H[2 * fs + (long)(j - 1) * ps];
      pv = {pv.x + (double)h.x * dz, pv.y + (double)h.y * dz};
      P[j * ps + m] = {(T)pv.x, (T)pv.y};
    }
  }
  {
    double* o = wbase;
    const double kx = solve ? kxd[x] : 0.0, ky = solve ? kyd[y] : 0.0;
    const double a_in = 1.0 / (dz * dz), c_in = 1.0 / (dz * dz);
    const double b_in = -(kx * kx + ky * ky + 2.0 / (dz * dz));
    for (int j = lane; j < n; j += 64) {
      double a = 0.0, b = 1.0, c = 0.0, rx = 0.0, ry = 0.0;
      if (solve) {
        if (j == 0) {  // b=-1, c=1, r=-dz*bottom
          b = -1.0;
          c = 1.0;
          rx = -dz * (double)BT[m].x;
          ry = -dz * (double)BT[m].y;
        } else if (j < nz) {
          a = a_in;
          b = b_in;
          c = c_in;
          const cx<T> hx = H[(long)(j - 1) * ps + m], hy = H[fs + (long)(j - 1) * ps + m];
          const cx<T> hzm = H[2 * fs + (long)(j - 1) * ps + m];
          const cx<T> hz = j < nz - 1 ? H[2 * fs + (long)j * ps + m] : cx<T>{T(0), T(0)};
          const double sr = kx * (double)hx.x + ky * (double)hy.x, si = kx * (double)hx.y + ky * (double)hy.y;
          rx = -si + ((double)hz.x - (double)hzm.x) / dz;
          ry = sr + ((double)hz.y - (double)hzm.y) / dz;
        } else {  // top row: a=-1, b=1, r=-dz*top
          a = -1.0;
          rx = -(double)BT[ps + m].x * dz;
          ry = -(double)BT[ps + m].y * dz;
        }
      }
      o[j] = a;
      o[n + j] = b;
      o[2 * n + j] = c;
      o[3 * n + j] = rx;
      o[4 * n + j] = ry;
    }
  }
  __syncthreads();
  int cur = 0;
  for (int st = 1; st < n; st *= 2) {
    const double* ia = wbase + cur * 5 * n;
    double* oa = wbase + (cur ^ 1) * 5 * n;
    for (int j = lane; j < n; j += 64) {
      const double aj = ia[j], bj = ia[n + j], cj = ia[2 * n + j];
      double a2 = 0.0, b2 = bj, c2 = 0.0, rx = ia[3 * n + j], ry = ia[4 * n + j];
      if (j - st >= 0) {
        const int q = j - st;
        const double k1 = aj / ia[n + q];
        a2 = -ia[q] * k1;
        b2 = b2 - ia[2 * n + q] * k1;
        rx = rx - ia[3 * n + q] * k1;
        ry = ry - ia[4 * n + q] * k1;
      }
      if (j + st < n) {
        const int q = j + st;
        const double k2 = cj / ia[n + q];
        c2 = -ia[2 * n + q] * k2;
        b2 = b2 - ia[q] * k2;
        rx = rx - ia[3 * n + q] * k2;
        ry = ry - ia[4 * n + q] * k2;
      }
      oa[j] = a2;
      oa[n + j] = b2;
      oa[2 * n + j] = c2;
      oa[3 * n + j] = rx;
      oa[4 * n + j] = ry;
    }
    __syncthreads();
    cur ^= 1;
  }
  if (solve) {
    const double* ia = wbase + cur * 5 * n;
    for (int j = lane; j < nz; j += 64) {
      const double b = ia[n + j];
      P[j * ps + m] = {(T)(ia[3 * n + j] / b), (T)(ia[4 * n + j] / b)};
    }
  }
}

// LDS bytes of a tridiagonal schedule with `block` modes per workgroup (0:
// the mode does not use LDS): gamma per mode, U per mode and component
size_t tridiag_lds(int mode, int block, int nz) {
  if (mode == 3) return (size_t)4 * 10 * (nz + 1) * sizeof(double);  // four waves, two row buffers each
  if (mode == 4) return (size_t)(nz + 1) * block * sizeof(double);    // gamma / delta per mode
  return (size_t)(nz + 1) * block * (mode >= 1 ? sizeof(double) : 0) +
         (size_t)(nz + 1) * block * (mode >= 2 ? sizeof(cx<double>) : 0);
}

template <typename T>
void launch_tridiag(les_ctx* c, const cx<T>* H, const cx<T>* BT, cx<T>* PH, double* gam, cx<double>* Ub, int* err) {
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1, nxh = g.nxh, mode = c->tune.tri_mode, bs = c->tune.tri_block;
  const long ps = (long)g.ny * nxh;
  const Tables<T>& t = tables<T>(c);
  const size_t sm = tridiag_lds(mode, bs, nz);
  // bs modes per block, two threads per mode
  const dim3 grid((unsigned)((ps + bs - 1) / bs)), block((unsigned)(2 * bs));
  // gamma in HBM for schedule 0, u in HBM for 0, 1 and 4
  ProfScope _ps(c, "tridiag", (3.0 * nl + 2 + nz) * ps * sizeof(cx<T>) +
                                  (nz + 1) * ps * 2.0 * ((mode >= 1 ? 0 : sizeof(double)) +
                                                         (mode == 2 || mode == 3 ? 0 : sizeof(cx<double>))));
  if (mode == 4) {  // bs modes per block, four threads per mode
    if (sm > 65536)
      LCHK(c, hipFuncSetAttribute((const void*)k_tridiag_tw<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL((k_tridiag_tw<T>), dim3((unsigned)((ps + bs - 1) / bs)), dim3((unsigned)(4 * bs)), sm,
                       c->stream, H, BT, PH, Ub, c->td.kx, c->td.ky, nxh, g.ny, nz, g.dz, err);
  } else if (mode == 3) {
    if (sm > 65536)
      LCHK(c, hipFuncSetAttribute((const void*)k_tridiag_pcr<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL((k_tridiag_pcr<T>), dim3((unsigned)((ps + 3) / 4)), dim3(256), sm, c->stream, H, BT, PH,
                       c->td.kx, c->td.ky, nxh, g.ny, nz, g.dz);
  } else if (mode == 0) {
    hipLaunchKernelGGL((k_tridiag<T, 0>), grid, block, 0, c->stream, H, BT, PH, gam, Ub, t.kx, t.ky, c->td.kx,
                       c->td.ky, nxh, g.ny, nz, g.dz, err);
  } else if (mode == 1) {
    if (sm > 65536)
      LCHK(c, hipFuncSetAttribute((const void*)k_tridiag<T, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL((k_tridiag<T, 1>), grid, block, sm, c->stream, H, BT, PH, gam, Ub, t.kx, t.ky, c->td.kx,
                       c->td.ky, nxh, g.ny, nz, g.dz, err);
  } else {
    if (sm > 65536)
      LCHK(c, hipFuncSetAttribute((const void*)k_tridiag<T, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL((k_tridiag<T, 2>), grid, block, sm, c->stream, H, BT, PH, gam, Ub, t.kx, t.ky, c->td.kx,
                       c->td.ky, nxh, g.ny, nz, g.dz, err);
  }
  LCHK(c, hipGetLastError());
}

template <typename T>
void op_press_stag(les_ctx* c, const T* u, const T* v, const T* w, const T* divtz, T* p, T* dpdx, T* dpdy) {
  OpTag _tag(c, "press_stag");
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1, nxh = g.nxh;
  const long ps = (long)g.ny * nxh;
  const long rows = 3L * nl * g.ny;
  cx<T>* X0 = (cx<T>*)scratch(c, S_XS0, (size_t)(rows + 2 * g.ny) * nxh * sizeof(cx<T>));
  const double cstd = 1.0 / ((double)g.nx * g.ny);
  LoadPress<T> L{u, v, w, g.ny, nl, g.ld, (T)(cstd / TADV1), (T)g.dt};
  launch_row_fwd<T>(c, L, X0, nxh, rows, g.nx, nxh, g.ny);
  LoadPlanes2<T> LB{divtz, g.ny, g.ld, nz - 1, (T)cstd};
  cx<T>* XB = X0 + (size_t)rows * nxh;
  launch_row_fwd<T>(c, LB, XB, nxh, 2L * g.ny, g.nx, nxh, g.ny);
  // forward y transforms only -> spectral H (3*nl planes) and bottom/top (2)
  cx<T>* H = (cx<T>*)scratch(c, S_PSPEC, (size_t)(3 * nl + 2) * ps * sizeof(cx<T>));
  ColArgs a = col_args(g.ny, g.ny, nxh, nxh, ps);
  a.do_inv = 0;
  col_out(a, 0, K_ID, nxh, nxh, ps);
  launch_col<T>(c, X0, nullptr, H, nullptr, nullptr, a, 3L * nl + 2, nxh);
  // tridiagonal solves (FP64)
  cx<T>* PH = (cx<T>*)scratch(c, S_PHAT, (size_t)nz * ps * sizeof(cx<T>));
  char* tri = (char*)scratch(c, S_TRI, (size_t)(nz + 1) * ps * (sizeof(double) + sizeof(cx<double>)) + 256);
  double* gam = (double*)tri;
  cx<double>* Ub = (cx<double>*)(tri + (size_t)(nz + 1) * ps * sizeof(double));
  int* err = (int*)scratch(c, S_TMP2, 256);
  const Tables<T>& t = tables<T>(c);
  (void)t;
  launch_tridiag<T>(c, H, H + (size_t)3 * nl * ps, PH, gam, Ub, err);
  // inverse y transforms: levels 1..nz-1 -> p, i kx p, i ky p ; level 0 -> p
  const long rowsP = 3L * nl * g.ny;
  cx<T>* X1 = (cx<T>*)scratch(c, S_XS1, (size_t)(rowsP + g.ny) * nxh * sizeof(cx<T>));
  ColArgs b = col_args(g.ny, g.ny, nxh, nxh, ps);
  b.do_fwd = 0;
  const size_t fo = (size_t)nl * ps;
  col_out(b, 0, K_ID, nxh, nxh, ps);
  col_out(b, 1, K_IKX, nxh, nxh, ps);
  col_out(b, 2, K_IKY, nxh, nxh, ps);
  launch_col<T>(c, PH + ps, nullptr, X1, X1 + fo, X1 + 2 * fo, b, nl, nxh);
  ColArgs b0 = col_args(g.ny, g.ny, nxh, nxh, ps);
  b0.do_fwd = 0;
  col_out(b0, 0, K_ID, nxh, nxh, ps);
  cx<T>* X1b = X1 + 3 * fo;
  launch_col<T>(c, PH, nullptr, X1b, nullptr, nullptr, b0, 1, nxh);
  StorePress<T> SP{p, dpdx, dpdy, nl, g.ld, g.nx, g.ny};
  launch_row_inv<T>(c, X1, nxh, nxh, SP, rowsP, g.nx, g.ld, g.ny);
  StoreFields<T> S0{{p, nullptr, nullptr, nullptr}, 1, g.ny, g.ld, g.nx};
  launch_row_inv<T>(c, X1b, nxh, nxh, S0, g.ny, g.nx, g.ld, g.ny);
  const size_t pl = (size_t)g.ld * g.ny;
  hipLaunchKernelGGL((k_fill_bogus_cplx<T>), grid1(pl), dim3(256), 0, c->stream, p + pl * nz, pl);
  hipLaunchKernelGGL((k_fill_bogus_cplx<T>), grid1(pl), dim3(256), 0, c->stream, dpdx + pl * (nz - 1), pl);
  hipLaunchKernelGGL((k_fill_bogus_cplx<T>), grid1(pl), dim3(256), 0, c->stream, dpdy + pl * (nz - 1), pl);
  LCHK(c, hipGetLastError());
}

// tridag_array.f90 on caller arrays (lh,ny,nz+1): one thread per mode
__global__ void k_tridag_array(const double* a, const double* b, const double* cc, const cx<double>* r,
                               cx<double>* u, double* gam, int lh, int ny, int n, int* err) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int nm = lh * ny;
  if (m >= nm) return;
  const int y = m / lh, x = m - y * lh;
  if (x < lh - 1 && b[m] == 0.0) atomicExch(err, 1);
  double bet = b[m];
  u[m] = {r[m].x / bet, r[m].y / bet};
  if (y == ny / 2 || x == lh - 1 || m == 0) return;
  for (int j = 1; j < n; ++j) {
    const long i = (long)j * nm + m;
    const double gj = cc[i - nm] / bet;
    gam[i] = gj;
    bet = b[i] - a[i] * gj;
    if (bet == 0.0) atomicExch(err, 2);
    u[i] = {(r[i].x - a[i] * u[i - nm].x) / bet, (r[i].y - a[i] * u[i - nm].y) / bet};
  }
  for (int j = n - 2; j >= 0; --j) {
    const long i = (long)j * nm + m;
    u[i] = {u[i].x - gam[i + nm] * u[i + nm].x, u[i].y - gam[i + nm] * u[i + nm].y};
  }
}

size_t tridag_array_gamma_bytes(const les_ctx* c) {
  return (size_t)(c->g.nz + 1) * c->g.lh * c->g.ny * sizeof(double);
}

int tridag_array_with(les_ctx* c, const double* a, const double* b, const double* cc, const double* r, double* u,
                      double* gam, int* err) {
  const Grid& g = c->g;
  const int n = g.nz + 1, nm = g.lh * g.ny;
  op_fill<int>(c, err, 1, 0);  // a fill kernel, not a memset node, so the call can sit in a captured graph
  hipLaunchKernelGGL(k_tridag_array, dim3((nm + 127) / 128), dim3(128), 0, c->stream, a, b, cc,
                     (const cx<double>*)r, (cx<double>*)u, gam, g.lh, g.ny, n, err);
  LCHK(c, hipGetLastError());
  int h = 0;
  LCHK(c, hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  LCHK(c, hipStreamSynchronize(c->stream));
  return h;
}

int op_tridag_array(les_ctx* c, const double* a, const double* b, const double* cc, const double* r, double* u) {
  double* gam = (double*)scratch(c, S_TMP0, tridag_array_gamma_bytes(c));
  int* err = (int*)scratch(c, S_TMP2, 256);
  return tridag_array_with(c, a, b, cc, r, u, gam, err);
}

// ------------------------------------------------------------------------
// forcing.f90:128-219 project
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_project(T* u, T* v, T* w, const T* dpdx, const T* dpdy, const T* dpdz, int nx, int ld, int ny,
                          int nz, T dt) {
  const long pl = (long)ld * ny;
  const long n = pl * (nz - 1);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    T uu = u[i], vv = v[i], ww = w[i];
    if (x < nx) {
      uu = (uu + dt * (-T(TADV1) * dpdx[i] + T(0)));
      vv = (vv + dt * (-T(TADV1) * dpdy[i] + T(0)));
      if (k >= 1) ww = (ww + dt * (-T(TADV1) * dpdz[i] + T(0)));
      u[i] = uu;
      v[i] = vv;
    }
    if (k == 0) ww = T(0);
    if (k >= 1 || x >= nx) w[i] = ww;
    else w[i] = T(0);
    if (k == nz - 2) {  // lid: u(nz) = u(nz-1), w(nz) = 0
      u[i + pl] = uu;
      v[i + pl] = vv;
      w[i + pl] = T(0);
    }
  }
}
template <typename T>
void op_project(les_ctx* c, T* u, T* v, T* w, const T* dpdx, const T* dpdy, const T* dpdz) {
  const Grid& g = c->g;
  hipLaunchKernelGGL((k_project<T>), grid1((size_t)g.ld * g.ny * (g.nz - 1)), dim3(256), 0, c->stream, u, v, w,
                     dpdx, dpdy, dpdz, g.nx, g.ld, g.ny, g.nz, (T)g.dt);
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// main.f90:309-390: RHS = -conv - divt ; AB2 predictor ; BOGUS lid
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_ab2(T* u, T* v, T* w, T* Rx, T* Ry, T* Rz, const T* Rxf, const T* Ryf, const T* Rzf,
                      const T* dx, const T* dy, const T* dz, long nv, long pl, T dt, T force, int euler) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nv + pl; i += (long)gridDim.x * blockDim.x) {
    if (i >= nv) {
      u[i] = T(BOGUS); v[i] = T(BOGUS); w[i] = T(BOGUS);
      continue;
    }
    const T rx = -Rx[i] - dx[i], ry = -Ry[i] - dy[i], rz = -Rz[i] - dz[i];
    Rx[i] = rx; Ry[i] = ry; Rz[i] = rz;
    const T fx = euler ? rx : Rxf[i], fy = euler ? ry : Ryf[i], fz = euler ? rz : Rzf[i];
    u[i] = u[i] + dt * (T(TADV1) * rx + T(TADV2) * fx + force);
    v[i] = v[i] + dt * (T(TADV1) * ry + T(TADV2) * fy);
    w[i] = w[i] + dt * (T(TADV1) * rz + T(TADV2) * fz);
  }
}
template <typename T>
void op_ab2(les_ctx* c, T* u, T* v, T* w, T* RHSx, T* RHSy, T* RHSz, const T* RHSx_f, const T* RHSy_f,
            const T* RHSz_f, const T* divtx, const T* divty, const T* divtz, int euler_first) {
  const Grid& g = c->g;
  const long pl = (long)g.ld * g.ny, nv = pl * (g.nz - 1);
  {
    ProfScope _ps(c, "ab2", 15.0 * (nv + pl) * sizeof(T));
    hipLaunchKernelGGL((k_ab2<T>), grid1((size_t)(nv + pl)), dim3(256), 0, c->stream, u, v, w, RHSx, RHSy, RHSz,
                       RHSx_f, RHSy_f, RHSz_f, divtx, divty, divtz, nv, pl, (T)g.dt, (T)g.force, euler_first);
  }
  LCHK(c, hipGetLastError());
}

// main.f90:403-415: dpdz, RHS -= grad p, then project (forcing.f90:128-219)
template <typename T>
__global__ void k_pressure_update(T* u, T* v, T* w, T* Rx, T* Ry, T* Rz, const T* p, const T* dpdx,
                                  const T* dpdy, T* dpdz, int nx, int ld, int ny, int nz, T rdz, T dt) {
  const long pl = (long)ld * ny;
  const long nv = pl * (nz - 1);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nv + pl; i += (long)gridDim.x * blockDim.x) {
    if (i >= nv) {
      dpdz[i] = T(BOGUS);
      continue;
    }
    const LevelCol kx_ = level_col(i, pl, ld);
    const int k = kx_.k, x = kx_.x;
    T dz = dpdz[i];
    if (x < nx) {
      dz = (p[i + pl] - p[i]) * rdz;  // p is (ld,ny,0:nz): p(k) at i+pl
      dpdz[i] = dz;
    }
    const T px = dpdx[i], py = dpdy[i];
    Rx[i] = Rx[i] - px;
    Ry[i] = Ry[i] - py;
    Rz[i] = Rz[i] - dz;
    T uu = u[i], vv = v[i], ww = w[i];
    if (x < nx) {
      uu = (uu + dt * (-T(TADV1) * px + T(0)));
      vv = (vv + dt * (-T(TADV1) * py + T(0)));
      if (k >= 1) ww = (ww + dt * (-T(TADV1) * dz + T(0)));
      u[i] = uu;
      v[i] = vv;
    }
    if (k == 0) ww = T(0);
    w[i] = ww;
    if (k == nz - 2) {
      u[i + pl] = uu;
      v[i + pl] = vv;
      w[i + pl] = T(0);
    }
  }
}
template <typename T>
void op_pressure_update(les_ctx* c, T* u, T* v, T* w, T* RHSx, T* RHSy, T* RHSz, const T* p, const T* dpdx,
                        const T* dpdy, T* dpdz) {
  const Grid& g = c->g;
  const long pl = (long)g.ld * g.ny, nv = pl * (g.nz - 1);
  {
    ProfScope _ps(c, "pressure_update", 16.0 * (nv + pl) * sizeof(T));
    hipLaunchKernelGGL((k_pressure_update<T>), grid1((size_t)(nv + pl)), dim3(256), 0, c->stream, u, v, w, RHSx,
                       RHSy, RHSz, p, dpdx, dpdy, dpdz, g.nx, g.ld, g.ny, g.nz, (T)(1.0 / g.dz), (T)g.dt);
  }
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// rmsdiv.f90 and scalars_module2.f90:22-70 (FP64 reductions)
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_rmsdiv(const T* du, const T* dv, const T* dw, int nx, int ld, int ny, int nl, double* part) {
  __shared__ double sh[256];
  const long n = (long)nx * ny * nl;
  double acc = 0.0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const unsigned pn = (unsigned)(nx * ny), qq = (unsigned)q;  // nx*ny*nz < 2^31
    const int k = (int)(qq / pn);
    const int r = (int)(qq - (unsigned)k * pn), y = r / nx, x = r - y * nx;
    const long i = ((long)k * ny + y) * ld + x;
    acc += fabs((double)du[i] + (double)dv[i] + (double)dw[i]);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
template <typename T>
double op_rmsdiv(les_ctx* c, const T* dudx, const T* dvdy, const T* dwdz) {
  const Grid& g = c->g;
  const int nb = 512;
  double* part = (double*)scratch(c, S_RED, nb * sizeof(double));
  hipLaunchKernelGGL((k_rmsdiv<T>), dim3(nb), dim3(256), 0, c->stream, dudx, dvdy, dwdz, g.nx, g.ld, g.ny, g.nz - 1,
                     part);
  LCHK(c, hipGetLastError());
  std::vector<double> h(nb);
  LCHK(c, hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  LCHK(c, hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (double x : h) s += x;
  return s / ((double)g.nx * g.ny * (g.nz - 1));
}

template <typename T>
__global__ void k_tsc(const T* u, const T* v, const T* w, const T* nu, int nx, int ld, int ny, int nl,
                      double* part) {
  __shared__ double sh[4][256];
  const long n = (long)nx * ny * nl;
  double m0 = -1.0, m1 = 0.0, m2 = 0.0, m3 = -1e300;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const unsigned pn = (unsigned)(nx * ny), qq = (unsigned)q;  // nx*ny*nz < 2^31
    const int k = (int)(qq / pn);
    const int r = (int)(qq - (unsigned)k * pn), y = r / nx, x = r - y * nx;
    const long i = ((long)k * ny + y) * ld + x;
    const double a = u[i], b = v[i], d = w[i];
    m0 = fmax(m0, a * a + b * b + d * d);
    m1 = fmax(m1, fabs(b));
    m2 = fmax(m2, fabs(d));
    m3 = fmax(m3, (double)nu[i]);
  }
  sh[0][threadIdx.x] = m0; sh[1][threadIdx.x] = m1; sh[2][threadIdx.x] = m2; sh[3][threadIdx.x] = m3;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int j = 0; j < 4; ++j) sh[j][threadIdx.x] = fmax(sh[j][threadIdx.x], sh[j][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int j = 0; j < 4; ++j) part[4 * blockIdx.x + j] = sh[j][0];
}
template <typename T>
void op_timestep_conditions(les_ctx* c, const T* u, const T* v, const T* w, const T* nu, double out[4]) {
  const Grid& g = c->g;
  const int nb = 256;
  double* part = (double*)scratch(c, S_RED, 4 * nb * sizeof(double));
  hipLaunchKernelGGL((k_tsc<T>), dim3(nb), dim3(256), 0, c->stream, u, v, w, nu, g.nx, g.ld, g.ny, g.nz - 1, part);
  LCHK(c, hipGetLastError());
  std::vector<double> h(4 * nb);
  LCHK(c, hipMemcpyAsync(h.data(), part, 4 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  LCHK(c, hipStreamSynchronize(c->stream));
  double m[4] = {-1.0, 0.0, 0.0, -1e300};
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < 4; ++j) m[j] = fmax(m[j], h[4 * b + j]);
  const double delta = fmin(g.dx, g.dy);
  out[0] = sqrt(m[0]) * g.dt / g.dx;
  out[1] = m[1] * g.dt / g.dy;
  out[2] = m[2] * g.dt / g.dz;
  out[3] = g.dt * m[3] / (delta * delta);
}

// ------------------------------------------------------------------------
// explicit instantiations
// ------------------------------------------------------------------------
#define LES_INST(T)                                                                                              \
  template void op_filt_da<T>(les_ctx*, T*, T*, T*, int, int, const cx<T>**);                                   \
  template void op_filt_da3<T>(les_ctx*, T*, T*, T*, T* const[9], int, int, const cx<T>**);                     \
  template void op_ddxy<T>(les_ctx*, T*, T*, const T*, int);                                                     \
  template void op_test_filter<T>(les_ctx*, T*, int, const T*, int);                                             \
  template void op_test_filter_level0<T>(les_ctx*, const T*, const T*, T*, const T*, int);                       \
  template void op_dealias1<T>(les_ctx*, const T*, T*, int);                                                     \
  template void op_dealias2<T>(les_ctx*, T*, T*, int);                                                           \
  template void op_padd<T>(les_ctx*, T*, const T*, int);                                                         \
  template void op_unpadd<T>(les_ctx*, T*, const T*, int);                                                       \
  template void op_ddz_uv<T>(les_ctx*, T*, const T*);                                                            \
  template void op_ddz_w<T>(les_ctx*, T*, const T*);                                                             \
  template void op_ddz3<T>(les_ctx*, T*, T*, T*, const T*, const T*, const T*);                                  \
  template void op_divstress<T>(les_ctx*, T*, const T*, const T*, const T*, int);                                \
  template void op_divstress3<T>(les_ctx*, T* const[3], T* const[6]);                                            \
  template void op_convec<T>(les_ctx*, const T*, const T*, const T*, const T*, const T*, const T*, const T*,     \
                             const T*, const T*, T*, T*, T*);                                                    \
  template void op_convec_ab2<T>(les_ctx*, T*, T*, T*, const T*, const T*, const T*, const T*, const T*,         \
                                 const T*, T*, T*, T*, T*, T*, T*, const T*, const T*, const T*, int, const T*,  \
                                 const ForcingArgs<T>*, const SpecIn<T>*);                                       \
  template void op_wallstress<T>(les_ctx*, const T*, const T*, T*, T*, T*, T*, const double*, const double*,     \
                                 const double*, double*);                                                        \
  template void op_calc_sij<T>(les_ctx*, const T* const[9], T* const[6]);                                        \
  template void op_interpolag<T>(les_ctx*, const SgsPtrs&);                                                      \
  template void op_interpolag4<T>(les_ctx*, T* const[4], T*, T*, T*);                                            \
  template T* op_interpolag4_into<T>(les_ctx*, const T* const[4], T*, T*, T*);                                   \
  template T* lag_filter_mom<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, int);                 \
  template void launch_tridiag<T>(les_ctx*, const cx<T>*, const cx<T>*, cx<T>*, double*, cx<double>*, int*);   \
  template void op_lagrange<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, const SgsPtrs&);       \
  template void op_lagrange_ssim<T>(les_ctx*, const T* const[6], const T*, const T*, const T*, const SgsPtrs&);  \
  template void op_interpolag_ssim<T>(les_ctx*, const SgsPtrs&);                                                 \
  template void op_sgs_stag<T>(les_ctx*, const T* const[9], const T*, const T*, const T*, T* const[6],           \
                               const SgsPtrs&, T*, const ScalarSgsPtrs*);                                        \
  template void op_press_stag<T>(les_ctx*, const T*, const T*, const T*, const T*, T*, T*, T*);                  \
  template void op_project<T>(les_ctx*, T*, T*, T*, const T*, const T*, const T*);                               \
  template double op_rmsdiv<T>(les_ctx*, const T*, const T*, const T*);                                          \
  template void op_timestep_conditions<T>(les_ctx*, const T*, const T*, const T*, const T*, double*);            \
  template void op_ab2<T>(les_ctx*, T*, T*, T*, T*, T*, T*, const T*, const T*, const T*, const T*, const T*,    \
                          const T*, int);                                                                        \
  template void op_pressure_update<T>(les_ctx*, T*, T*, T*, T*, T*, T*, const T*, const T*, const T*, T*);       \
  template void op_fill<T>(les_ctx*, T*, size_t, T);                                                             \
  template void op_copy<T>(les_ctx*, T*, const T*, size_t);

LES_INST(double)
LES_INST(float)

}  // namespace les
