// les_xform.h -- transform kernels shared by the hot-path translation units:
// batched row (x) r2c / c2r transforms with fused loader / storer functors,
// the LDS column (y) kernel with its fused spectral operations, and small
// elementwise helpers.  Everything here is a template or has internal
// linkage, so each .hip file that includes it gets its own instantiations.
//
//   k_row_fwd : real rows (x) -> first `kxo` Fourier coefficients, two real
//               rows packed into one complex FFT (x-spectral rows)
//   k_col     : LDS tile of TX adjacent kx columns x all ky of one plane;
//               forward y-FFT, a fused spectral operation (i kx, i ky, test
//               filter G, 3/2 zero-pad / truncate, divergence), inverse y-FFT
//   k_row_inv : x-spectral rows -> real rows (c2r), with a fused epilogue
#pragma once
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "les_internal.h"

namespace les {

constexpr double BOGUS = -1234567890.0;  // param.f90:50
constexpr double VONK = 0.41;            // param.f90:98
constexpr double TADV1 = 1.5, TADV2 = -0.5;

__device__ __forceinline__ double fq(double x) { return (double)(float)x; }  // Fortran real(x)

// |S| = sqrt(2 S_ij S_ij) in the reference's evaluation order
template <typename T>
__device__ __forceinline__ T smag(T s11, T s12, T s13, T s22, T s23, T s33) {
  return sqrt(T(2) * (s11 * s11 + s22 * s22 + s33 * s33 + T(2) * (s12 * s12 + s13 * s13 + s23 * s23)));
}

#define LCHK(c, call) check((c), (call), #call)

// e / d for 0 <= e, 0 < d and e + d < 2^22 through a float reciprocal:
// (e + 1/2)/d sits at least 1/(2d) from an integer and the two roundings move
// it by less than (e/d + 1) 2^-23 < 1/(2d), so the truncation is exact
__device__ __forceinline__ int div_small(int e, float inv_d) { return (int)(((float)e + 0.5f) * inv_d); }

// Offset of element (y, k) in a plane of ny x-spectral rows of nk
// coefficients: row-major (y nk + k), or tiled, where each 8-coefficient
// column tile of the plane is contiguous ((k / 8) 8 ny + 8 y + k % 8, nk a
// multiple of 8).  In the tiled layout the two rows of a row-kernel pair fill
// whole 128-byte lines of every tile and a column kernel's 8-column tile is
// one contiguous block, where the row-major layout gives a column kernel
// 64-byte row segments (FP32).
__host__ __device__ __forceinline__ long xs_off(int y, int k, int ny, long nk, int tiled) {
  return tiled ? (long)(k >> 3) * (8L * ny) + 8 * y + (k & 7) : (long)y * nk + k;
}


// The r2c split of a packed pair of real rows: coefficient k of row 2pp
// (h = 0) or of row 2pp + 1 (h = 1) from Z_k and Z_{n-k} of their complex FFT.
template <typename T>
__device__ __forceinline__ cx<T> r2c_half(cx<T> zk, cx<T> zn, int h) {
  return h == 0 ? cx<T>{T(0.5) * (zk.x + zn.x), T(0.5) * (zk.y - zn.y)}
                : cx<T>{T(0.5) * (zk.y + zn.y), T(-0.5) * (zk.x - zn.x)};
}
// Tiled output of a row pair: element f of the pair's 2 ceil8(kxo) slots is
// coefficient k = 8 (f / 16) + f % 8 of row 2pp + (f / 8) % 2, at offset
// (f / 16) 8 ny + 16 pp + f % 16 of the plane (xs_off), so 16 consecutive
// threads write one whole 128-byte line (FP32) rather than two halves of it
// in two store instructions.
__device__ __forceinline__ int tiled_k(int f) { return 8 * (f >> 4) + (f & 7); }
__device__ __forceinline__ long tiled_pair_off(int f, int pp, int ny) {
  return (long)(f >> 4) * (8L * ny) + 16 * pp + (f & 15);
}

// the x-spectral layout of rows of nk coefficients: the context's, where nk allows tiles
static inline int xs_layout(const les_ctx* c, long nk) { return c->tune.xs_tiled && nk % 8 == 0 ? 1 : 0; }

template <typename T>
static const cx<T>* tw_for(const les_ctx* c, int n) {
  const Tables<T>& t = tables<T>(c);
  if (n == c->g.nx) return t.tw_nx;
  if (n == c->g.ny) return t.tw_ny;
  if (n == c->g.nx2) return t.tw_nx2;
  return t.tw_ny2;
}
static const FftPlan& plan_for(const les_ctx* c, int n) {
  if (n == c->g.nx) return c->p_nx;
  if (n == c->g.ny) return c->p_ny;
  if (n == c->g.nx2) return c->p_nx2;
  return c->p_ny2;
}

// ------------------------------------------------------------------------
// row transforms
// ------------------------------------------------------------------------
// Rows are grouped by plane (field x level): blockIdx.y = plane,
// blockIdx.x = group of P row pairs inside it, so loaders and storers get a
// block-uniform plane index (no per-element integer division).  N > 0 is the
// compile-time transform length; N == 0 runs the same code with run-time n.
// PLAN: R8 selects radix-8 passes (plan 2) from 64 points, else radix-4
// passes (plan 0).  FP32 forward (r2c) kernels and FP32 inverse kernels with
// plain storers use plan 2; inverse epilogues with a load phase would hold
// twice the points per thread at plan 2's thread count, and FP64 rows measured
// faster with plan 0.  TPF threads per transform (n/8 with plan 2, whose last
// radix-4 or -2 pass takes two butterflies per thread; n/4 with plan 0), P
// transforms per block
template <int N, bool R8 = false>
struct RowGeom {
  static constexpr int PLAN = R8 && N >= 64 ? 2 : 0;
  static constexpr int TPF = N > 0 ? (PLAN == 2 ? N / 8 : N / 4) : 64;
  static constexpr int P = (256 / TPF) > 0 ? 256 / TPF : 1;
  static constexpr int BLOCK = P * TPF;
};
template <typename T, int N>
using RowGeomF = RowGeom<N, sizeof(T) == 4>;  // forward (r2c) row kernels: radix-8 in FP32 (measured)

// Loaders whose value at (plane, y, x) is xf(row_ptr(plane, y)[x]) over rows
// of ld contiguous elements (field reads with a pointwise scaling).  When the launcher has
// found every row aligned to two elements (v2), the forward row kernel loads
// x, x + 1 of a row with one two-element load: half the load instructions,
// the same values.
template <class L, class = void>
struct HasRows : std::false_type {};
template <class L>
struct HasRows<L, std::void_t<decltype(std::declval<const L&>().row_ptr(0, 0))>> : std::true_type {};

template <typename T, int N, class Loader>
__global__ void __launch_bounds__((N > 0 ? RowGeomF<T, N>::BLOCK : 1024))
    k_row_fwd(Loader L, cx<T>* __restrict__ out, long orow, int ny, FftPlan p, const cx<T>* __restrict__ twg,
              int kxo, int tiled) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  const int n = N > 0 ? N : p.n;
  const int P = N > 0 ? RowGeomF<T, N>::P : (int)(blockDim.x * 4) / n;
  const int BLOCK = N > 0 ? RowGeomF<T, N>::BLOCK : (int)blockDim.x;
  constexpr int RS = N > 0 ? lds_row(N) : 0;  // padded row stride (compile-time lengths)
  auto ix = [&](int t, int e) { return N > 0 ? t * RS + lds_pad(e) : t * n + e; };
  cx<T>* buf = tw + n;
  if (N > 0)
    for (int i = threadIdx.x; i < n - 1; i += BLOCK) tw[i] = twg[n + RowGeomF<T, N>::PLAN * n + i];
  else
    for (int i = threadIdx.x; i < n; i += BLOCK) tw[i] = twg[i];
  const int plane = blockIdx.y;
  const int npp = ny / 2;
  const int pair0 = blockIdx.x * P;
  bool staged = false;
  if constexpr (N > 0 && N % 2 == 0 && HasRows<Loader>::value) {
    if (L.v2) {
      constexpr int NE = RowGeomF<T, N>::P * N / 2, BL = RowGeomF<T, N>::BLOCK, MAXE = (NE + BL - 1) / BL;
      using V2 = HIP_vector_type<T, 2>;
      V2 a[MAXE], b[MAXE];
#pragma unroll
      for (int i = 0; i < MAXE; ++i) {
        const int e = threadIdx.x + i * BL;
        const int t = e / (N / 2), x = 2 * (e - t * (N / 2)), pp = pair0 + t;
        a[i] = V2{T(0), T(0)};
        b[i] = a[i];
        if ((NE % BL == 0 || e < NE) && pp < npp) {
          a[i] = *reinterpret_cast<const V2*>(L.row_ptr(plane, 2 * pp) + x);
          b[i] = *reinterpret_cast<const V2*>(L.row_ptr(plane, 2 * pp + 1) + x);
        }
      }
#pragma unroll
      for (int i = 0; i < MAXE; ++i) {
        const int e = threadIdx.x + i * BL;
        const int t = e / (N / 2), x = 2 * (e - t * (N / 2));
        if (NE % BL == 0 || e < NE) {
          buf[ix(t, x)] = cx<T>{L.xf(a[i].x), L.xf(b[i].x)};
          buf[ix(t, x + 1)] = cx<T>{L.xf(a[i].y), L.xf(b[i].y)};
        }
      }
      staged = true;
    }
  }
  if (staged) {
  } else if constexpr (N > 0) {
    // every row load of the block is issued before the first LDS store
    constexpr int NE = RowGeomF<T, N>::P * (N > 0 ? N : 1), BL = RowGeomF<T, N>::BLOCK, MAXE = (NE + BL - 1) / BL;
    cx<T> z[MAXE];
#pragma unroll
    for (int i = 0; i < MAXE; ++i) {
      const int e = threadIdx.x + i * BL;
      const int t = e / N, x = e - t * N, pp = pair0 + t;
      z[i] = {T(0), T(0)};
      if ((NE % BL == 0 || e < NE) && pp < npp) {
        z[i].x = L(plane, 2 * pp, x);
        z[i].y = L(plane, 2 * pp + 1, x);
      }
    }
#pragma unroll
    for (int i = 0; i < MAXE; ++i) {
      const int e = threadIdx.x + i * BL;
      const int t = e / N, x = e - t * N;
      if (NE % BL == 0 || e < NE) buf[ix(t, x)] = z[i];
    }
  } else {
    for (int e = threadIdx.x; e < P * n; e += BLOCK) {
      const int t = e / n, x = e - t * n, pp = pair0 + t;
      cx<T> z = {T(0), T(0)};
      if (pp < npp) {
        z.x = L(plane, 2 * pp, x);
        z.y = L(plane, 2 * pp + 1, x);
      }
      buf[ix(t, x)] = z;
    }
  }
  __syncthreads();
  if constexpr (N > 0)
    fft_lds_c<T, N, 1, RS, RowGeomF<T, N>::P, RowGeomF<T, N>::BLOCK, false, true, RowGeomF<T, N>::PLAN>(buf, tw, -1);
  else
    fft_lds<T, false, 2>(buf, p, 1, n, P, tw, -1);
  if (tiled) {
    // 16-thread groups walk the block's P pairs within a tile first, so a
    // tile's lines from this block (P consecutive pairs) go out together
    const int per = 2 * ((kxo + 7) & ~7);
    cx<T>* op = out + (long)plane * ny * orow;
    for (int e = threadIdx.x; e < P * per; e += BLOCK) {
      const int grp = e >> 4, t = grp % P, f = 16 * (grp / P) + (e & 15), pp = pair0 + t, k = tiled_k(f);
      if (pp >= npp || k >= kxo) continue;
      const cx<T> zk = buf[ix(t, k)], zn = buf[ix(t, k ? n - k : 0)];
      op[tiled_pair_off(f, pp, ny)] = r2c_half(zk, zn, (f >> 3) & 1);
    }
    return;
  }
  const float inv_kxo = 1.0f / (float)kxo;
  for (int e = threadIdx.x; e < P * kxo; e += BLOCK) {
    const int t = div_small(e, inv_kxo), k = e - t * kxo, pp = pair0 + t;
    if (pp >= npp) continue;
    const cx<T> zk = buf[ix(t, k)], zn = buf[ix(t, k ? n - k : 0)];
    const long row = (long)plane * ny + 2 * pp;
    out[row * orow + k] = r2c_half(zk, zn, 0);
    out[(row + 1) * orow + k] = r2c_half(zk, zn, 1);
  }
}

// a storer with a load phase: Pre pre(plane, y, x) reads what the update
// needs, fin(plane, y, x, v, pre) computes and writes
template <class S, class = void>
struct StorerHasPre {
  static constexpr bool value = false;
};
template <class S>
struct StorerHasPre<S, std::void_t<typename S::Pre>> {
  static constexpr bool value = true;
};
// inverse rows: the radix-8 plan (eight transforms per block) in FP32 for
// plain storers; storers with a load phase keep the radix-4 plan, whose
// block of four transforms holds half the points per thread
template <typename T, int N, class S>
using RowGeomI = RowGeom<N, sizeof(T) == 4 && !StorerHasPre<S>::value>;

template <typename T, int N, class Storer>
__global__ void __launch_bounds__((N > 0 ? RowGeomI<T, N, Storer>::BLOCK : 1024))
    k_row_inv(const cx<T>* __restrict__ in, long irow, int kxi, Storer S, int ny, FftPlan p,
              const cx<T>* __restrict__ twg, int ldo, const cx<T>* __restrict__ in2, T c2, int tiled) {
  using RG = RowGeomI<T, N, Storer>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  const int n = N > 0 ? N : p.n;
  const int P = N > 0 ? RG::P : (int)(blockDim.x * 4) / n;
  const int BLOCK = N > 0 ? RG::BLOCK : (int)blockDim.x;
  constexpr int RS = N > 0 ? lds_row(N) : 0;  // padded row stride (compile-time lengths)
  auto ix = [&](int t, int e) { return N > 0 ? t * RS + lds_pad(e) : t * n + e; };
  cx<T>* buf = tw + n;
  if (N > 0)
    for (int i = threadIdx.x; i < n - 1; i += BLOCK) tw[i] = twg[n + RG::PLAN * n + i];
  else
    for (int i = threadIdx.x; i < n; i += BLOCK) tw[i] = twg[i];
  const int plane = blockIdx.y;
  const int npp = ny / 2;
  const int pair0 = blockIdx.x * P;
  // the block's plane (uniform), then 32-bit offsets inside it
  const cx<T>* inp = in + (long)plane * ny * irow;
  const cx<T>* in2p = in2 ? in2 + (long)plane * ny * irow : nullptr;
  // element j of row pair t: z_j = a_j + i b_j (conjugate partners above n/2)
  auto element = [&](int t, int j) {
    const int pp = pair0 + t;
    cx<T> z = {T(0), T(0)};
    if (pp < npp) {
      const int a = 2 * pp, b = a + 1;
      // optional second spectral input added with weight c2 (same layout)
      auto get = [&](int r, int jj) {
        const unsigned o = (unsigned)xs_off(r, jj, ny, irow, tiled);
        cx<T> v = inp[o];
        if (in2) {
          const cx<T> w = in2p[o];
          v = {v.x + c2 * w.x, v.y + c2 * w.y};
        }
        return v;
      };
      if (j == 0) {
        const cx<T> A = get(a, 0), B = get(b, 0);
        z = {A.x, B.x};  // Im of the DC entries is ignored (halfcomplex c2r)
      } else if (j < kxi) {
        const cx<T> A = get(a, j), B = get(b, j);
        z = {A.x - B.y, A.y + B.x};
      } else if (n - j < kxi) {
        const cx<T> A = get(a, n - j), B = get(b, n - j);
        z = {A.x + B.y, B.x - A.y};
      }
    }
    return z;
  };
  if constexpr (N > 0) {
    // every row load of the block is issued before the first LDS store; one
    // item is coefficient j < N/2 of both rows of a pair, loaded once and
    // stored as element j and as its conjugate partner N - j (kxi <= N/2:
    // the Nyquist element N/2 is zero and goes with item 0)
    constexpr int H = (N > 0 ? N : 2) / 2, NI = RG::P * H, BL = RG::BLOCK, MAXI = (NI + BL - 1) / BL;
    cx<T> za[MAXI], zb[MAXI];
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
      const int e = threadIdx.x + i * BL;
      const int t = e / H, j = e - t * H, pp = pair0 + t;
      za[i] = zb[i] = cx<T>{T(0), T(0)};
      if ((NI % BL == 0 || e < NI) && pp < npp && j < kxi) {
        const unsigned a = (unsigned)xs_off(2 * pp, j, ny, irow, tiled), b = (unsigned)xs_off(2 * pp + 1, j, ny, irow, tiled);
        za[i] = inp[a];
        zb[i] = inp[b];
        if (in2) {
          const cx<T> wa = in2p[a], wb = in2p[b];
          za[i] = {za[i].x + c2 * wa.x, za[i].y + c2 * wa.y};
          zb[i] = {zb[i].x + c2 * wb.x, zb[i].y + c2 * wb.y};
        }
      }
    }
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
      const int e = threadIdx.x + i * BL;
      const int t = e / H, j = e - t * H;
      if (NI % BL == 0 || e < NI) {
        const cx<T> A = za[i], B = zb[i];
        if (j == 0) {
          buf[ix(t, 0)] = {A.x, B.x};  // Im of the DC entries is ignored (halfcomplex c2r)
          buf[ix(t, H)] = {T(0), T(0)};
        } else {
          buf[ix(t, j)] = {A.x - B.y, A.y + B.x};
          buf[ix(t, N - j)] = {A.x + B.y, B.x - A.y};
        }
      }
    }
  } else {
    for (int e = threadIdx.x; e < P * n; e += BLOCK) {
      const int t = e / n, j = e - t * n;
      buf[ix(t, j)] = element(t, j);
    }
  }
  __syncthreads();
  if constexpr (N > 0)
    fft_lds_c<T, N, 1, RS, RG::P, RG::BLOCK, false, true, RG::PLAN>(buf, tw, +1);
  else
    fft_lds<T, false, 2>(buf, p, 1, n, P, tw, +1);
  if constexpr (N > 0 && StorerHasPre<Storer>::value) {
    // storers that read the arrays they update: every read of the block's
    // points is issued before the first write (full rows, ldo = N + 2)
    if (ldo == N + 2) {
      constexpr int LDO = N + 2, NS = RG::P * LDO, BL = RG::BLOCK, MS = (NS + BL - 1) / BL;
      // load phases of at most PB elements at a time (about 48 registers of
      // loaded values), so that a large Pre does not cut the waves per SIMD
      constexpr int W = 2 * (int)(sizeof(typename Storer::Pre) / 4);
      constexpr int PB = 48 / W < 1 ? 1 : (48 / W > MS ? MS : 48 / W);
#pragma unroll
      for (int i0 = 0; i0 < MS; i0 += PB) {
        typename Storer::Pre pre[PB][2];
#pragma unroll
        for (int ii = 0; ii < PB; ++ii) {
          const int i = i0 + ii, e = threadIdx.x + i * BL;
          const int t = e / LDO, x = e - t * LDO, pp = pair0 + t;
          if (i < MS && (NS % BL == 0 || e < NS) && pp < npp) {
            pre[ii][0] = S.pre(plane, 2 * pp, x);
            pre[ii][1] = S.pre(plane, 2 * pp + 1, x);
          }
        }
#pragma unroll
        for (int ii = 0; ii < PB; ++ii) {
          const int i = i0 + ii, e = threadIdx.x + i * BL;
          const int t = e / LDO, x = e - t * LDO, pp = pair0 + t;
          if (i < MS && (NS % BL == 0 || e < NS) && pp < npp) {
            T v0 = T(0), v1 = T(0);
            if (x < N) {
              v0 = buf[ix(t, x)].x;
              v1 = buf[ix(t, x)].y;
            }
            S.fin(plane, 2 * pp, x, v0, pre[ii][0]);
            S.fin(plane, 2 * pp + 1, x, v1, pre[ii][1]);
          }
        }
      }
      return;
    }
  }
  if constexpr (N > 0) {
    // the transform's N columns (a whole number of passes over the block),
    // then the ldo - N pad columns, which the storers set to 0
    for (int e = threadIdx.x; e < RG::P * N; e += BLOCK) {
      const int t = e / N, x = e - t * N, pp = pair0 + t;
      if (pp >= npp) continue;
      const cx<T> v = buf[ix(t, x)];
      S(plane, 2 * pp, x, v.x);
      S(plane, 2 * pp + 1, x, v.y);
    }
    const int npad = ldo - N;
    for (int e = threadIdx.x; e < RG::P * npad; e += BLOCK) {
      const int t = e / npad, x = N + (e - t * npad), pp = pair0 + t;
      if (pp >= npp) continue;
      S(plane, 2 * pp, x, T(0));
      S(plane, 2 * pp + 1, x, T(0));
    }
  } else {
    const float inv_ldo = 1.0f / (float)ldo;
    for (int e = threadIdx.x; e < P * ldo; e += BLOCK) {
      const int t = div_small(e, inv_ldo), x = e - t * ldo, pp = pair0 + t;
      if (pp >= npp) continue;
      T v0 = T(0), v1 = T(0);
      if (x < n) {
        v0 = buf[ix(t, x)].x;
        v1 = buf[ix(t, x)].y;
      }
      S(plane, 2 * pp, x, v0);
      S(plane, 2 * pp + 1, x, v1);
    }
  }
}

// compile-time lengths with specialised kernels; anything else runs the
// run-time-length kernels
#define LES_FFT_SIZES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(144) X(192) X(256) X(384)

template <typename T, int N, class L>
static void launch_row_fwd_n(les_ctx* c, const L& ld, cx<T>* out, long orow, long nplanes, int ny, int n, int kxo,
                             int tiled) {
  const FftPlan& p = plan_for(c, n);
  int P = 1, threads = 64;
  if constexpr (N > 0) {
    P = RowGeomF<T, N>::P;
    threads = RowGeomF<T, N>::BLOCK;
  } else {  // run-time length: radix-4/2/3 passes, n/4 threads per transform
    const int tpf = n / 4;
    P = 256 / tpf > 0 ? 256 / tpf : 1;
    threads = P * tpf;
  }
  const long blocks = (ny / 2 + P - 1) / P;
  const size_t sm = (size_t)(n + P * (N > 0 ? lds_row(n) : n)) * sizeof(cx<T>);
  ProfScope ps(c, "row_fwd", (double)nplanes * ny * (n * sizeof(T) * ld.reads() + kxo * sizeof(cx<T>)));
  L lv = ld;
  if constexpr (HasRows<L>::value) lv.v2 = lv.rows_aligned() ? 1 : 0;
  hipLaunchKernelGGL((k_row_fwd<T, N, L>), dim3((unsigned)blocks, (unsigned)nplanes), dim3(threads), sm, c->stream,
                     lv, out, orow, ny, p, tw_for<T>(c, n), kxo, tiled);
  LCHK(c, hipGetLastError());
}

template <typename T, class L>
static void launch_row_fwd(les_ctx* c, const L& ld, cx<T>* out, long orow, long nrows, int n, int kxo, int ny,
                           int tiled = 0) {
  const long nplanes = nrows / ny;
  if (tiled && orow % 8) {
    c->err = "row forward: a tiled x-spectral layout needs rows of a multiple of 8 coefficients";
    return;
  }
  switch (n) {
#define LES_CASE(NN) \
  case NN:           \
    return launch_row_fwd_n<T, NN>(c, ld, out, orow, nplanes, ny, n, kxo, tiled);
    LES_FFT_SIZES(LES_CASE)
#undef LES_CASE
    default:
      return launch_row_fwd_n<T, 0>(c, ld, out, orow, nplanes, ny, n, kxo, tiled);
  }
}

// Row transforms of NF fields formed together: one workgroup per (level,
// row pair); the loader ML forms all NF field values of a point at once
// (ML::all), so the arrays it combines are read once for all fields, where the
// plane-batched k_row_fwd reads them once per field.  Two rows of a field
// share one complex transform as in k_row_fwd (same values, same transform,
// same output rows ((f nz + k) ny + y)).
template <typename T, int N, class ML>
__global__ void __launch_bounds__(256)
    k_row_fwd_mf(ML L, cx<T>* __restrict__ out, long orow, int ny, int nz, const cx<T>* __restrict__ twg,
                 int kxo, int tiled) {
  constexpr int NF = ML::NF, RS = lds_row(N), BLOCK = 256;
  constexpr int NP = 2 * N, MP = (NP + BLOCK - 1) / BLOCK;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* buf = tw + N;
  for (int i = threadIdx.x; i < N - 1; i += BLOCK) tw[i] = twg[N + RowGeomF<T, N>::PLAN * N + i];
  const int k = blockIdx.y, y0 = 2 * blockIdx.x;
  T v[MP][NF];
#pragma unroll
  for (int i = 0; i < MP; ++i) {
    const int e = threadIdx.x + i * BLOCK;
    if (NP % BLOCK == 0 || e < NP) {
      const int r = e / N, x = e - r * N;
      L.all(k, y0 + r, x, v[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < MP; ++i) {
    const int e = threadIdx.x + i * BLOCK;
    if (NP % BLOCK == 0 || e < NP) {
      const int r = e / N, x = e - r * N;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        cx<T>& z = buf[f * RS + lds_pad(x)];
        if (r == 0) z.x = v[i][f];
        else z.y = v[i][f];
      }
    }
  }
  __syncthreads();
  fft_lds_c<T, N, 1, RS, NF, BLOCK, false, true, RowGeomF<T, N>::PLAN>(buf, tw, -1);
  if (tiled) {
    const int per = 2 * ((kxo + 7) & ~7);
    const float inv_per = 1.0f / (float)per;
    for (int e = threadIdx.x; e < NF * per; e += BLOCK) {
      const int f = div_small(e, inv_per), q = e - f * per, kk = tiled_k(q);
      if (kk >= kxo) continue;
      const cx<T> zk = buf[f * RS + lds_pad(kk)], zn = buf[f * RS + lds_pad(kk ? N - kk : 0)];
      out[((long)f * nz + k) * ny * orow + tiled_pair_off(q, blockIdx.x, ny)] = r2c_half(zk, zn, (q >> 3) & 1);
    }
    return;
  }
  const float inv_kxo = 1.0f / (float)kxo;
  for (int e = threadIdx.x; e < NF * kxo; e += BLOCK) {
    const int f = div_small(e, inv_kxo), kk = e - f * kxo;
    const cx<T> zk = buf[f * RS + lds_pad(kk)], zn = buf[f * RS + lds_pad(kk ? N - kk : 0)];
    const long row = ((long)f * nz + k) * ny + y0;
    out[row * orow + kk] = r2c_half(zk, zn, 0);
    out[(row + 1) * orow + kk] = r2c_half(zk, zn, 1);
  }
}

// false when there is no compile-time variant for this row length (or its
// LDS tile does not fit): the caller then runs the plane-batched kernel
// nlev levels (the grid's nz when < 0), output rows ((f nlev + k) ny + y)
template <typename T, class ML>
static bool launch_row_fwd_mf(les_ctx* c, const ML& L, cx<T>* out, int kxo, double read_fields, int nlev = -1,
                              int tiled = 0) {
  const Grid& g = c->g;
  if (tiled && kxo % 8) {
    c->err = "row forward: a tiled x-spectral layout needs rows of a multiple of 8 coefficients";
    return true;
  }
  const int n = g.nx;
  const int nl = nlev < 0 ? g.nz : nlev;
  auto go = [&](auto nn) -> bool {
    constexpr int N = decltype(nn)::value;
    constexpr size_t sm = (size_t)(N + ML::NF * lds_row(N)) * sizeof(cx<T>);
    if constexpr (sm > 160 * 1024) {
      return false;
    } else {
      if (sm > 65536)
        LCHK(c, hipFuncSetAttribute((const void*)k_row_fwd_mf<T, N, ML>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)sm));
      ProfScope ps(c, "row_fwd",
                   (double)nl * g.ny * (n * sizeof(T) * read_fields + 2.0 * ML::NF * kxo * sizeof(cx<T>) / 2));
      hipLaunchKernelGGL((k_row_fwd_mf<T, N, ML>), dim3((unsigned)(g.ny / 2), (unsigned)nl), dim3(256), sm,
                         c->stream, L, out, (long)kxo, g.ny, nl, tw_for<T>(c, N), kxo, tiled);
      LCHK(c, hipGetLastError());
      return true;
    }
  };
  switch (n) {
#define LES_CASE(NN) \
  case NN:           \
    return go(std::integral_constant<int, NN>{});
    LES_FFT_SIZES(LES_CASE)
#undef LES_CASE
    default:
      return false;
  }
}

template <typename T, int N, class S>
static void launch_row_inv_n(les_ctx* c, const cx<T>* in, long irow, int kxi, const S& st, long nplanes, int ny,
                             int n, int ldo, const cx<T>* in2, T c2, int tiled) {
  const FftPlan& p = plan_for(c, n);
  int P = 1, threads = 64;
  if constexpr (N > 0) {
    P = RowGeomI<T, N, S>::P;
    threads = RowGeomI<T, N, S>::BLOCK;
  } else {  // run-time length: radix-4/2/3 passes, n/4 threads per transform
    const int tpf = n / 4;
    P = 256 / tpf > 0 ? 256 / tpf : 1;
    threads = P * tpf;
  }
  if (ldo < n) {  // the output rows hold the n physical values (plus pads)
    c->err = "row inverse: output row shorter than the transform";
    return;
  }
  if (2 * kxi > n) {  // the Nyquist coefficient is never stored
    c->err = "row inverse: more input coefficients than n/2";
    return;
  }
  const long blocks = (ny / 2 + P - 1) / P;
  const size_t sm = (size_t)(n + P * (N > 0 ? lds_row(n) : n)) * sizeof(cx<T>);
  ProfScope ps(c, "row_inv",
               (double)nplanes * ny * (kxi * sizeof(cx<T>) * (in2 ? 2 : 1) + ldo * sizeof(T) * st.writes()));
  hipLaunchKernelGGL((k_row_inv<T, N, S>), dim3((unsigned)blocks, (unsigned)nplanes), dim3(threads), sm, c->stream,
                     in, irow, kxi, st, ny, p, tw_for<T>(c, n), ldo, in2, c2, tiled);
  LCHK(c, hipGetLastError());
}

// in2 (optional): a second x-spectral input of the same layout, added with
// weight c2 before the inverse transform
template <typename T, class S>
static void launch_row_inv(les_ctx* c, const cx<T>* in, long irow, int kxi, const S& st, long nrows, int n,
                           int ldo, int ny, const cx<T>* in2 = nullptr, T c2 = T(1), int tiled = 0) {
  const long nplanes = nrows / ny;
  if (tiled && irow % 8) {
    c->err = "row inverse: a tiled x-spectral layout needs rows of a multiple of 8 coefficients";
    return;
  }
  switch (n) {
#define LES_CASE(NN) \
  case NN:           \
    return launch_row_inv_n<T, NN>(c, in, irow, kxi, st, nplanes, ny, n, ldo, in2, c2, tiled);
    LES_FFT_SIZES(LES_CASE)
#undef LES_CASE
    default:
      return launch_row_inv_n<T, 0>(c, in, irow, kxi, st, nplanes, ny, n, ldo, in2, c2, tiled);
  }
}

// loaders / storers ------------------------------------------------------
// row index = (field * nlev + level) * ny + y; fields are (ld, ny, nlev).
// plane = field * planes_per_field + level; fields are (ld, ny, levels)
template <typename T>
struct LoadFields {
  const T* p[4];
  int ppf, ny, ld;
  T scale;
  int v2 = 0;  // rows aligned to two elements (set by the launcher)
  static constexpr double reads() { return 1.0; }
  __device__ T operator()(int plane, int y, int x) const { return xf(row_ptr(plane, y)[x]); }
  __device__ T xf(T v) const { return scale * v; }
  __device__ const T* row_ptr(int plane, int y) const {
    const int f = plane / ppf, r = plane - f * ppf;
    return p[f] + (long)r * ny * ld + (unsigned)(y * ld);
  }
  bool rows_aligned() const {
    for (const T* q : p)
      if (q && (uintptr_t)q % (2 * sizeof(T))) return false;
    return ld % 2 == 0;
  }
};
template <typename T, int NF>
struct LoadFieldsN {  // up to NF fields, plane = field * ppf + level
  const T* p[NF];
  int ppf, ny, ld;
  T scale;
  int v2 = 0;  // rows aligned to two elements (set by the launcher)
  static constexpr double reads() { return 1.0; }
  __device__ T operator()(int plane, int y, int x) const { return xf(row_ptr(plane, y)[x]); }
  __device__ T xf(T v) const { return scale * v; }
  __device__ const T* row_ptr(int plane, int y) const {
    const int f = plane / ppf, r = plane - f * ppf;
    return p[f] + (long)r * ny * ld + (unsigned)(y * ld);
  }
  bool rows_aligned() const {
    for (const T* q : p)
      if (q && (uintptr_t)q % (2 * sizeof(T))) return false;
    return ld % 2 == 0;
  }
};
template <typename T>
struct StoreFields {
  T* p[12];
  int ppf, ny;
  int ld, nx;  // pads (x >= nx) are written 0
  static constexpr double writes() { return 1.0; }
  __device__ void operator()(int plane, int y, int x, T v) const {
    const int f = plane / ppf, r = plane - f * ppf;
    (p[f] + (long)r * ny * ld)[(unsigned)(y * ld + x)] = x < nx ? v : T(0);
  }
};

// ------------------------------------------------------------------------
// column (y) kernel with fused spectral operation
// ------------------------------------------------------------------------
enum Kind : int { K_ID = 0, K_IKX = 1, K_IKY = 2, K_G = 3 };

struct ColArgs {
  int ny_in, ny_out, do_fwd, do_inv;
  int nki;               // input kx columns present
  int nout;              // 1..3 outputs
  int kind[3];
  int nko[3];            // output kx columns written
  long in_row, in_plane, in2_plane;
  long out_row[3], out_plane[3];
  int lhG;               // row length of the G tables (lh)
  int in_tl, out_tl[3];  // x-spectral layout of the input(s) / each output: 0 row-major, 1 tiled (xs_off)
};

__device__ __forceinline__ int map_row(int yo, int nyi, int nyo) {
  if (nyi == nyo) return (yo == nyo / 2) ? -1 : yo;
  if (nyo > nyi) {  // zero-pad (padd.f90)
    if (yo < nyi / 2) return yo;
    if (yo >= nyo - nyi / 2 + 1) return yo - nyo + nyi;
    return -1;
  }
  if (yo < nyo / 2) return yo;  // truncate (unpadd.f90)
  if (yo == nyo / 2) return -1;
  return yo + nyi - nyo;
}

// XCD-aware tile order.  Workgroups go to the eight XCDs round-robin by linear
// id, and each XCD has its own L2.  Tiles whose rows are 64 bytes come in
// pairs that share every 128-byte line (kx tiles 2j, 2j + 1 of a plane); this
// order gives a pair to ids p and p + 8, which run on the same XCD one after
// the other, so the second tile finds its lines in that L2.  A bijection of
// the first 16 floor(n/16) workgroups of a grid of even width (the rest, and
// grids of odd width, keep their own tile).
__device__ __forceinline__ void xcd_pair_tiles(unsigned& bx, unsigned& by) {
  const unsigned gx = gridDim.x, n16 = (gx * gridDim.y) & ~15u;
  const unsigned p = blockIdx.x + gx * blockIdx.y;
  if ((gx & 1u) != 0 || p >= n16) return;
  const unsigned x = p & 7u, q = p >> 3;
  const unsigned tile = 2u * (x + 8u * (q >> 1)) + (q & 1u);
  bx = tile % gx;
  by = tile / gx;
}

// block: TX * n/8 threads (plans 0 and 2: at most two butterflies per thread
// in any pass).  WIDE (FP32): same-length passes, which keep the spectrum in
// registers, take TX * n/4 threads where that fits 1024, so each thread holds
// half as many tile elements (65 registers instead of 102 at TX 8, n 256:
// filt_da's columns 0.231 -> 0.219 ms at 256^2); in FP64 the wider block was
// slower and is not used
template <int TX, int NYI, int NYO, bool WIDE = false>
struct ColGeom {
  static constexpr int NMAX = NYI > NYO ? NYI : NYO;
  static constexpr int B8 = (TX * NMAX / 8 + 63) / 64 * 64, B4 = (TX * NMAX / 4 + 63) / 64 * 64;
  static constexpr int B0 = WIDE && NYI == NYO && B4 <= 1024 ? B4 : B8;
  static constexpr int BLOCK = B0 < 64 ? 64 : B0;
};
template <typename T, int TX, int NYI, int NYO>
using ColGeomT = ColGeom<TX, NYI, NYO, sizeof(T) == 4>;

// NYI/NYO > 0: compile-time y lengths (block size ColGeom::BLOCK); 0: run time.
// IP: the single identity output of a zero-pad or truncation (padd.f90 /
// unpadd.f90) resampled in place in the input tile: the rows that move never
// overlap their destinations, so the tile needs max(NYI, NYO) rows and no
// second buffer (more workgroups per CU, the same values).
// waves per SIMD the register allocator is asked for: FP32 same-length
// passes of 512 threads fit four workgroups per CU (eight waves per SIMD) at
// 64 registers, three at the 65 they take unasked (a two-register spill is
// the cheaper side: filt_da's columns 0.198 -> 0.186, divstress's 0.086 ->
// 0.078 ms per step); elsewhere no request
template <typename T, int TX, int NYI, int NYO, bool IP, bool IN2>
constexpr int col_wpe() {
  return sizeof(T) == 4 && NYI > 0 && NYI == NYO && !IP && ColGeomT<T, TX, NYI, NYO>::BLOCK == 512 ? 8 : 1;
}

template <typename T, int TX, int NYI, int NYO, int PLAN = 0, bool IP = false, bool IN2 = false>
__global__ void __launch_bounds__((NYI > 0 ? ColGeomT<T, TX, NYI, NYO>::BLOCK : 1024))
    __attribute__((amdgpu_waves_per_eu(col_wpe<T, TX, NYI, NYO, IP, IN2>())))
    k_col(const cx<T>* __restrict__ in, const cx<T>* __restrict__ in2, cx<T>* o0, cx<T>* o1, cx<T>* o2, ColArgs a,
          FftPlan pin, FftPlan pout, const cx<T>* __restrict__ twin, const cx<T>* __restrict__ twout,
          const T* __restrict__ kxt, const T* __restrict__ kyt, const T* __restrict__ G0, const T* __restrict__ G1,
          const T* __restrict__ G2) {
  constexpr bool CT = NYI > 0;
  constexpr int BLK = CT ? ColGeomT<T, TX, NYI, NYO>::BLOCK : 0;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nyi = CT ? NYI : a.ny_in;
  const int nyo = CT ? NYO : a.ny_out;
  const int bd = CT ? BLK : (int)blockDim.x;
  cx<T>* twi = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* two = twi + nyi;
  cx<T>* A = two + nyo;
  cx<T>* B = A + nyi * TX;
  for (int i = threadIdx.x; i < nyi; i += bd)
    twi[i] = CT ? (i < nyi - 1 ? twin[nyi + PLAN * nyi + i] : cx<T>{}) : twin[i];
  for (int i = threadIdx.x; i < nyo; i += bd)
    two[i] = CT ? (i < nyo - 1 ? twout[nyo + PLAN * nyo + i] : cx<T>{}) : twout[i];
  auto fwd = [&](cx<T>* buf) {
    if constexpr (CT)
      fft_lds_c<T, (CT ? NYI : 4), TX, 1, TX, (CT ? BLK : 64), true, false, PLAN>(buf, twi, -1);
    else
      fft_lds<T, true, 4>(buf, pin, TX, 1, TX, twi, -1);
  };
  auto inv = [&](cx<T>* buf) {
    if constexpr (CT)
      fft_lds_c<T, (CT ? NYO : 4), TX, 1, TX, (CT ? BLK : 64), true, false, PLAN>(buf, two, +1);
    else
      fft_lds<T, true, 4>(buf, pout, TX, 1, TX, two, +1);
  };
  unsigned bx = blockIdx.x, by = blockIdx.y;
  if constexpr (TX * sizeof(cx<T>) == 64)
    if (!a.in_tl) xcd_pair_tiles(bx, by);  // two row-major tiles share 128-byte rows
  const int kx0 = bx * TX;
  const long pl = by;
  // input tile -> LDS; with compile-time lengths every load of the tile is
  // issued before the first LDS store, so the HBM latency is paid once
  auto load_tile = [&](cx<T>* dst, const cx<T>* src) {
    if constexpr (CT) {
      constexpr int NI = (CT ? NYI : 1) * TX, MAXI = (NI + (CT ? BLK : 64) - 1) / (CT ? BLK : 64);
      cx<T> rg[MAXI];
#pragma unroll
      for (int i = 0; i < MAXI; ++i) {
        const int e = threadIdx.x + i * BLK;
        rg[i] = cx<T>{T(0), T(0)};
        if (NI % BLK == 0 || e < NI) {
          const int y = e / TX, c = e - y * TX, kk = kx0 + c;
          if (kk < a.nki) rg[i] = src[xs_off(y, kk, nyi, a.in_row, a.in_tl)];
        }
      }
#pragma unroll
      for (int i = 0; i < MAXI; ++i) {
        const int e = threadIdx.x + i * BLK;
        if (NI % BLK == 0 || e < NI) dst[e] = rg[i];
      }
    } else {
      for (int e = threadIdx.x; e < nyi * TX; e += bd) {
        const int y = e / TX, c = e - y * TX, kk = kx0 + c;
        dst[e] = kk < a.nki ? src[xs_off(y, kk, nyi, a.in_row, a.in_tl)] : cx<T>{T(0), T(0)};
      }
    }
  };
  if constexpr (IP && NYI > 0 && NYI != NYO) {
    cx<T>* R = two + nyo;  // TX * max(NYI, NYO), in place of A and B
    load_tile(R, in + pl * a.in_plane);
    __syncthreads();
    fwd(R);
    if constexpr (NYO > NYI) {
      // rows NYI/2+1 .. NYI-1 move up by NYO-NYI, then rows NYI/2 .. NYO-NYI/2 are zero
      constexpr int NMV = (NYI / 2 - 1) * TX, NZR = (NYO - NYI + 1) * TX;
      for (int e = threadIdx.x; e < NMV; e += BLK) {
        const int yi = NYI / 2 + 1 + e / TX, c = e % TX;
        R[(yi + NYO - NYI) * TX + c] = R[yi * TX + c];
      }
      __syncthreads();
      for (int e = threadIdx.x; e < NZR; e += BLK) R[(NYI / 2) * TX + e] = cx<T>{T(0), T(0)};
    } else {
      // rows NYI-NYO/2+1 .. NYI-1 move down by NYI-NYO, row NYO/2 is zero
      constexpr int NMV = (NYO / 2 - 1) * TX;
      for (int e = threadIdx.x; e < NMV; e += BLK) {
        const int yo = NYO / 2 + 1 + e / TX, c = e % TX;
        R[yo * TX + c] = R[(yo + NYI - NYO) * TX + c];
      }
      for (int e = threadIdx.x; e < TX; e += BLK) R[(NYO / 2) * TX + e] = cx<T>{T(0), T(0)};
    }
    __syncthreads();
    inv(R);
    for (int e = threadIdx.x; e < NYO * TX; e += BLK) {
      const int y = e / TX, c = e - y * TX, kk = kx0 + c;
      if (kk < a.nko[0]) o0[pl * a.out_plane[0] + xs_off(y, kk, nyo, a.out_row[0], a.out_tl[0])] = R[e];
    }
    return;
  }
  if constexpr (CT && NYI == NYO && !IP) {
    // same-length passes: the spectrum stays in registers (each thread its
    // own tile elements) and one LDS tile serves every transform
    constexpr int NE = (CT ? NYI : 1) * TX, ME = (NE + (CT ? BLK : 64) - 1) / (CT ? BLK : 64);
    cx<T>* R = two + nyo;
    cx<T> ra[ME], rb[ME];
    // offset of tile element e = y TX + c in a plane of nk-coefficient rows
    // of layout tl: with TX = 8 and tiles, the tile is contiguous
    auto eoff = [&](int e, long nk, int ny_, int tl) -> long {
      if (TX == 8 && tl) return (long)(kx0 >> 3) * (8L * ny_) + e;
      const int y = e / TX, c = e - y * TX;
      return xs_off(y, kx0 + c, ny_, nk, tl);
    };
    auto fetch = [&](cx<T> (&r)[ME], const cx<T>* src) {
#pragma unroll
      for (int i = 0; i < ME; ++i) {
        const int e = threadIdx.x + i * BLK;
        r[i] = cx<T>{T(0), T(0)};
        if ((NE % BLK == 0 || e < NE) && kx0 + e % TX < a.nki) r[i] = src[eoff(e, a.in_row, nyi, a.in_tl)];
      }
    };
    auto put = [&](const cx<T> (&r)[ME]) {
#pragma unroll
      for (int i = 0; i < ME; ++i) {
        const int e = threadIdx.x + i * BLK;
        if (NE % BLK == 0 || e < NE) R[e] = r[i];
      }
    };
    auto get = [&](cx<T> (&r)[ME]) {
#pragma unroll
      for (int i = 0; i < ME; ++i) {
        const int e = threadIdx.x + i * BLK;
        if (NE % BLK == 0 || e < NE) r[i] = R[e];
      }
    };
    fetch(ra, in + pl * a.in_plane);
    if (IN2 && in2) fetch(rb, in2 + pl * a.in2_plane);
    if (a.do_fwd) {
      put(ra);
      __syncthreads();
      fwd(R);
      get(ra);
    }
    if (IN2 && in2) {
      // divergence mode: out = i kx FFT(in) + i ky FFT(in2)
      const int c0 = threadIdx.x % TX;
      const T kx = kx0 + c0 < a.nki ? kxt[kx0 + c0] : T(0);
      put(rb);
      __syncthreads();
      fwd(R);
#pragma unroll
      for (int i = 0; i < ME; ++i) {
        const int e = threadIdx.x + i * BLK;
        if (NE % BLK == 0 || e < NE) {
          const int y = e / TX, kk = kx0 + c0;
          cx<T> r = {T(0), T(0)};
          if (y != NYI / 2 && kk < a.nki) {
            const T ky = kyt[y];
            const cx<T> X = ra[i], Y = R[e];
            r = {-kx * X.y - ky * Y.y, kx * X.x + ky * Y.x};
          }
          R[e] = r;
        }
      }
      __syncthreads();
      if (a.do_inv) inv(R);
      for (int e = threadIdx.x; e < NYO * TX; e += BLK)
        if (kx0 + e % TX < a.nko[0]) o0[pl * a.out_plane[0] + eoff(e, a.out_row[0], nyo, a.out_tl[0])] = R[e];
      return;
    }
    // an i kx output beside an identity output is i kx times the identity
    // output's inverse (kx is constant along a column): no transform of its own
    int jid = -1, jkx = -1;
    for (int j = 0; j < a.nout; ++j) {
      if (a.kind[j] == K_ID && jid < 0) jid = j;
      if (a.kind[j] == K_IKX && jkx < 0) jkx = j;
    }
    if (jid < 0 || !a.do_inv) jkx = -1;
    for (int j = 0; j < a.nout; ++j) {
      if (kx0 >= a.nko[j] || j == jkx) continue;  // uniform over the block
      const int kind = a.kind[j];
      const T* G = j == 0 ? G0 : (j == 1 ? G1 : G2);
#pragma unroll
      for (int i = 0; i < ME; ++i) {
        const int e = threadIdx.x + i * BLK;
        if (NE % BLK == 0 || e < NE) {
          const int yo = e / TX, c = e - yo * TX, kk = kx0 + c;
          cx<T> v = {T(0), T(0)};
          if (yo != NYO / 2 && kk < a.nki) {  // map_row for equal lengths: the Nyquist row is zero
            v = ra[i];
            if (kind == K_IKX) {
              const T k = kxt[kk];
              v = {-k * v.y, k * v.x};
            } else if (kind == K_IKY) {
              const T k = kyt[yo];
              v = {-k * v.y, k * v.x};
            } else if (kind == K_G) {
              const T g = G[(long)yo * a.lhG + kk];
              v = {g * v.x, g * v.y};
            }
          }
          R[e] = v;
        }
      }
      __syncthreads();
      if (a.do_inv) inv(R);
      cx<T>* o = j == 0 ? o0 : (j == 1 ? o1 : o2);
      for (int e = threadIdx.x; e < NYO * TX; e += BLK)
        if (kx0 + e % TX < a.nko[j]) o[pl * a.out_plane[j] + eoff(e, a.out_row[j], nyo, a.out_tl[j])] = R[e];
      if (j == jid && jkx >= 0 && kx0 < a.nko[jkx]) {
        cx<T>* ox = jkx == 0 ? o0 : (jkx == 1 ? o1 : o2);
        const int c0 = threadIdx.x % TX;
        const T k = kx0 + c0 < a.nki ? kxt[kx0 + c0] : T(0);
        for (int e = threadIdx.x; e < NYO * TX; e += BLK) {
          const cx<T> v = R[e];
          if (kx0 + c0 < a.nko[jkx])
            ox[pl * a.out_plane[jkx] + eoff(e, a.out_row[jkx], nyo, a.out_tl[jkx])] = cx<T>{-k * v.y, k * v.x};
        }
      }
    }
    return;
  }
  load_tile(A, in + pl * a.in_plane);
  if (IN2 && in2) load_tile(B, in2 + pl * a.in2_plane);
  __syncthreads();
  if (a.do_fwd) fwd(A);
  if (IN2 && in2) {
    // divergence mode: out = i kx FFT(in) + i ky FFT(in2)
    if constexpr (CT) {
      // wavenumbers first (a thread keeps one column), then the transform
      constexpr int NE = (CT ? NYI : 1) * TX, ME = (NE + (CT ? BLK : 64) - 1) / (CT ? BLK : 64);
      const int c0 = threadIdx.x % TX;
      const T kx = kx0 + c0 < a.nki ? kxt[kx0 + c0] : T(0);
      T ky[ME];
#pragma unroll
      for (int i = 0; i < ME; ++i) {
        const int e = threadIdx.x + i * BLK;
        ky[i] = (NE % BLK == 0 || e < NE) ? kyt[e / TX] : T(0);
      }
      fwd(B);
#pragma unroll
      for (int i = 0; i < ME; ++i) {
        const int e = threadIdx.x + i * BLK;
        if (NE % BLK == 0 || e < NE) {
          const int y = e / TX, kk = kx0 + c0;
          cx<T> r = {T(0), T(0)};
          if (y != NYI / 2 && kk < a.nki) {
            const cx<T> X = A[e], Y = B[e];
            r = {-kx * X.y - ky[i] * Y.y, kx * X.x + ky[i] * Y.x};
          }
          A[e] = r;
        }
      }
    } else {
      fwd(B);
      for (int e = threadIdx.x; e < nyi * TX; e += bd) {
        const int y = e / TX, c = e - y * TX, kk = kx0 + c;
        cx<T> r = {T(0), T(0)};
        if (y != nyi / 2 && kk < a.nki) {
          const T kx = kxt[kk], ky = kyt[y];
          const cx<T> X = A[e], Y = B[e];
          r = {-kx * X.y - ky * Y.y, kx * X.x + ky * Y.x};
        }
        A[e] = r;
      }
    }
    __syncthreads();
    if (a.do_inv) inv(A);
    for (int e = threadIdx.x; e < nyo * TX; e += bd) {
      const int y = e / TX, c = e - y * TX, kk = kx0 + c;
      if (kk < a.nko[0]) o0[pl * a.out_plane[0] + xs_off(y, kk, nyo, a.out_row[0], a.out_tl[0])] = A[e];
    }
    return;
  }
  for (int j = 0; j < a.nout; ++j) {
    if (kx0 >= a.nko[j]) continue;  // uniform over the block
    const int kind = a.kind[j];
    const T* G = j == 0 ? G0 : (j == 1 ? G1 : G2);
    for (int e = threadIdx.x; e < nyo * TX; e += bd) {
      const int yo = e / TX, c = e - yo * TX, kk = kx0 + c;
      const int yi = map_row(yo, nyi, nyo);
      cx<T> v = {T(0), T(0)};
      if (yi >= 0 && kk < a.nki) {
        v = A[yi * TX + c];
        if (kind == K_IKX) {
          const T k = kxt[kk];
          v = {-k * v.y, k * v.x};
        } else if (kind == K_IKY) {
          const T k = kyt[yo];
          v = {-k * v.y, k * v.x};
        } else if (kind == K_G) {
          const T g = G[(long)yo * a.lhG + kk];
          v = {g * v.x, g * v.y};
        }
      }
      B[e] = v;
    }
    __syncthreads();
    if (a.do_inv) inv(B);
    cx<T>* o = j == 0 ? o0 : (j == 1 ? o1 : o2);
    for (int e = threadIdx.x; e < nyo * TX; e += bd) {
      const int y = e / TX, c = e - y * TX, kk = kx0 + c;
      if (kk < a.nko[j]) o[pl * a.out_plane[j] + xs_off(y, kk, nyo, a.out_row[j], a.out_tl[j])] = B[e];
    }
    __syncthreads();
  }
}

// column tile width: TX kx columns per block (a 64-lane wavefront reads
// TX*sizeof(cx<T>) contiguous bytes per row); the default is 128 B rows,
// les_autotune may pick another (les_coltx.hip holds those variants)
template <typename T>
constexpr int col_tx_default() { return sizeof(T) == 8 ? 8 : 16; }
template <typename T, int TX, int NYI, int NYO>
constexpr bool col_tile_fits() {
  return ColGeomT<T, TX, NYI, NYO>::BLOCK <= 1024 &&
         (size_t)(NYI + NYO + TX * NYI + TX * NYO) * sizeof(cx<T>) <= 160 * 1024;
}

template <typename T, int TX, int NYI, int NYO, int PLAN = 0, bool IP = false>
static void launch_col_n(les_ctx* c, const cx<T>* in, const cx<T>* in2, cx<T>* o0, cx<T>* o1, cx<T>* o2,
                         const ColArgs& a, long nplanes, int ncols, const T* G0, const T* G1, const T* G2) {
  const int nmax = a.ny_in > a.ny_out ? a.ny_in : a.ny_out;
  // the FFT stages need blockDim >= TX*n/8 (at most 4 butterflies per thread)
  int threads = NYI > 0 ? ColGeomT<T, TX, NYI, NYO>::BLOCK : (TX * nmax / 8 + 63) / 64 * 64;
  if (threads < 64) threads = 64;
  if (threads > 1024) {
    c->err = "k_col: column tile too large";
    return;
  }
  // LDS tiles: in-place resampling and same-length passes (spectrum kept in
  // registers) use one, the rest two
  constexpr bool ONE = NYI > 0 && (IP || NYI == NYO);
  const size_t sm = (size_t)(a.ny_in + a.ny_out + (ONE ? TX * nmax : TX * a.ny_in + TX * a.ny_out)) * sizeof(cx<T>);
  // the divergence mode (a second input) only with equal lengths, in its own
  // instantiation
  constexpr bool CAN2 = NYI == NYO && !IP;
  if (!CAN2 && in2) {
    c->err = "k_col: a second input needs equal lengths";
    return;
  }
  const void* kern = CAN2 && in2 ? (const void*)k_col<T, TX, NYI, NYO, PLAN, IP, CAN2>
                                 : (const void*)k_col<T, TX, NYI, NYO, PLAN, IP, false>;
  if (sm > 65536) LCHK(c, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  const Tables<T>& t = tables<T>(c);
  const unsigned gx = (unsigned)((ncols + TX - 1) / TX);
  double bytes = 0;
  {
    const double es = 2.0 * sizeof(T);
    bytes += (double)nplanes * a.ny_in * (ncols < a.nki ? ncols : a.nki) * es * (in2 ? 2 : 1);
    for (int j = 0; j < (in2 ? 1 : a.nout); ++j) bytes += (double)nplanes * a.ny_out * a.nko[j] * es;
  }
  ProfScope ps(c, a.ny_in < a.ny_out ? "col_pad" : (a.ny_in > a.ny_out ? "col_trunc" : "col"), bytes);
  if (CAN2 && in2)
    hipLaunchKernelGGL((k_col<T, TX, NYI, NYO, PLAN, IP, CAN2>), dim3(gx, (unsigned)nplanes), dim3(threads), sm,
                       c->stream, in, in2, o0, o1, o2, a, plan_for(c, a.ny_in), plan_for(c, a.ny_out),
                       tw_for<T>(c, a.ny_in), tw_for<T>(c, a.ny_out), t.kx, t.ky, G0, G1, G2);
  else
    hipLaunchKernelGGL((k_col<T, TX, NYI, NYO, PLAN, IP, false>), dim3(gx, (unsigned)nplanes), dim3(threads), sm,
                       c->stream, in, nullptr, o0, o1, o2, a, plan_for(c, a.ny_in), plan_for(c, a.ny_out),
                       tw_for<T>(c, a.ny_in), tw_for<T>(c, a.ny_out), t.kx, t.ky, G0, G1, G2);
  LCHK(c, hipGetLastError());
}

#define LES_COL_PAIRS(X)                                                                                       \
  X(16, 16) X(16, 24) X(24, 16) X(32, 32) X(32, 48) X(48, 32) X(64, 64) X(64, 96) X(96, 64) X(96, 96) X(96, 144) \
  X(144, 96) X(128, 128) X(128, 192) X(192, 128) X(256, 256) X(256, 384) X(384, 256)

// zero-pad / truncation passes with a single identity output: the in-place
// tile at 128-byte rows (16 FP32 / 8 FP64 columns); FP32 in the radix-8 plan
// (at 256^2 its twelve-field zero-pad takes 0.216 ms against 0.261 with the
// radix-4 plan, whichever plan suits the same-length passes), FP64 in the
// plan the autotuner chose; false for lengths without a compile-time variant
#define LES_COL_RESAMPLE(X) \
  X(16, 24) X(24, 16) X(32, 48) X(48, 32) X(64, 96) X(96, 64) X(96, 144) X(144, 96) X(128, 192) X(192, 128) \
  X(256, 384) X(384, 256)
template <typename T>
static bool launch_col_ip(les_ctx* c, int plan, const cx<T>* in, cx<T>* o0, const ColArgs& a, long nplanes,
                          int ncols) {
  constexpr int TXI = sizeof(T) == 8 ? 8 : 16;
#define LES_CASE(I, O)                                                                                             \
  if (a.ny_in == I && a.ny_out == O) {                                                                             \
    if (plan == 2)                                                                                                 \
      launch_col_n<T, TXI, I, O, 2, true>(c, in, nullptr, o0, nullptr, nullptr, a, nplanes, ncols, nullptr,      \
                                          nullptr, nullptr);                                                       \
    else                                                                                                           \
      launch_col_n<T, TXI, I, O, 0, true>(c, in, nullptr, o0, nullptr, nullptr, a, nplanes, ncols, nullptr,      \
                                          nullptr, nullptr);                                                       \
    return true;                                                                                                   \
  }
  LES_COL_RESAMPLE(LES_CASE)
#undef LES_CASE
  return false;
}

// the tuned tile widths / FFT plans other than the default (les_coltx.hip);
// false when that variant does not exist for these lengths
template <typename T>
bool launch_col_alt(les_ctx* c, int tx, int plan, const cx<T>* in, const cx<T>* in2, cx<T>* o0, cx<T>* o1,
                    cx<T>* o2, const ColArgs& a, long nplanes, int ncols, const T* G0, const T* G1, const T* G2);

template <typename T>
static void launch_col(les_ctx* c, const cx<T>* in, const cx<T>* in2, cx<T>* o0, cx<T>* o1, cx<T>* o2,
                       const ColArgs& a, long nplanes, int ncols, const T* G0 = nullptr, const T* G1 = nullptr,
                       const T* G2 = nullptr) {
  constexpr int TXD = col_tx_default<T>();
  const int tx = c->tune.col_tx[sizeof(T) == 8 ? 1 : 0], plan = c->tune.col_plan[sizeof(T) == 8 ? 1 : 0];
  if (a.nout == 1 && a.kind[0] == K_ID && !in2 && a.do_fwd && a.do_inv && a.ny_in != a.ny_out &&
      launch_col_ip<T>(c, sizeof(T) == 4 ? 2 : plan, in, o0, a, nplanes, ncols))
    return;
  if ((tx != TXD || plan != 0) && launch_col_alt<T>(c, tx, plan, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2))
    return;
#define LES_CASE(I, O)                                                                        \
  if (a.ny_in == I && a.ny_out == O)                                                          \
    return launch_col_n<T, TXD, I, O>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  LES_COL_PAIRS(LES_CASE)
#undef LES_CASE
  launch_col_n<T, TXD, 0, 0>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
}

// ------------------------------------------------------------------------
// Zero-pad column pass (padd.f90 on the y direction) whose input planes a
// functor forms: fn.comb(plane) gives the weighted x-spectral planes (rows of
// nki coefficients) whose sum is the input.  Same
// tile, transforms and output layout as the in-place zero-pad pass of k_col
// (the FP32 radix-8 plan, FP64 the tuned one), so the convective inputs can
// be formed from filt_da's x-spectral rows instead of forward-transforming
// physical fields again.
// ------------------------------------------------------------------------
template <typename T, int TX, int NYI, int NYO, int PLAN, class Fn>
__global__ void __launch_bounds__((ColGeomT<T, TX, NYI, NYO>::BLOCK))
    k_col_pad_fn(Fn fn, cx<T>* __restrict__ out, int nki, long out_row, long out_plane,
                 const cx<T>* __restrict__ twin, const cx<T>* __restrict__ twout, int in_tl) {
  static_assert(NYO > NYI && NYO < 2 * NYI, "3/2 zero-pad pass");
  constexpr int BLK = ColGeomT<T, TX, NYI, NYO>::BLOCK;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* twi = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* two = twi + NYI;
  cx<T>* R = two + NYO;  // TX * NYO
  for (int i = threadIdx.x; i < NYI; i += BLK) twi[i] = i < NYI - 1 ? twin[NYI + PLAN * NYI + i] : cx<T>{};
  for (int i = threadIdx.x; i < NYO; i += BLK) two[i] = i < NYO - 1 ? twout[NYO + PLAN * NYO + i] : cx<T>{};
  unsigned bx = blockIdx.x, by = blockIdx.y;
  if constexpr (TX * sizeof(cx<T>) == 64) xcd_pair_tiles(bx, by);
  const int kx0 = bx * TX;
  const long pl = by;
  {
    constexpr int NI = NYI * TX, MAXI = (NI + BLK - 1) / BLK;
    const auto lc = fn.comb(pl);
    // every load of the tile issued before the first combination
    cx<T> raw[MAXI][4];
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
      const int e = threadIdx.x + i * BLK;
      const int y = e / TX, c = e - y * TX, kk = kx0 + c;
      const bool ok = (NI % BLK == 0 || e < NI) && kk < nki;
      const long off = ok ? xs_off(y, kk, NYI, nki, in_tl) : 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) raw[i][q] = (ok && q < lc.n) ? lc.p[q][off] : cx<T>{T(0), T(0)};
    }
    cx<T> rg[MAXI];
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
      rg[i] = cx<T>{T(0), T(0)};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        rg[i].x += lc.c[q] * raw[i][q].x;
        rg[i].y += lc.c[q] * raw[i][q].y;
      }
    }
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
      const int e = threadIdx.x + i * BLK;
      if (NI % BLK == 0 || e < NI) R[e] = rg[i];
    }
  }
  __syncthreads();
  fft_lds_c<T, NYI, TX, 1, TX, BLK, true, false, PLAN>(R, twi, -1);
  // rows NYI/2+1 .. NYI-1 move up by NYO-NYI (never onto a source row), then
  // rows NYI/2 .. NYO-NYI/2 are zero
  constexpr int NMV = (NYI / 2 - 1) * TX, NZR = (NYO - NYI + 1) * TX;
  for (int e = threadIdx.x; e < NMV; e += BLK) {
    const int yi = NYI / 2 + 1 + e / TX, c = e % TX;
    R[(yi + NYO - NYI) * TX + c] = R[yi * TX + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NZR; e += BLK) R[(NYI / 2) * TX + e] = cx<T>{T(0), T(0)};
  __syncthreads();
  fft_lds_c<T, NYO, TX, 1, TX, BLK, true, false, PLAN>(R, two, +1);
  for (int e = threadIdx.x; e < NYO * TX; e += BLK) {
    const int y = e / TX, c = e - y * TX, kk = kx0 + c;
    if (kk < nki) out[pl * out_plane + (long)y * out_row + kk] = R[e];
  }
}

template <typename T, int TX, int NYI, int NYO, int PLAN, class Fn>
static bool launch_col_pad_fn_n(les_ctx* c, const Fn& fn, cx<T>* out, long nplanes, int nki, long out_row,
                                long out_plane, double reads_per_elem, int in_tl) {
  using G = ColGeomT<T, TX, NYI, NYO>;
  const size_t sm = (size_t)(NYI + NYO + TX * NYO) * sizeof(cx<T>);
  if (sm > 65536)
    LCHK(c, hipFuncSetAttribute((const void*)k_col_pad_fn<T, TX, NYI, NYO, PLAN, Fn>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  const unsigned gx = (unsigned)((nki + TX - 1) / TX);
  const double es = 2.0 * sizeof(T);
  ProfScope ps(c, "col_pad", (double)nplanes * nki * es * (NYI * reads_per_elem + NYO));
  hipLaunchKernelGGL((k_col_pad_fn<T, TX, NYI, NYO, PLAN, Fn>), dim3(gx, (unsigned)nplanes), dim3(G::BLOCK), sm,
                     c->stream, fn, out, nki, out_row, out_plane, tw_for<T>(c, NYI), tw_for<T>(c, NYO), in_tl);
  LCHK(c, hipGetLastError());
  return true;
}

// false where no compile-time variant exists for (ny, 3 ny / 2)
template <typename T, class Fn>
static bool launch_col_pad_fn(les_ctx* c, const Fn& fn, cx<T>* out, long nplanes, int nki, long out_row,
                              long out_plane, double reads_per_elem, int in_tl = 0) {
  constexpr int TX = sizeof(T) == 8 ? 8 : 16;
  const int plan = sizeof(T) == 4 ? 2 : c->tune.col_plan[1];
#define LES_CASE(I, O)                                                                                        \
  if (c->g.ny == I && c->g.ny2 == O) {                                                                         \
    if (plan == 2)                                                                                             \
      return launch_col_pad_fn_n<T, TX, I, O, 2>(c, fn, out, nplanes, nki, out_row, out_plane, reads_per_elem, \
                                                 in_tl);                                                          \
    return launch_col_pad_fn_n<T, TX, I, O, 0>(c, fn, out, nplanes, nki, out_row, out_plane, reads_per_elem,   \
                                               in_tl);                                                            \
  }
  LES_CASE(16, 24) LES_CASE(32, 48) LES_CASE(64, 96) LES_CASE(96, 144) LES_CASE(128, 192) LES_CASE(256, 384)
#undef LES_CASE
  return false;
}

// The convective inputs of convec.f90:45-114 (u, v, w, the three vorticity
// components; with NS = 2 also dtheta/dx, dy, dz and dq/dx, dy, dz), as the
// x-spectral rows the zero-pad pass reads (scale 1/(nx ny) of the physical
// fields), formed from filt_da's x-spectral rows of the filtered fields
// (kinds f, d/dx, d/dy; scale 1/nx) instead of a forward transform of the
// physical fields.  The vertical differences of ddz_uv.f90 and scalar_dz are
// linear, so they are taken on the x-spectral rows; the planes those
// differences do not give (the wall values of dudz, dvdz from wallstress, the
// surface and lid gradients of the scalars) come from one small forward
// transform (W).  Plane order: field-major, (f * nz + k).
// up to four x-spectral planes (rows of nxh coefficients, ny rows each) and
// real weights: one input plane of the zero-pad pass as a linear combination
template <typename T>
struct LinComb4 {
  const cx<T>* p[4];
  T c[4];
  int n;
  __device__ cx<T> at(long off) const {
    cx<T> v{T(0), T(0)};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < n) {
        const cx<T> a = p[q][off];
        v.x += c[q] * a.x;
        v.y += c[q] * a.y;
      }
    return v;
  }
};

template <typename T, int NS>
struct ConvecSpecIn {
  const cx<T>* U;      // filt_da of u, v, w: rows (((j * 3 + f) * nz + k) * ny + y), j = f, d/dx, d/dy
  const cx<T>* S[2];   // filt_da of theta, q: rows ((j * nz + k) * ny + y)
  const cx<T>* W;      // rows (p * ny + y): dudz(0), dvdz(0), then per scalar its dz at levels 0 and nz-1
  int nz, ny, nxh;
  T s, rdz;            // 1 / ny, 1 / dz
  // the weights of input plane `plane` (uniform over a workgroup)
  __device__ LinComb4<T> comb(long plane) const {
    const int f = (int)(plane / nz), k = (int)(plane - (long)f * nz);
    const long pl = (long)ny * nxh;
    auto A = [&](int j, int fi, int kk) { return U + (((long)j * 3 + fi) * nz + kk) * pl; };
    auto Wp = [&](int p) { return W + (long)p * pl; };
    const T h = T(0.5) * s, sd = s * rdz;
    LinComb4<T> L{{nullptr, nullptr, nullptr, nullptr}, {T(0), T(0), T(0), T(0)}, 0};
    auto add = [&](const cx<T>* p, T c) { L.p[L.n] = p; L.c[L.n] = c; ++L.n; };
    if (f < 3) {
      add(A(0, f, k), s);
    } else if (f == 3) {  // dw/dy - dv/dz (dw/dy averaged onto the first uv-node; dv/dz = 0 at the lid)
      if (k == 0) { add(A(2, 2, 0), h); add(A(2, 2, 1), h); add(Wp(1), -s); }
      else if (k == nz - 1) add(A(2, 2, k), s);
      else { add(A(2, 2, k), s); add(A(0, 1, k), -sd); add(A(0, 1, k - 1), sd); }
    } else if (f == 4) {  // du/dz - dw/dx
      if (k == 0) { add(Wp(0), s); add(A(1, 2, 0), -h); add(A(1, 2, 1), -h); }
      else if (k == nz - 1) add(A(1, 2, k), -s);
      else { add(A(0, 0, k), sd); add(A(0, 0, k - 1), -sd); add(A(1, 2, k), -s); }
    } else if (f == 5) {  // dv/dx - du/dy
      add(A(1, 1, k), s);
      add(A(2, 0, k), -s);
    } else if constexpr (NS > 0) {
      const int q = (f - 6) / 3, d = (f - 6) - 3 * q;
      auto Sj = [&](int j, int kk) { return S[q] + ((long)j * nz + kk) * pl; };
      if (d < 2) add(Sj(d + 1, k), s);
      else if (k == 0) add(Wp(2 + 2 * q), s);
      else if (k == nz - 1) add(Wp(3 + 2 * q), s);
      else { add(Sj(0, k), sd); add(Sj(0, k - 1), -sd); }
    }
    return L;
  }
};

// the physical planes ConvecSpecIn takes from W, forward-transformed with the
// filt_da scale 1/nx
template <typename T>
struct LoadPlanes6 {
  const T* p[6];
  int ld;
  T scale;
  static constexpr double reads() { return 1.0; }
  __device__ T operator()(int plane, int y, int x) const { return scale * p[plane][(unsigned)(y * ld + x)]; }
};

static ColArgs col_args(int ny_in, int ny_out, int nki, long in_row, long in_plane) {
  ColArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ny_in = ny_in;
  a.ny_out = ny_out;
  a.do_fwd = 1;
  a.do_inv = 1;
  a.nki = nki;
  a.in_row = in_row;
  a.in_plane = in_plane;
  return a;
}
static void col_out(ColArgs& a, int j, int kind, int nko, long out_row, long out_plane) {
  a.kind[j] = kind;
  a.nko[j] = nko;
  a.out_row[j] = out_row;
  a.out_plane[j] = out_plane;
  if (a.nout < j + 1) a.nout = j + 1;
}

// ------------------------------------------------------------------------
// simple elementwise kernels
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* a, size_t n, T v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}
// the same value into the first n entries of up to 9 arrays, one launch
template <typename T>
struct FillPtrs {
  T* p[9];
};
template <typename T>
__global__ void k_fill_n(FillPtrs<T> a, int cnt, unsigned n, T v) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (int)blockIdx.y < cnt) a.p[blockIdx.y][i] = v;
}
template <typename T>
void op_fill_n(les_ctx* c, T* const* ptrs, int cnt, size_t n, T val) {
  if (!n || cnt <= 0) return;
  if (cnt > 9 || n >= (1u << 31)) {
    c->err = "op_fill_n: at most 9 arrays below 2^31 entries";
    return;
  }
  FillPtrs<T> a{};
  for (int q = 0; q < cnt; ++q) a.p[q] = ptrs[q];
  {
    ProfScope _ps(c, "fill", (double)n * cnt * sizeof(T));
    hipLaunchKernelGGL((k_fill_n<T>), dim3((unsigned)((n + 255) / 256), (unsigned)cnt), dim3(256), 0, c->stream, a,
                       cnt, (unsigned)n, val);
  }
  LCHK(c, hipGetLastError());
}
template <typename T>
void op_fill(les_ctx* c, T* a, size_t n, T val) {
  if (!n) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  {
    ProfScope _ps(c, "fill", (double)n * sizeof(T));
    hipLaunchKernelGGL((k_fill<T>), dim3((unsigned)blocks), dim3(256), 0, c->stream, a, n, val);
  }
  LCHK(c, hipGetLastError());
}
template <typename T>
__global__ void k_copy(T* d, const T* s, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i];
}
template <typename T>
void op_copy(les_ctx* c, T* dst, const T* src, size_t n) {
  if (!n) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  {
    ProfScope _ps(c, "copy", 2.0 * n * sizeof(T));
    hipLaunchKernelGGL((k_copy<T>), dim3((unsigned)blocks), dim3(256), 0, c->stream, dst, src, n);
  }
  LCHK(c, hipGetLastError());
}
// complex-(BOGUS,0) pattern written by press_stag_array.f90:288-290
template <typename T>
__global__ void k_fill_bogus_cplx(T* a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = (i & 1) ? T(0) : T(BOGUS);
}

static dim3 grid1(size_t n, int bs = 256) {
  size_t b = (n + bs - 1) / bs;
  if (b > 1048576) b = 1048576;
  return dim3((unsigned)b);
}

// ------------------------------------------------------------------------
// convec.f90 loader / AB2 epilogue (shared by the neutral and the convective
// steps)
// ------------------------------------------------------------------------
template <typename T>
struct LoadConvec {  // fields: u1,u2,u3, vort1, vort2, vort3 ; levels 0..nz-1
  const T *u1, *u2, *u3, *du1d2, *du1d3, *du2d1, *du2d3, *du3d1, *du3d2;
  int ny, nz, ld;
  T cst;
  static constexpr double reads() { return 1.5; }  // 9 fields feed 6 rows
  __device__ T operator()(int plane, int y, int x) const {
    const int f = plane / nz, k = plane - f * nz;
    const long i = ((long)k * ny + y) * ld + x, pl = (long)ld * ny;
    switch (f) {
      case 0: return cst * u1[i];
      case 1: return cst * u2[i];
      case 2: return cst * u3[i];
      case 3:
        if (k == 0) return cst * (T(0.5) * (du3d2[i] + du3d2[i + pl]) - du2d3[i]);
        return cst * (du3d2[i] - du2d3[i]);
      case 4:
        if (k == 0) return cst * (du1d3[i] - T(0.5) * (du3d1[i] + du3d1[i + pl]));
        return cst * (du1d3[i] - du3d1[i]);
      default: return cst * (du2d1[i] - du1d2[i]);
    }
  }
};

// main.f90:319-348 after the buoyancy term: Coriolis on whole rows, then
// the sponge on the nx*ny points of level r, each term rounded as written.
// f: velocity component, u0: its value before the update, cor: the Coriolis
// product of the other component.
template <typename T>
__device__ __forceinline__ T forcing_terms(const ForcingArgs<T>& fo, T rhs, int f, int r, int x, int nx, int nl,
                                           T u0, T cor) {
#pragma clang fp contract(off)
  if (fo.cor[0]) {
    if (f == 0) rhs = rhs + cor - fo.cvg;
    else if (f == 1) rhs = rhs - cor + fo.cug;
  }
  if (x < nx) {
    if (fo.damp == 2) {
      const T m = (T)(fo.sums[f * nl + r] / fo.nxy);
      rhs = rhs - (T)fo.coef[(f == 2 ? nl : 0) + r] * (u0 - m);
    } else if (fo.damp == 1 && f == 2) {
      rhs = rhs - (T)fo.coef[nl + r] * u0;
    }
  }
  return rhs;
}

// main.f90:309-378 fused into the c2r epilogue of convec: RHS = -conv - divt,
// u* = u + dt*(1.5 RHS - 0.5 RHS_f + force) (force on u only).  Arithmetic
// and evaluation order as k_ab2.
template <typename T>
struct StoreConvecAB2 {
  T* R[3];
  T* Rf[3];
  const T* dv[3];
  T* U[3];
  int ppf, ny, ld, nx;
  T dt, force;
  int euler;
  const T* beta;  // buoyancy added to RHSz (main.f90:312-317), or null
  ForcingArgs<T> fo;  // Coriolis and sponge terms (none by default)
  static constexpr double writes() { return 5.0; }
  struct Pre {
    T dv, beta, rf, u, cor;
  };
  __device__ Pre pre(int plane, int y, int x) const {
    const int f = plane / ppf, r = plane - f * ppf;
    const long i = ((long)r * ny + y) * ld + x;
    Pre p;
    p.dv = dv[f][i];
    p.beta = (f == 2 && beta) ? beta[i] : T(0);
    p.rf = euler ? T(0) : Rf[f][i];
    p.u = U[f][i];
    p.cor = (fo.cor[0] && f < 2) ? fo.cor[f][i] : T(0);
    return p;
  }
  __device__ void fin(int plane, int y, int x, T v, const Pre& p) const {
    const int f = plane / ppf, r = plane - f * ppf;
    const long i = ((long)r * ny + y) * ld + x;
    const T conv = x < nx ? v : T(0);
    T rhs = -conv - p.dv;
    if (f == 2 && beta) rhs = rhs + p.beta;
    if (fo.cor[0] || fo.damp) rhs = forcing_terms(fo, rhs, f, r, x, nx, ppf, p.u, p.cor);
    R[f][i] = rhs;
    T fo;
    if (euler) {
      fo = rhs;
      Rf[f][i] = rhs;
    } else {
      fo = p.rf;
    }
    if (f == 0)
      U[f][i] = p.u + dt * (T(TADV1) * rhs + T(TADV2) * fo + force);
    else
      U[f][i] = p.u + dt * (T(TADV1) * rhs + T(TADV2) * fo);
  }
  __device__ void operator()(int plane, int y, int x, T v) const { fin(plane, y, x, v, pre(plane, y, x)); }
};


// ------------------------------------------------------------------------
// 3/2-dealiased pointwise products on the padded grid, level-batched.
//
// XB holds x-spectral rows (kx < nx/2) of NF real fields on the padded y grid,
// field-major: row ((f * nz + k) * ny2 + y).  One workgroup owns one padded
// row yb and CH consecutive output levels.  It loads the CH + HB + HA levels
// it needs (HB below, HA above), pairs the fields two per complex transform,
// runs ONE batched inverse x-FFT of length nx2 over all of them in LDS,
// forms the policy's products for the CH levels, runs ONE batched forward
// FFT and stores the kx < nx/2 coefficients of the NOUT outputs to XP (row
// ((o * (nz-1) + k) * ny2 + y)).  Batching the levels keeps the number of
// workgroup barriers per level small; the padded physical fields stay in LDS.
// ------------------------------------------------------------------------
template <typename T, int N, class Pol, int CH, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    k_pad_products(const cx<T>* __restrict__ XB, cx<T>* __restrict__ XP, int ny2, int nz, int nxh, FftPlan p,
                   const cx<T>* __restrict__ twg, Pol pol) {
  constexpr int NF = Pol::NF, NP = (NF + 1) / 2, NO = Pol::NOUT, NOP = (NO + 1) / 2;
  constexpr int HB = Pol::HB, NL = CH + Pol::HB + Pol::HA;
  constexpr int NB = N > 0 ? N : 4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = N > 0 ? N : p.n;
  constexpr int RS = N > 0 ? lds_row(N) : 0;  // padded transform stride (compile-time lengths)
  const int rs = N > 0 ? RS : n;
  auto ix = [&](int e) { return N > 0 ? lds_pad(e) : e; };
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* win = tw + n;              // [NL][NP][rs]: field pairs, physical after the inverse FFT
  cx<T>* out = win + NL * NP * rs;  // [CH][NOP][rs]: product pairs
  load_twiddles_for(tw, twg, n, N > 0, 1);
  const int yb = blockIdx.x;
  const int nl = nz - 1;
  const int k0 = blockIdx.y * CH;
  for (int e = threadIdx.x; e < NL * NP * n; e += BLOCK) {
    const int lv = e / (NP * n), r = e - lv * NP * n, t = r / n, j = r - t * n;
    const int kk = k0 - HB + lv;
    cx<T> z = {T(0), T(0)};
    if (kk >= 0 && kk < nz) {
      const bool hb = 2 * t + 1 < NF;
      const cx<T>* a = XB + (((long)(2 * t) * nz + kk) * ny2 + yb) * nxh;
      const cx<T>* b = XB + (((long)(2 * t + 1) * nz + kk) * ny2 + yb) * nxh;
      const cx<T> zero = {T(0), T(0)};
      if (j == 0) {
        z = {a[0].x, hb ? b[0].x : T(0)};
      } else if (j < nxh) {
        const cx<T> A = a[j], B = hb ? b[j] : zero;
        z = {A.x - B.y, A.y + B.x};
      } else if (n - j < nxh) {
        const cx<T> A = a[n - j], B = hb ? b[n - j] : zero;
        z = {A.x + B.y, B.x - A.y};
      }
    }
    win[(lv * NP + t) * rs + ix(j)] = z;
  }
  __syncthreads();
  if constexpr (N > 0)
    fft_lds_c<T, NB, 1, RS, NL * NP, BLOCK, false, true, 1>(win, tw, +1);
  else
    fft_lds<T, false, 8>(win, p, 1, n, NL * NP, tw, +1);
  for (int e = threadIdx.x; e < CH * n; e += BLOCK) {
    const int lv = e / n, x = e - lv * n, k = k0 + lv;
    T o[2 * NOP];
#pragma unroll
    for (int q = 0; q < 2 * NOP; ++q) o[q] = T(0);
    if (k < nl) {
      auto F = [&](int dk, int f) -> T {
        const cx<T> z = win[((lv + HB + dk) * NP + (f >> 1)) * rs + ix(x)];
        return (f & 1) ? z.y : z.x;
      };
      pol.product(k, F, o);
    }
#pragma unroll
    for (int q = 0; q < NOP; ++q) out[(lv * NOP + q) * rs + ix(x)] = {o[2 * q], o[2 * q + 1]};
  }
  __syncthreads();
  if constexpr (N > 0)
    fft_lds_c<T, NB, 1, RS, CH * NOP, BLOCK, false, true, 1>(out, tw, -1);
  else
    fft_lds<T, false, 8>(out, p, 1, n, CH * NOP, tw, -1);
  for (int e = threadIdx.x; e < CH * NO * nxh; e += BLOCK) {
    const int lv = e / (NO * nxh), r = e - lv * NO * nxh, o = r / nxh, kk = r - o * nxh;
    const int k = k0 + lv;
    if (k >= nl) continue;
    const cx<T>* z = out + (lv * NOP + (o >> 1)) * rs;
    const cx<T> zk = z[ix(kk)], zn = z[ix(kk ? n - kk : 0)];
    const cx<T> v = (o & 1) ? cx<T>{T(0.5) * (zk.y + zn.y), T(-0.5) * (zk.x - zn.x)}
                            : cx<T>{T(0.5) * (zk.x + zn.x), T(0.5) * (zk.y - zn.y)};
    XP[(((long)Pol::slot(o) * nl + k) * ny2 + yb) * nxh + kk] = v;
  }
}

// LDS bytes of k_pad_products (padded rows for compile-time lengths)
template <typename T, class Pol, int CH>
constexpr size_t pad_products_lds(int n, bool padded) {
  return sizeof(cx<T>) * ((size_t)n + (size_t)(padded ? lds_row(n) : n) *
                                          ((size_t)(CH + Pol::HB + Pol::HA) * ((Pol::NF + 1) / 2) +
                                           (size_t)CH * ((Pol::NOUT + 1) / 2)));
}

template <typename T, int N, class Pol, int CH>
static void launch_pad_products_n(les_ctx* c, const cx<T>* XB, cx<T>* XP, const Pol& pol, const char* tag) {
  constexpr int BLOCK = (N >= 192 || N == 0) ? 1024 : 256;
  static_assert(N == 0 || pad_products_lds<T, Pol, CH>(N, true) <= 163840, "padded-product tile exceeds LDS");
  const Grid& g = c->g;
  const int n = g.nx2, nz = g.nz;
  const size_t sm = pad_products_lds<T, Pol, CH>(n, N > 0);
  if (sm > 163840) {
    c->err = "padded-product tile exceeds LDS for this grid";
    return;
  }
  if (sm > 65536)
    LCHK(c, hipFuncSetAttribute((const void*)k_pad_products<T, N, Pol, CH, BLOCK>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  const unsigned gy = (unsigned)((nz - 1 + CH - 1) / CH);
  ProfScope ps(c, tag,
               ((double)Pol::NF * nz + (double)Pol::NOUT * (nz - 1)) * g.ny2 * g.nxh * sizeof(cx<T>));
  hipLaunchKernelGGL((k_pad_products<T, N, Pol, CH, BLOCK>), dim3(g.ny2, gy), dim3(BLOCK), sm, c->stream, XB, XP,
                     g.ny2, nz, g.nxh, plan_for(c, n), tw_for<T>(c, n), pol);
  LCHK(c, hipGetLastError());
}

// ------------------------------------------------------------------------
// Twiddles of a compile-time transform kept in the padding words of 16 padded
// LDS rows: word 17 m + 16 of a padded row (m < N/16) is never an element,
// since lds_pad skips it.  Stage s >= 1 of the stage-major table takes hole
// numbers h = H(s) + (r - 1) NS + k, H(s) rounded up to a multiple of 16, and
// hole h is word 17 (h / 16) + 16 of row h % 16: for a given stage the row and
// word split into a part that depends on k and a compile-time part of r.  The
// first stage's twiddles are all 1 and are not stored.
// ------------------------------------------------------------------------
template <int N, int PLAN>
constexpr int tw_hole_off(int s) {
  int h = 0, ns = fft_radix(N, 0, PLAN);
  for (int i = 1; i < s; ++i) {
    const int r = fft_radix(N, i, PLAN);
    h = (h + (r - 1) * ns + 15) / 16 * 16;
    ns *= r;
  }
  return h;
}
template <int N, int PLAN>
constexpr int tw_stage_at(int off) {
  for (int s = 0; s < fft_nstages(N, PLAN); ++s)
    if (fft_tw_off(N, s, PLAN) == off) return s;
  return -1;
}

template <typename T, int N, int PLAN>
struct TwHoles {
  static constexpr int RS = lds_row(N);
  static_assert(N % 16 == 0 && tw_hole_off<N, PLAN>(fft_nstages(N, PLAN)) <= N, "twiddles exceed the padding words");
  cx<T>* rows;  // 16 padded rows
  __device__ static int word(int h) { return (h & 15) * RS + 17 * (h >> 4) + 16; }
  template <int OFF, int NS>
  __device__ cx<T> at(int r, int k) const {
    constexpr int S = tw_stage_at<N, PLAN>(OFF), H = tw_hole_off<N, PLAN>(S);
    static_assert(S >= 1 && H % 16 == 0 && (NS % 16 == 0 || 16 % NS == 0), "twiddle hole layout");
    if constexpr (NS >= 16) {
      return rows[(k & 15) * RS + 17 * (k >> 4) + (17 * (H / 16 + (r - 1) * (NS / 16)) + 16)];
    } else {
      constexpr int GR = 16 / NS;  // stage rows r that share one word column
      return rows[((r - 1) % GR * NS + k) * RS + (17 * (H / 16 + (r - 1) / GR) + 16)];
    }
  }
  // copy the stage-major table of N for PLAN (from load_twiddles_for's source)
  __device__ void fill(const cx<T>* twg) const {
    const cx<T>* tab = twg + N + PLAN * N;
    int ns = fft_radix(N, 0, PLAN), h0 = 0;
    for (int s = 1; s < fft_nstages(N, PLAN); ++s) {
      const int r = fft_radix(N, s, PLAN), cnt = (r - 1) * ns, off = fft_tw_off(N, s, PLAN);
      for (int e = threadIdx.x; e < cnt; e += blockDim.x) rows[word(h0 + e)] = tab[off + e];
      h0 = (h0 + cnt + 15) / 16 * 16;
      ns *= r;
    }
  }
};

// ------------------------------------------------------------------------
// Streaming variant of the padded products.  One workgroup owns padded row yb
// and a run of output levels [k0, k1).  LDS rows: [X (NP)][O (NOP)][Y (NP)]
// [M (NPM)]; level L's physical field pairs live in X when L - k0 is even, in
// Y when odd, and M keeps the pairs the products read below level k.  A level
// step (1) forms level k's products from X/Y/M into O, moving level k's lower
// pairs to M; (2) stages level k + 2, whose spectral rows were loaded into
// registers during the previous step, into level k's slot and starts the HBM
// reads of level k + 3; (3) transforms level k + 2 and level k's products in
// one pass over 9 contiguous rows (X O or O Y).  The products go through the
// backward-signed pass conjugated, which gives the conjugate of their forward
// transform exactly (every butterfly and twiddle product is symmetric under
// conjugation), so one pass of one sign serves both.  The product of a column
// x reads only column x, so the thread that owns x may overwrite it.  With 16
// rows the twiddles sit in the rows' padding words (TwHoles), which keeps
// three 512-thread workgroups on a CU at nx2 = 384 in FP32.  Each level is
// inverse-transformed once per run (the batched kernel redoes the halo levels
// of every CH-level group), with the radix-8 plan.  The physical values equal
// the batched kernel's up to the rounding of the transform plan.
// ------------------------------------------------------------------------
template <typename T, int N, class Pol>
struct PadStream {
  static constexpr int NF = Pol::NF, NP = (NF + 1) / 2, NO = Pol::NOUT, NOP = (NO + 1) / 2;
  static constexpr int HB = Pol::HB, HA = Pol::HA, NPM = HB > 0 ? Pol::NPM : 0;
  static_assert(HA == 1 && HB <= 1 && NPM <= NP && NOP <= NP, "streaming products read levels k-1..k+1");
  static constexpr int RS = lds_row(N), NXH = N / 3, H2 = NXH / 2;  // nx2 = 3 nxh
  static constexpr int NROWS = 2 * NP + NOP + NPM;
  static constexpr int PLAN = 2;
  static constexpr bool HOLES = NROWS >= 16 && N % 16 == 0;  // twiddles in the padding words
  static constexpr int BLOCK = N >= 384 ? 512 : 256;
  static constexpr int ITEMS = NP * H2;  // one item: two coefficients of a field pair
  static constexpr int MAXI = (ITEMS + BLOCK - 1) / BLOCK;
  static constexpr size_t LDS = sizeof(cx<T>) * ((size_t)RS * NROWS + (HOLES ? 0 : (size_t)N));
  // waves per SIMD asked of the register allocator: at nx2 = 384 in FP32 the
  // 16-row ring fits three workgroups (24 waves) per CU, which holds at 80
  // registers; elsewhere no bound
  static constexpr int WPE = (HOLES && N == 384 && sizeof(T) == 4) ? 6 : 1;
};

template <typename T, int N, class Pol>
__global__ void __launch_bounds__((PadStream<T, N, Pol>::BLOCK))
    __attribute__((amdgpu_waves_per_eu(PadStream<T, N, Pol>::WPE))) k_pad_stream(const cx<T>* __restrict__ XB, cx<T>* __restrict__ XP, int ny2, int nz, int lc,
                 const cx<T>* __restrict__ twg, Pol pol) {
  using G = PadStream<T, N, Pol>;
  constexpr int NF = G::NF, NP = G::NP, NO = G::NO, NOP = G::NOP, NPM = G::NPM, NROWS = G::NROWS;
  constexpr int RS = G::RS, NXH = G::NXH, H2 = G::H2, BLOCK = G::BLOCK, MAXI = G::MAXI, PLAN = G::PLAN;
  constexpr int ITEMS = G::ITEMS;
  struct alignas(2 * sizeof(cx<T>)) cx2 {
    cx<T> a, b;
  };
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* X = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* O = X + NP * RS;
  cx<T>* Y = O + NOP * RS;
  cx<T>* M = Y + NP * RS;
  const auto tw = [&] {
    if constexpr (G::HOLES) {
      const TwHoles<T, N, PLAN> h{X};
      h.fill(twg);
      return h;
    } else {
      cx<T>* t = X + NROWS * RS;
      load_twiddles_for(t, twg, N, true, PLAN);
      return static_cast<const cx<T>*>(t);
    }
  }();
  const int yb = blockIdx.x, nl = nz - 1;
  const int k0 = blockIdx.y * lc, k1 = min(k0 + lc, nl);
  auto slot = [&](int L) { return ((L - k0) & 1) ? Y : X; };
  cx2 ra[MAXI], rb[MAXI];
  auto fetch = [&](int L) {  // spectral rows of level L -> registers (zero outside 0..nz-1)
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
      const int w = threadIdx.x + i * BLOCK;
      ra[i] = cx2{{T(0), T(0)}, {T(0), T(0)}};
      rb[i] = ra[i];
      if ((ITEMS % BLOCK == 0 || w < ITEMS) && L >= 0 && L < nz) {
        const int t = w / H2, jj = w - t * H2;
        ra[i] = reinterpret_cast<const cx2*>(XB + (((long)(2 * t) * nz + L) * ny2 + yb) * NXH)[jj];
        if (2 * t + 1 < NF)
          rb[i] = reinterpret_cast<const cx2*>(XB + (((long)(2 * t + 1) * nz + L) * ny2 + yb) * NXH)[jj];
      }
    }
  };
  // field pair (a, b) as one complex row: z_j = a_j + i b_j and its conjugate
  // partner z_{N-j} = conj(a_j) + i conj(b_j); kx = 0 keeps the real parts
  auto put = [](cx<T>* row, int j, cx<T> A, cx<T> B) {
    if (j == 0) {
      row[lds_pad(0)] = {A.x, B.x};
    } else {
      row[lds_pad(j)] = {A.x - B.y, A.y + B.x};
      row[lds_pad(N - j)] = {A.x + B.y, B.x - A.y};
    }
  };
  auto stage = [&](int L) {  // registers -> slot of level L, zero band [NXH, N - NXH]
    cx<T>* s = slot(L);
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
      const int w = threadIdx.x + i * BLOCK;
      if (ITEMS % BLOCK == 0 || w < ITEMS) {
        const int t = w / H2, jj = w - t * H2;
        put(s + t * RS, 2 * jj, ra[i].a, rb[i].a);
        put(s + t * RS, 2 * jj + 1, ra[i].b, rb[i].b);
      }
    }
    constexpr int NZB = N - 2 * NXH + 1;
    for (int e = threadIdx.x; e < NP * NZB; e += BLOCK) {
      const int t = e / NZB, j = NXH + (e - t * NZB);
      s[t * RS + lds_pad(j)] = {T(0), T(0)};
    }
  };
  if constexpr (NPM > 0) {  // level k0 - 1 (slot Y): only its lower pairs are read
    fetch(k0 - 1);
    stage(k0 - 1);
    __syncthreads();
    fft_lds_c<T, N, 1, RS, NPM, BLOCK, false, true, PLAN>(Y, tw, +1);
    for (int e = threadIdx.x; e < NPM * N; e += BLOCK) {
      const int p = e / N, x = e - p * N;
      M[p * RS + lds_pad(x)] = Y[p * RS + lds_pad(x)];
    }
  }
  fetch(k0);
  stage(k0);
  fetch(k0 + 1);
  __syncthreads();  // level k0 - 1 has left Y
  stage(k0 + 1);
  if (k0 + 2 <= k1) fetch(k0 + 2);
  __syncthreads();
  fft_lds_c<T, N, 1, RS, NP, BLOCK, false, true, PLAN>(X, tw, +1);
  fft_lds_c<T, N, 1, RS, NP, BLOCK, false, true, PLAN>(Y, tw, +1);
  for (int k = k0; k < k1; ++k) {
    cx<T>* cur = slot(k);
    cx<T>* nxt = slot(k + 1);
    if (k > k0) __syncthreads();  // the previous level's products in O have been stored
    for (int x = threadIdx.x; x < N; x += BLOCK) {
      T o[2 * NOP];
#pragma unroll
      for (int q = 0; q < 2 * NOP; ++q) o[q] = T(0);
      auto F = [&](int dk, int f) -> T {
        const cx<T>* b = dk < 0 ? M : (dk > 0 ? nxt : cur);
        const cx<T> z = b[(f >> 1) * RS + lds_pad(x)];
        return (f & 1) ? z.y : z.x;
      };
      pol.product(k, F, o);
#pragma unroll
      for (int p = 0; p < NPM; ++p) M[p * RS + lds_pad(x)] = cur[p * RS + lds_pad(x)];
#pragma unroll
      for (int q = 0; q < NOP; ++q) O[q * RS + lds_pad(x)] = {o[2 * q], -o[2 * q + 1]};
    }
    __syncthreads();
    const bool more = k + 2 <= k1;  // level k + 2 is read by the products of level k + 1
    if (more) {
      stage(k + 2);  // into level k's slot
      if (k + 3 <= k1) fetch(k + 3);  // in flight during the next level's work
    }
    __syncthreads();
    if (more)
      fft_lds_c<T, N, 1, RS, NP + NOP, BLOCK, false, true, PLAN>(cur == X ? X : O, tw, +1);
    else
      fft_lds_c<T, N, 1, RS, NOP, BLOCK, false, true, PLAN>(O, tw, +1);
    for (int e = threadIdx.x; e < NO * NXH; e += BLOCK) {
      const int o = e / NXH, kx = e - o * NXH;
      const cx<T>* z = O + (o >> 1) * RS;
      const cx<T> a = z[lds_pad(kx)], b = z[lds_pad(kx ? N - kx : 0)];
      const cx<T> zk = {a.x, -a.y}, zn = {b.x, -b.y};  // forward transform = conj of the pass
      const cx<T> v = (o & 1) ? cx<T>{T(0.5) * (zk.y + zn.y), T(-0.5) * (zk.x - zn.x)}
                              : cx<T>{T(0.5) * (zk.x + zn.x), T(0.5) * (zk.y - zn.y)};
      XP[(((long)Pol::slot(o) * nl + k) * ny2 + yb) * NXH + kx] = v;
    }
  }
}

// level runs per padded row of the streaming kernel (workgroups = ny2 * runs)
constexpr int PAD_STREAM_RUNS = 4;

template <typename T, int N, class Pol>
static void launch_pad_stream_n(les_ctx* c, const cx<T>* XB, cx<T>* XP, const Pol& pol, const char* tag) {
  using G = PadStream<T, N, Pol>;
  static_assert(G::LDS <= 163840, "streaming padded-product ring exceeds LDS");
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1;
  if (g.nx2 != N || 3 * g.nxh != N || g.nxh % 2) {
    c->err = "streaming padded products: grid does not match the kernel";
    return;
  }
  if (G::LDS > 65536)
    LCHK(c, hipFuncSetAttribute((const void*)k_pad_stream<T, N, Pol>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)G::LDS));
  const int lc = (nl + PAD_STREAM_RUNS - 1) / PAD_STREAM_RUNS;
  const unsigned gy = (unsigned)((nl + lc - 1) / lc);
  ProfScope ps(c, tag,
               ((double)Pol::NF * nz + (double)Pol::NOUT * (nz - 1)) * g.ny2 * g.nxh * sizeof(cx<T>));
  hipLaunchKernelGGL((k_pad_stream<T, N, Pol>), dim3(g.ny2, gy), dim3(G::BLOCK), G::LDS, c->stream, XB, XP, g.ny2,
                     nz, lc, tw_for<T>(c, N), pol);
  LCHK(c, hipGetLastError());
}

// CH32 / CH64: levels per workgroup of the batched kernel in FP32 / FP64
// (sized to the 160 KB LDS); c->tune.pad_mode 1 selects the streaming kernel
// where it is instantiated (nx2 >= 96)
template <typename T>
struct PolConvecScalar;
template <typename T>
static bool launch_pad_wave(les_ctx* c, const cx<T>* XB, cx<T>* XP, const PolConvecScalar<T>& pol, const char* tag);

// pad_mode 2: the one-transform-per-wavefront kernel where it is
// instantiated (the convective step with both scalars at nx2 = 384), the
// streaming kernel elsewhere
template <typename T, class Pol, int CH32, int CH64>
static void launch_pad_products(les_ctx* c, const cx<T>* XB, cx<T>* XP, const Pol& pol, const char* tag) {
  constexpr int CH = sizeof(T) == 4 ? CH32 : CH64;
  const int mode = c->tune.pad_mode[sizeof(T) == 8 ? 1 : 0];
  if constexpr (std::is_same_v<Pol, PolConvecScalar<T>>)
    if (mode == 2 && launch_pad_wave<T>(c, XB, XP, pol, tag)) return;
  const bool stream = mode >= 1;
  switch (c->g.nx2) {
#define LES_CASE(NN) \
  case NN:           \
    return launch_pad_products_n<T, NN, Pol, CH>(c, XB, XP, pol, tag);
#define LES_CASE_S(NN)                                                           \
  case NN:                                                                       \
    if (stream) return launch_pad_stream_n<T, NN, Pol>(c, XB, XP, pol, tag);    \
    return launch_pad_products_n<T, NN, Pol, CH>(c, XB, XP, pol, tag);
    LES_CASE(24) LES_CASE(48) LES_CASE_S(96) LES_CASE_S(144) LES_CASE_S(192) LES_CASE_S(384)
#undef LES_CASE
#undef LES_CASE_S
    default:
      return launch_pad_products_n<T, 0, Pol, CH>(c, XB, XP, pol, tag);
  }
}

// convec.f90:117-233: -(u x omega) at uvp (cx, cy) and w nodes (cz) from
// u1,u2,u3, omega1..3 (fields 0..5); cst = 1/(nx2 ny2)
template <typename T>
struct PolConvec {
  static constexpr int NF = 6, NOUT = 3, HB = 1, HA = 1;
  static constexpr int NPM = 1;  // field pairs read at level k - 1: (u1, u2)
  __device__ static int slot(int o) { return o; }  // output o -> plane block of XP
  T cst;
  // cx, cy at level k (reads levels k and k + 1)
  template <class Fn>
  __device__ void product_xy(int k, Fn F, T* o) const {
    const T u1 = F(0, 0), u2 = F(0, 1), u3 = F(0, 2), v1 = F(0, 3), v2 = F(0, 4), v3 = F(0, 5);
    const T u3p = F(1, 2), v1p = F(1, 3), v2p = F(1, 4);
    if (k == 0) {
      o[0] = cst * (u2 * (-v3) + T(0.5) * u3p * (v2p));
      o[1] = cst * (u1 * (v3) + T(0.5) * u3p * (-v1p));
    } else {
      o[0] = cst * (u2 * (-v3) + T(0.5) * (u3p * (v2p) + u3 * (v2)));
      o[1] = cst * (u1 * (v3) + T(0.5) * (u3p * (-v1p) + u3 * (-v1)));
    }
  }
  // cz at a level k >= 1 from u1, u2, omega1, omega2 there and u1, u2 at k - 1
  __device__ T cz(T u1, T u2, T v1, T v2, T u1m, T u2m) const {
    return cst * T(0.5) * ((u1 + u1m) * (-v2) + (u2 + u2m) * (v1));
  }
  template <class Fn>
  __device__ void product(int k, Fn F, T* o) const {
    product_xy(k, F, o);
    o[2] = k == 0 ? T(0) : cz(F(0, 0), F(0, 1), F(0, 3), F(0, 4), F(-1, 0), F(-1, 1));
  }
};

// scalar_RHS_calc advection (scalars_module.f90:322-338) of NS scalars: fields
// u, v, w (F0 + 0..2) and ds/dx, ds/dy, ds/dz per scalar (F0 + 3 + 3s ..)
template <typename T>
__device__ __forceinline__ T scalar_adv(int k, T u, T v, T w, T wp, T sx, T sy, T sz, T szp, T cst) {
  if (k == 0) return cst * (u * sx + v * sy + (T(0.5) * wp) * szp);
  return cst * (u * sx + v * sy + (w * sz + wp * szp) / T(2));
}
template <typename T, int NS>
struct PolScalar {
  static constexpr int NF = 3 + 3 * NS, NOUT = NS, HB = 0, HA = 1;
  static constexpr int NPM = 0;
  __device__ static int slot(int o) { return o; }
  T cst;
  template <class Fn>
  __device__ void product(int k, Fn F, T* o) const {
#pragma unroll
    for (int s = 0; s < NS; ++s)
      o[s] = scalar_adv<T>(k, F(0, 0), F(0, 1), F(0, 2), F(1, 2), F(0, 3 + 3 * s), F(0, 4 + 3 * s), F(0, 5 + 3 * s),
                           F(1, 5 + 3 * s), cst);
  }
};
// convec and both scalars in one pass: fields u1,u2,u3, omega1..3, then
// dtheta/dx,dy,dz and dq/dx,dy,dz.  Two real outputs share one complex
// transform, so the small scalar advection terms are paired with each other
// (o = cx, cy | adv theta, adv q | cz) and not with a momentum term, whose
// rounding would swamp them in FP32; XP keeps the order cx, cy, cz, theta, q.
template <typename T>
struct PolConvecScalar {
  static constexpr int NF = 12, NOUT = 5, HB = 1, HA = 1;
  static constexpr int NPM = 1;  // field pairs read at level k - 1: (u1, u2)
  __device__ static int slot(int o) { return o < 2 ? o : (o == 4 ? 2 : o + 1); }
  T cst;
  // the outputs at level k that read levels k and k + 1 only: cx, cy, theta, q
  template <class Fn>
  __device__ void product_same(int k, Fn F, T* o) const {
    PolConvec<T>{cst}.product_xy(k, F, o);
    const T u = F(0, 0), v = F(0, 1), w = F(0, 2), wp = F(1, 2);
#pragma unroll
    for (int s = 0; s < 2; ++s)
      o[2 + s] = scalar_adv<T>(k, u, v, w, wp, F(0, 6 + 3 * s), F(0, 7 + 3 * s), F(0, 8 + 3 * s), F(1, 8 + 3 * s),
                               cst);
  }
  template <class Fn>
  __device__ void product(int k, Fn F, T* o) const {
    product_same(k, F, o);
    o[4] = k == 0 ? T(0) : PolConvec<T>{cst}.cz(F(0, 0), F(0, 1), F(0, 3), F(0, 4), F(-1, 0), F(-1, 1));
  }
};


// ------------------------------------------------------------------------
// Streaming padded products of the convective step with both scalars, one
// transform per wavefront (tune.pad_mode 2, nx2 = 384).  A workgroup of eight
// wavefronts owns padded row yb and a run of output levels [k0, k1).  LDS
// rows: X and Y (six field pairs each, levels k and k + 1 alternately),
// O (two rows: (cx, cy) and (theta, q) advection at level k) and C (two
// rows, each pairing cz of two consecutive levels).  cz(k + 1) reads u1, u2
// at levels k and k + 1, so it is formed while both are resident and no
// level k - 1 row is kept.  A level step is
//   (1) the products: cx, cy, theta, q at level k into O, cz(k + 1) into C;
//   (2) wavefronts 0-5 stage field pair w of level k + 2 into level k's slot
//       and transform it, issuing the HBM reads of level k + 3; wavefront 6
//       transforms O's first row and stores it, and runs the first pass of
//       a C row at the even step that completes it and the second pass at
//       the next step; wavefront 7 transforms O's second row and stores it,
//       and finishes that C row two steps after its completion (every
//       wavefront then has at most four passes per level step),
// with two workgroup barriers per level: every transform belongs to one
// wavefront, whose passes need only wave_sync.  The last pass of each
// product row stays in registers and stores the kx < nxh coefficients
// directly (conjugate partners exchanged across lanes).  The twiddles sit in
// the rows' padding words.  Products enter conjugated, as in k_pad_stream;
// the cz pairs round differently from k_pad_stream's (cz, 0) rows.
// ------------------------------------------------------------------------
template <typename T, int N>
struct PadWave {
  static constexpr int NP = 6, NXH = N / 3, RS = lds_row(N), PLAN = 2, BLOCK = 512, NROWS = 16;
  static constexpr int NST = fft_nstages(N, PLAN);
  static constexpr bool OK = N % 16 == 0 && NXH == 128 && N / fft_radix(N, NST - 1, PLAN) == 64;
  static constexpr size_t LDS = sizeof(cx<T>) * (size_t)RS * NROWS;
  // at 52 KB (FP32) three workgroups share a CU: 24 wavefronts, 80 registers
  static constexpr int WPE = sizeof(T) == 4 ? 6 : 1;
  static constexpr int NCARRY = sizeof(T) == 4 ? 2 : 0;  // product rows kept in registers across levels
};

template <typename T, int N>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(PadWave<T, N>::WPE)))
    k_pad_wave(const cx<T>* __restrict__ XB, cx<T>* __restrict__ XP, int ny2, int nz, int lc,
               const cx<T>* __restrict__ twg, PolConvecScalar<T> pol) {
  using G = PadWave<T, N>;
  constexpr int NP = G::NP, RS = G::RS, NXH = G::NXH, PLAN = G::PLAN, NST = G::NST;
  constexpr int RL = fft_radix(N, NST - 1, PLAN);
  static_assert(G::OK && RL == 6 && N == 384 && NST == 3 && fft_radix(N, 0, PLAN) == 8 && fft_radix(N, 1, PLAN) == 8,
                "384 = 8 x 8 x 6: one field pair (128 coefficients) per wavefront, 64 last-stage butterflies");
  struct alignas(2 * sizeof(cx<T>)) cx2 {
    cx<T> a, b;
  };
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* X = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* Y = X + NP * RS;
  cx<T>* O = Y + NP * RS;
  cx<T>* C = O + 2 * RS;
  const TwHoles<T, N, PLAN> tw{X};
  tw.fill(twg);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int yb = blockIdx.x, nl = nz - 1;
  const int k0 = blockIdx.y * lc, k1 = min(k0 + lc, nl), n = k1 - k0;
  auto slot = [&](int L) { return ((L - k0) & 1) ? Y : X; };
  constexpr int OFF1 = fft_tw_off(N, 1, PLAN), OFF2 = fft_tw_off(N, 2, PLAN);
  cx<T> rtw[RL - 1];  // the last pass's twiddles of this lane (k = lane), filled once the table is in LDS
  cx2 ra, rb;  // coefficients 2 lane, 2 lane + 1 of fields 2w, 2w + 1
  auto fetch = [&](int L) {
    ra = cx2{{T(0), T(0)}, {T(0), T(0)}};
    rb = ra;
    if (L >= 0 && L < nz) {
      ra = reinterpret_cast<const cx2*>(XB + (((long)(2 * w) * nz + L) * ny2 + yb) * NXH)[lane];
      rb = reinterpret_cast<const cx2*>(XB + (((long)(2 * w + 1) * nz + L) * ny2 + yb) * NXH)[lane];
    }
  };
  // field pair (a, b) as one complex row (z_j = a_j + i b_j, z_{N-j} its
  // conjugate partner), zero band [NXH, N - NXH]
  auto put = [](cx<T>* row, int j, cx<T> A, cx<T> B) {
    if (j == 0) {
      row[lds_pad(0)] = {A.x, B.x};
    } else {
      row[lds_pad(j)] = {A.x - B.y, A.y + B.x};
      row[lds_pad(N - j)] = {A.x + B.y, B.x - A.y};
    }
  };
  // the zero band [NXH, N - NXH] is not written: the first pass of a
  // staged row takes it as zero without reading it
  auto stage_row = [&](cx<T>* row) {
    put(row, 2 * lane, ra.a, rb.a);
    put(row, 2 * lane + 1, ra.b, rb.b);
  };
  // the first two passes (radix 8, 8): the second with the bank-disjoint
  // butterfly order; all three passes of a staged row; the last pass with
  // register twiddles
  auto pass1 = [&](cx<T>* row) { wave_stage_lane<T, N, 8, 1, 0, false, false>(row, tw, lane, +1, rtw); };
  auto pass2 = [&](cx<T>* row) { wave_stage_lane<T, N, 8, 8, OFF1, true, false>(row, tw, lane, +1, rtw); };
  auto passes12 = [&](cx<T>* row) {
    pass1(row);
    pass2(row);
  };
  auto transform = [&](cx<T>* row) {
    wave_stage_lane<T, N, 8, 1, 0, false, false, NXH, N - NXH>(row, tw, lane, +1, rtw);
    wave_stage_lane<T, N, 8, 8, OFF1, true, false>(row, tw, lane, +1, rtw);
    wave_stage_lane<T, N, RL, 64, OFF2, false, true>(row, tw, lane, +1, rtw);
  };
  // last pass in registers, then the kx < NXH coefficients of the row's two
  // real outputs to XP planes sa, sb at levels la, lb (none where < 0)
  auto finish_store = [&](const cx<T>* row, int sa, int la, int sb, int lb) {
    cx<T> v[RL];
    wave_last_stage<T, N, PLAN>(row, tw, lane, +1, v, rtw);
    const int src = (64 - lane) & 63;  // holds the partners N - kx
    const cx<T> p4 = {__shfl(v[4].x, src), __shfl(v[4].y, src)};
    const cx<T> p5 = {__shfl(v[5].x, src), __shfl(v[5].y, src)};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const cx<T> a = v[r];
      const cx<T> b = r == 0 ? (lane ? p5 : v[0]) : (lane ? p4 : p5);
      const int kx = lane + 64 * r;
      const cx<T> zk = {a.x, -a.y}, zn = {b.x, -b.y};  // forward transform = conj of the pass
      if (la >= 0)
        XP[(((long)sa * nl + la) * ny2 + yb) * NXH + kx] = cx<T>{T(0.5) * (zk.x + zn.x), T(0.5) * (zk.y - zn.y)};
      if (lb >= 0)
        XP[(((long)sb * nl + lb) * ny2 + yb) * NXH + kx] = cx<T>{T(0.5) * (zk.y + zn.y), T(-0.5) * (zk.x - zn.x)};
    }
  };
  // cz at level L >= 1 from the pairs of level L (a) and pair 0 of level L - 1 (m)
  auto cz_at = [&](const cx<T>* a, const cx<T>* m, int x) -> T {
    const cx<T> p0 = a[lds_pad(x)], p1 = a[RS + lds_pad(x)], p2 = a[2 * RS + lds_pad(x)], q0 = m[lds_pad(x)];
    return PolConvec<T>{pol.cst}.cz(p0.x, p0.y, p1.y, p2.x, q0.x, q0.y);
  };
  constexpr int S_CX = 0, S_CY = 1, S_CZ = 2, S_T = 3, S_Q = 4;  // XP planes (PolConvecScalar::slot)
  __syncthreads();  // twiddles in place
#pragma unroll
  for (int r = 1; r < RL; ++r) {
    rtw[r - 1] = tw_at<OFF2, 64>(tw, r, lane);
    rtw[r - 1].y = -rtw[r - 1].y;
  }
  // level k0 into X; wavefront 0 also puts pair 0 of level k0 - 1 into Y for cz(k0)
  if (w < NP) {
    fetch(k0);
    stage_row(X + w * RS);
    if (w == 0) {
      fetch(k0 - 1);
      stage_row(Y);
    }
    wave_sync();
    transform(X + w * RS);
    if (w == 0) transform(Y);
    fetch(k0 + 1);
  }
  __syncthreads();
  for (int x = threadIdx.x; x < N; x += G::BLOCK) C[lds_pad(x)] = {k0 == 0 ? T(0) : cz_at(X, Y, x), T(0)};
  __syncthreads();
  if (w < NP) {
    stage_row(Y + w * RS);
    if (k0 + 2 <= k1) fetch(k0 + 2);
    wave_sync();
    transform(Y + w * RS);
  }
  __syncthreads();
  // the first NC rows (u, v, w and the vorticity) of level k + 1 at this
  // thread's column: level k's products read them and level k + 1's take
  // them from registers (a thread keeps its column from level to level);
  // FP32 only, where the registers allow it without spills
  static_assert(N <= G::BLOCK, "at most one product column per thread");
  constexpr int NC = G::NCARRY;
  cx<T> carry[NC > 0 ? NC : 1];
  for (int k = k0; k < k1; ++k) {
    const int c = k - k0;
    cx<T>* cur = slot(k);
    cx<T>* nxt = slot(k + 1);
    // column x of wavefront w < 6: the two 16-column blocks a 32-lane read
    // group takes lie 16 blocks apart where possible (padded, their LDS
    // banks do not overlap); every column once
    for (int x0 = threadIdx.x; x0 < N; x0 += G::BLOCK) {
      const int blk = w < 4 ? 2 * w + (lane >> 5) + 16 * ((lane >> 4) & 1) : 8 + 4 * (w - 4) + (lane >> 4);
      const int x = 16 * blk + (lane & 15);
      T o[4];
      cx<T> cr[NP], nr[NP];  // the field pairs of levels k and k + 1 at x (nxt row 3 is not used)
#pragma unroll
      for (int r = 0; r < NP; ++r)
        if (r >= NC || c == 0) cr[r] = cur[r * RS + lds_pad(x)];
      if (c > 0) {
#pragma unroll
        for (int r = 0; r < NC; ++r) cr[r] = carry[r];
      }
#pragma unroll
      for (int r = 0; r < NP; ++r) nr[r] = r == 3 ? cx<T>{T(0), T(0)} : nxt[r * RS + lds_pad(x)];
      auto F = [&](int dk, int f) -> T {
        const cx<T> z = dk > 0 ? nr[f >> 1] : cr[f >> 1];
        return (f & 1) ? z.y : z.x;
      };
      pol.product_same(k, F, o);
      O[lds_pad(x)] = {o[0], -o[1]};
      O[RS + lds_pad(x)] = {o[2], -o[3]};
      // cz(k + 1): part (c + 1) & 1 of C row ((c + 1) >> 1) & 1, conjugated
      const T z = PolConvec<T>{pol.cst}.cz(nr[0].x, nr[0].y, nr[1].y, nr[2].x, cr[0].x, cr[0].y);
#pragma unroll
      for (int r = 0; r < NC; ++r) carry[r] = nr[r];
      T* cp = reinterpret_cast<T*>(C + (((c + 1) >> 1) & 1) * RS + lds_pad(x));
      if ((c + 1) & 1)
        cp[1] = -z;
      else
        cp[0] = z;
    }
    __syncthreads();
    // wavefronts 6 and 7 carry the longest work of this phase (four passes)
    if (w >= NP) __builtin_amdgcn_s_setprio(1);
    if (w < NP) {
      if (k + 2 <= k1) {  // level k + 2 is read by the products of level k + 1
        cx<T>* row = cur + w * RS;
        stage_row(row);
        if (k + 3 <= k1) fetch(k + 3);  // in flight during the next level's work
        wave_sync();
        transform(row);
      }
    } else if (w == NP) {
      // a C row completed at an even step c: its first pass at c, its second at c + 1
      passes12(O);
      finish_store(O, S_CX, k, S_CY, k);
      if ((c & 1) == 0)
        pass1(C + ((c >> 1) & 1) * RS);
      else
        pass2(C + ((c >> 1) & 1) * RS);
    } else {
      // ... and its last pass at c + 2, storing cz(k0 + c), cz(k0 + c + 1)
      passes12(O + RS);
      finish_store(O + RS, S_T, k, S_Q, k);
      if ((c & 1) == 0 && c >= 2) finish_store(C + (((c >> 1) + 1) & 1) * RS, S_CZ, k - 2, S_CZ, k - 1);
    }
    if (w >= NP) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
  }
  // the C rows whose passes the run's last steps did not reach: completed at
  // n - 1 (n odd; first pass done; its second level, cz(k1), another run
  // stores) or at n - 2 (n even; two passes done)
  if (w == NP + 1) {
    if (n & 1) {
      cx<T>* row = C + (((n - 1) >> 1) & 1) * RS;
      pass2(row);
      finish_store(row, S_CZ, k1 - 1, S_CZ, -1);
    } else if (n >= 2) {
      finish_store(C + (((n - 2) >> 1) & 1) * RS, S_CZ, k1 - 2, S_CZ, k1 - 1);
    }
  }
}

template <typename T, int N>
static void launch_pad_wave_n(les_ctx* c, const cx<T>* XB, cx<T>* XP, const PolConvecScalar<T>& pol,
                              const char* tag) {
  using G = PadWave<T, N>;
  static_assert(G::LDS <= 163840, "padded-product ring exceeds LDS");
  const Grid& g = c->g;
  const int nz = g.nz, nl = nz - 1;
  if (g.nx2 != N || 3 * g.nxh != N) {
    c->err = "padded products (one transform per wavefront): grid does not match the kernel";
    return;
  }
  if (G::LDS > 65536)
    LCHK(c, hipFuncSetAttribute((const void*)k_pad_wave<T, N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)G::LDS));
  const int lc = (nl + PAD_STREAM_RUNS - 1) / PAD_STREAM_RUNS;
  const unsigned gy = (unsigned)((nl + lc - 1) / lc);
  ProfScope ps(c, tag, (12.0 * nz + 5.0 * (nz - 1)) * g.ny2 * g.nxh * sizeof(cx<T>));
  hipLaunchKernelGGL((k_pad_wave<T, N>), dim3(g.ny2, gy), dim3(G::BLOCK), G::LDS, c->stream, XB, XP, g.ny2, nz, lc,
                     tw_for<T>(c, N), pol);
  LCHK(c, hipGetLastError());
}
template <typename T>
static bool launch_pad_wave(les_ctx* c, const cx<T>* XB, cx<T>* XP, const PolConvecScalar<T>& pol, const char* tag) {
  if (c->g.nx2 != 384) return false;
  launch_pad_wave_n<T, 384>(c, XB, XP, pol, tag);
  return true;
}

}  // namespace les
