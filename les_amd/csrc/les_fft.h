// les_fft.h -- LDS-resident mixed-radix (4/2/3) FFT building blocks for gfx950.
//
// All horizontal transforms of the LES hot path (2^a 3^b lengths: 16..384)
// run through one in-place Stockham routine that works on a batch of
// transforms staged in LDS.  Element e of transform t lives at
// buf[t*ts + e*es]; the work-item -> (transform, butterfly) mapping is a
// template flag so that both layouts used here are bank-conflict free:
//   * row transforms   (es = 1,  ts = n): lanes walk butterflies of one row,
//   * column tiles     (es = TX, ts = 1): lanes walk adjacent columns, so a
//     64-wide wavefront reads 64 consecutive complex words per access.
// Twiddles come from a per-length table exp(-2*pi*i*k/n) copied to LDS.
#pragma once
#include <hip/hip_runtime.h>

// The transform arithmetic is evaluated as written, without contracting a
// multiply and an add into an FMA: the compiler's choice would otherwise
// depend on the surrounding kernel (tile width, unrolling), and a transform
// must give the same bits whichever launch shape runs it.
#pragma clang fp contract(off)

namespace les {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

struct FftPlan {
  int n;        // transform length
  int ns;       // number of stages
  int rad[12];  // radices, product == n
};

// R-point DFT in registers, sign -1 (forward) or +1 (backward).
template <typename T, int R>
__device__ __forceinline__ void butterfly(cx<T> (&v)[R], int sign) {
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = {a.x + b.x, a.y + b.y};
    v[1] = {a.x - b.x, a.y - b.y};
  } else if constexpr (R == 4) {
    cx<T> a0 = {v[0].x + v[2].x, v[0].y + v[2].y};
    cx<T> a1 = {v[0].x - v[2].x, v[0].y - v[2].y};
    cx<T> a2 = {v[1].x + v[3].x, v[1].y + v[3].y};
    cx<T> d = {v[1].x - v[3].x, v[1].y - v[3].y};
    cx<T> a3 = sign < 0 ? cx<T>{d.y, -d.x} : cx<T>{-d.y, d.x};
    v[0] = {a0.x + a2.x, a0.y + a2.y};
    v[1] = {a1.x + a3.x, a1.y + a3.y};
    v[2] = {a0.x - a2.x, a0.y - a2.y};
    v[3] = {a1.x - a3.x, a1.y - a3.y};
  } else {  // R == 3
    const T s3 = T(0.86602540378443864676372317075294);
    cx<T> sm = {v[1].x + v[2].x, v[1].y + v[2].y};
    cx<T> df = {v[1].x - v[2].x, v[1].y - v[2].y};
    cx<T> t1 = {v[0].x - T(0.5) * sm.x, v[0].y - T(0.5) * sm.y};
    T f = sign < 0 ? -s3 : s3;
    cx<T> t2 = {-f * df.y, f * df.x};
    v[0] = {v[0].x + sm.x, v[0].y + sm.y};
    v[1] = {t1.x + t2.x, t1.y + t2.y};
    v[2] = {t1.x - t2.x, t1.y - t2.y};
  }
}

// One Stockham stage over `ntr` transforms.  Every thread of the block must
// call it (it contains block barriers).  Requires ntr*(n/R) <= MAXB*blockDim.
template <typename T, int R, bool LANE_T, int MAXB>
__device__ __forceinline__ void fft_stage(cx<T>* buf, int n, int es, int ts, int ntr, int Ns,
                                          const cx<T>* tw, int sign) {
  const int nb = n / R;
  const int items = ntr * nb;
  const int tws = n / (Ns * R);
  cx<T> v[MAXB][R];
  int tt[MAXB], dd[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int w = threadIdx.x + i * blockDim.x;
    tt[i] = -1;
    dd[i] = 0;
    if (w < items) {
      const int t = LANE_T ? (w % ntr) : (w / nb);
      const int b = LANE_T ? (w / ntr) : (w % nb);
      const int k = b % Ns;
      cx<T>* base = buf + (long)t * ts;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> a = base[(b + r * nb) * es];
        if (r > 0 && k > 0) {
          cx<T> wv = tw[r * k * tws];
          if (sign > 0) wv.y = -wv.y;
          a = cmul(a, wv);
        }
        v[i][r] = a;
      }
      butterfly<T, R>(v[i], sign);
      tt[i] = t;
      dd[i] = (b / Ns) * Ns * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    if (tt[i] >= 0) {
      cx<T>* base = buf + (long)tt[i] * ts;
#pragma unroll
      for (int r = 0; r < R; ++r) base[(dd[i] + r * Ns) * es] = v[i][r];
    }
  }
  __syncthreads();
}

// Full in-place transform of `ntr` transforms held in LDS.
template <typename T, bool LANE_T, int MAXB>
__device__ void fft_lds(cx<T>* buf, const FftPlan& p, int es, int ts, int ntr, const cx<T>* tw,
                        int sign) {
  int Ns = 1;
  for (int s = 0; s < p.ns; ++s) {
    const int R = p.rad[s];
    if (R == 4)
      fft_stage<T, 4, LANE_T, MAXB>(buf, p.n, es, ts, ntr, Ns, tw, sign);
    else if (R == 2)
      fft_stage<T, 2, LANE_T, MAXB>(buf, p.n, es, ts, ntr, Ns, tw, sign);
    else
      fft_stage<T, 3, LANE_T, MAXB>(buf, p.n, es, ts, ntr, Ns, tw, sign);
    Ns *= R;
  }
}

// Copy a global twiddle table of length n into LDS (all threads call).
template <typename T>
__device__ __forceinline__ void load_twiddles(cx<T>* dst, const cx<T>* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}
// Device twiddle arrays hold the base table exp(-2 pi i k/n), k < n (run-time
// lengths), followed by the stage-major table (compile-time lengths).
template <typename T>
__device__ __forceinline__ void load_twiddles_for(cx<T>* dst, const cx<T>* src, int n, bool stage_major) {
  if (stage_major)
    load_twiddles(dst, src + n, n - 1);
  else
    load_twiddles(dst, src, n);
}

}  // namespace les

namespace les {

// ---- compile-time-length variant: every index is a constant expression ----
constexpr int fft_nstages(int n) {
  int s = 0;
  while (n % 4 == 0) { n /= 4; ++s; }
  while (n % 2 == 0) { n /= 2; ++s; }
  while (n % 3 == 0) { n /= 3; ++s; }
  return n == 1 ? s : -1;
}
constexpr int fft_radix(int n, int s) {
  int i = 0;
  while (n % 4 == 0) { if (i == s) return 4; n /= 4; ++i; }
  while (n % 2 == 0) { if (i == s) return 2; n /= 2; ++i; }
  while (n % 3 == 0) { if (i == s) return 3; n /= 3; ++i; }
  return 1;
}

// Stage-major twiddle table of the compile-time transforms: stage s with
// radix R and NS = product of the earlier radices stores W^(r k N/(NS R)) at
// [fft_tw_off(N, s) + (r-1) NS + k], r = 1..R-1, k = 0..NS-1 (N-1 entries in
// all).  Lanes that work on consecutive k read consecutive words, where a
// single base table exp(-2 pi i j/N) indexed by r k N/(NS R) makes them
// stride through the LDS banks.
constexpr int fft_tw_off(int n, int s) {
  int off = 0, ns = 1;
  for (int i = 0; i < s; ++i) {
    const int r = fft_radix(n, i);
    off += (r - 1) * ns;
    ns *= r;
  }
  return off;
}

// Padded LDS index of element e inside one contiguous (ES = 1) transform:
// one spare slot every 16 complex words, so the stride-R writes of the early
// Stockham stages spread over the LDS banks instead of piling onto a few.
// A padded transform occupies lds_row(N) words.
__host__ __device__ constexpr int lds_pad(int e) { return e + (e >> 4); }
__host__ __device__ constexpr int lds_row(int n) { return n + (n >> 4) + 1; }

// One Stockham stage of NTR transforms of length N in LDS (element e of
// transform t at buf[t*TS + e*ES], or buf[t*TS + lds_pad(e)] with PAD)
// executed by BLOCK threads.
template <typename T, int N, int R, int NS, int ES, int TS, int NTR, int BLOCK, bool LANE_T, bool PAD, int OFF>
__device__ __forceinline__ void fft_stage_c(cx<T>* buf, const cx<T>* tw, int sign) {
  constexpr int NB = N / R;
  constexpr int ITEMS = NTR * NB;
  constexpr int MAXB = (ITEMS + BLOCK - 1) / BLOCK;
  auto at = [](int e) { return PAD ? lds_pad(e) : e * ES; };
  cx<T> v[MAXB][R];
  int dd[MAXB], tt[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int w = threadIdx.x + i * BLOCK;
    tt[i] = -1;
    dd[i] = 0;
    if (ITEMS % BLOCK == 0 || w < ITEMS) {
      const int t = LANE_T ? (w % NTR) : (w / NB);
      const int b = LANE_T ? (w / NTR) : (w % NB);
      const int k = b % NS;
      cx<T>* base = buf + t * TS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> a = base[at(b + r * NB)];
        if (r > 0 && NS > 1 && k > 0) {
          cx<T> wv = tw[OFF + (r - 1) * NS + k];
          if (sign > 0) wv.y = -wv.y;
          a = cmul(a, wv);
        }
        v[i][r] = a;
      }
      butterfly<T, R>(v[i], sign);
      tt[i] = t;
      dd[i] = (b / NS) * NS * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    if (tt[i] >= 0) {
      cx<T>* base = buf + tt[i] * TS;
#pragma unroll
      for (int r = 0; r < R; ++r) base[at(dd[i] + r * NS)] = v[i][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int S, int NS, int ES, int TS, int NTR, int BLOCK, bool LANE_T, bool PAD>
__device__ __forceinline__ void fft_stages_c(cx<T>* buf, const cx<T>* tw, int sign) {
  if constexpr (S < fft_nstages(N)) {
    constexpr int R = fft_radix(N, S);
    fft_stage_c<T, N, R, NS, ES, TS, NTR, BLOCK, LANE_T, PAD, fft_tw_off(N, S)>(buf, tw, sign);
    fft_stages_c<T, N, S + 1, NS * R, ES, TS, NTR, BLOCK, LANE_T, PAD>(buf, tw, sign);
  }
}

// tw: the stage-major table of N (fft_tw_off).
// PAD (with ES = 1, TS = lds_row(N)): padded contiguous transforms
template <typename T, int N, int ES, int TS, int NTR, int BLOCK, bool LANE_T, bool PAD = false>
__device__ __forceinline__ void fft_lds_c(cx<T>* buf, const cx<T>* tw, int sign) {
  static_assert(fft_nstages(N) > 0, "FFT length must be 2^a 3^b");
  static_assert(!PAD || (ES == 1 && TS >= lds_row(N)), "padded transforms are contiguous");
  fft_stages_c<T, N, 0, 1, ES, TS, NTR, BLOCK, LANE_T, PAD>(buf, tw, sign);
}

}  // namespace les

#pragma clang fp contract(fast)
