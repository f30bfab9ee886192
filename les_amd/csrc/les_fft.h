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

#include <cmath>

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

// complex product with two fused multiply-adds (explicit, so the rounding
// does not depend on what the compiler would contract)
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {std::fma(a.x, b.x, -(a.y * b.y)), std::fma(a.x, b.y, a.y * b.x)};
}

struct FftPlan {
  int n;        // transform length
  int ns;       // number of stages
  int rad[12];  // radices, product == n
};

// R-point DFT in registers, sign -1 (forward) or +1 (backward).
template <typename T, int R>
__host__ __device__ __forceinline__ void butterfly(cx<T> (&v)[R], int sign) {
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
// lengths), then the stage-major table of each compile-time plan, n entries
// apart (plan p at n + p n).
template <typename T>
__device__ __forceinline__ void load_twiddles_for(cx<T>* dst, const cx<T>* src, int n, bool stage_major,
                                                  int plan = 0) {
  if (stage_major)
    load_twiddles(dst, src + n + plan * n, n - 1);
  else
    load_twiddles(dst, src, n);
}

}  // namespace les

namespace les {

// ---- compile-time-length variant: every index is a constant expression ----
// Two plans.  Plan 0: radix-4, -2, -3 Stockham passes.  Plan 1: two passes
// through LDS, N = R0 * R1 with R0 = 16 (8 where that fits better) and
// R1 <= 24, each an in-register R-point DFT (reg_dft); the first pass writes
// at stride 16, which the lds_pad spare slot every 16 words keeps free of
// bank conflicts, the second writes where it read.  Plan 1 moves 2-2.5x less
// data through LDS but gives one transform fewer threads: it pays where a
// block holds many transforms (k_pad_products), plan 0 elsewhere.
__host__ __device__ constexpr bool fft_smooth(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}
__host__ __device__ constexpr int fft_first_radix(int n) {
  if (n <= 24) return n;
  if (n % 16 == 0 && n / 16 <= 24) return 16;
  if (n % 8 == 0 && n / 8 <= 24) return 8;
  return 0;  // no two-pass plan
}
// in-register DFT sizes (reg_dft)
__host__ __device__ constexpr bool fft_reg_size(int r) {
  return r == 2 || r == 3 || r == 4 || r == 6 || r == 8 || r == 9 || r == 12 || r == 16 || r == 18 || r == 24;
}
// plan 2: radix-8 passes while more than 24 points remain, then one pass of
// the rest (256 = 8 8 4, 384 = 8 8 6): the thread count of plan 0 with one
// LDS pass in two or three fewer
__host__ __device__ constexpr int fft_r8_stages(int n) {
  int s = 0;
  while (n > 24 && n % 8 == 0) { n /= 8; ++s; }
  if (n == 1) return s;
  return fft_reg_size(n) ? s + 1 : -1;
}
__host__ __device__ constexpr int fft_r8_radix(int n, int s) {
  int i = 0;
  while (n > 24 && n % 8 == 0) { if (i == s) return 8; n /= 8; ++i; }
  return i == s ? n : 1;
}
// plan 0: radix 4, then 2, then 3 passes (many threads per transform, short
// register DFTs); plan 1: the two-pass plan; plan 2: radix-8 passes.  A
// length without a plan-1/2 decomposition uses plan 0.
__host__ __device__ constexpr int fft_nstages(int n, int plan = 0) {
  if (!fft_smooth(n) || n < 2) return -1;
  const int r0 = plan == 1 ? fft_first_radix(n) : 0;
  if (r0 == n) return 1;
  if (r0 > 0) return 2;
  if (plan == 2 && fft_r8_stages(n) > 0) return fft_r8_stages(n);
  int s = 0;
  while (n % 4 == 0) { n /= 4; ++s; }
  while (n % 2 == 0) { n /= 2; ++s; }
  while (n % 3 == 0) { n /= 3; ++s; }
  return s;
}
__host__ __device__ constexpr int fft_radix(int n, int s, int plan = 0) {
  const int r0 = plan == 1 ? fft_first_radix(n) : 0;
  if (r0 > 0) return s == 0 ? r0 : (s == 1 && r0 < n ? n / r0 : 1);
  if (plan == 2 && fft_r8_stages(n) > 0) return fft_r8_radix(n, s);
  int i = 0;
  while (n % 4 == 0) { if (i == s) return 4; n /= 4; ++i; }
  while (n % 2 == 0) { if (i == s) return 2; n /= 2; ++i; }
  while (n % 3 == 0) { if (i == s) return 3; n /= 3; ++i; }
  return 1;
}
// most butterflies per transform over the passes (threads one transform uses)
__host__ __device__ constexpr int fft_max_nb(int n, int plan = 0) {
  int m = 0;
  for (int s = 0; s < fft_nstages(n, plan); ++s)
    m = n / fft_radix(n, s, plan) > m ? n / fft_radix(n, s, plan) : m;
  return m;
}
constexpr int FFT_NPLANS = 3;

// cos(2 pi j / 144) (all the in-register DFT sizes divide 144), from a
// quarter-period table evaluated to 22 digits
__host__ __device__ constexpr double dft_cos144_q(int j) {
  switch (j) {
    case 0: return 1.0000000000000000000000e+0;
    case 1: return 9.9904822158185776240372e-1;
    case 2: return 9.9619469809174553229501e-1;
    case 3: return 9.9144486137381041114456e-1;
    case 4: return 9.8480775301220805936674e-1;
    case 5: return 9.7629600711993336597089e-1;
    case 6: return 9.6592582628906828674974e-1;
    case 7: return 9.5371695074822692114385e-1;
    case 8: return 9.3969262078590838405411e-1;
    case 9: return 9.2387953251128675612818e-1;
    case 10: return 9.0630778703664996324255e-1;
    case 11: return 8.8701083317822170105461e-1;
    case 12: return 8.6602540378443864676372e-1;
    case 13: return 8.4339144581288570127286e-1;
    case 14: return 8.1915204428899178968449e-1;
    case 15: return 7.9335334029123516457978e-1;
    case 16: return 7.6604444311897803520239e-1;
    case 17: return 7.3727733681012404138429e-1;
    case 18: return 7.0710678118654752440084e-1;
    case 19: return 6.7559020761566024434834e-1;
    case 20: return 6.4278760968653932632264e-1;
    case 21: return 6.0876142900872063941610e-1;
    case 22: return 5.7357643635104609610803e-1;
    case 23: return 5.3729960834682383184079e-1;
    case 24: return 5.0000000000000000000000e-1;
    case 25: return 4.6174861323503393056293e-1;
    case 26: return 4.2261826174069943618698e-1;
    case 27: return 3.8268343236508977172846e-1;
    case 28: return 3.4202014332566873304410e-1;
    case 29: return 3.0070579950427312162255e-1;
    case 30: return 2.5881904510252076234890e-1;
    case 31: return 2.1643961393810287975955e-1;
    case 32: return 1.7364817766693034885172e-1;
    case 33: return 1.3052619222005159154841e-1;
    case 34: return 8.7155742747658173558064e-2;
    case 35: return 4.3619387365335999781753e-2;
    case 36: return 0.0;
    default: return 0.0;
  }
}
__host__ __device__ constexpr double dft_cos144(int j) {
  j %= 144;
  if (j < 0) j += 144;
  if (j <= 36) return dft_cos144_q(j);
  if (j <= 72) return -dft_cos144_q(72 - j);
  if (j <= 108) return -dft_cos144_q(j - 72);
  return dft_cos144_q(144 - j);
}
__host__ __device__ constexpr double dft_sin144(int j) { return dft_cos144(j - 36); }

// a * W_R^m, W_R = exp(sign 2 pi i / R); quarter turns exactly
template <typename T, int R>
__host__ __device__ __forceinline__ cx<T> twiddle_const(cx<T> a, int m, int sign) {
  m %= R;
  if (m == 0) return a;
  if (4 * m == R) return sign < 0 ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x};
  if (2 * m == R) return cx<T>{-a.x, -a.y};
  if (4 * m == 3 * R) return sign < 0 ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
  const int j = m * (144 / R);
  const T c = (T)dft_cos144(j), sn = (T)(sign < 0 ? -dft_sin144(j) : dft_sin144(j));
  return cx<T>{std::fma(a.x, c, -(a.y * sn)), std::fma(a.x, sn, a.y * c)};
}

// R = P Q split of the in-register DFT sizes
template <int R>
struct DftSplit {
  static constexpr int P = R == 6 ? 2 : R == 8 ? 4 : R == 9 ? 3 : R == 12 ? 4 : R == 16 ? 4 : R == 18 ? 6 : R == 24 ? 8 : R;
};

// R-point DFT in registers: radix 2/3/4 butterflies, larger R by one
// Cooley-Tukey split R = P Q (n = Q n1 + n2, k = k1 + P k2): P-point DFTs over
// n1, the twiddles W_R^(n2 k1), Q-point DFTs over n2
template <typename T, int R>
__host__ __device__ __forceinline__ void reg_dft(cx<T> (&v)[R], int sign) {
  if constexpr (R == 2 || R == 3 || R == 4) {
    butterfly<T, R>(v, sign);
  } else if constexpr (R > 4) {
    constexpr int P = DftSplit<R>::P, Q = R / P;
    static_assert(P > 1 && Q > 1 && P * Q == R, "no split for this DFT size");
    cx<T> y[Q][P];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      cx<T> t[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) t[n1] = v[Q * n1 + n2];
      reg_dft<T, P>(t, sign);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) y[n2][k1] = twiddle_const<T, R>(t[k1], n2 * k1, sign);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      cx<T> t[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) t[n2] = y[n2][k1];
      reg_dft<T, Q>(t, sign);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
    }
  }
}

// Stage-major twiddle table of the compile-time transforms: stage s with
// radix R and NS = product of the earlier radices stores W^(r k N/(NS R)) at
// [fft_tw_off(N, s) + (r-1) NS + k], r = 1..R-1, k = 0..NS-1 (N-1 entries in
// all).  Lanes that work on consecutive k read consecutive words, where a
// single base table exp(-2 pi i j/N) indexed by r k N/(NS R) makes them
// stride through the LDS banks.
constexpr int fft_tw_off(int n, int s, int plan = 0) {
  int off = 0, ns = 1;
  for (int i = 0; i < s; ++i) {
    const int r = fft_radix(n, i, plan);
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

// Whether, in a stage of radix R after NS points, every butterfly's R padded
// positions lie at constant distances r s + (r s)/16 from its first one
// (reads: s = N/R from b; writes: s = NS from the output base), so the
// addresses fold into the LDS instructions' immediate offsets.
__host__ __device__ constexpr bool pad_offsets_const(int n, int R, int NS, bool rd) {
  const int NB = n / R, s = rd ? NB : NS;
  for (int b = 0; b < NB; ++b) {
    const int e0 = rd ? b : (b / NS) * NS * R + b % NS;
    for (int r = 0; r < R; ++r)
      if (lds_pad(e0 + r * s) - lds_pad(e0) != r * s + ((r * s) >> 4)) return false;
  }
  return true;
}

// Twiddle r, k of the stage whose entries start at OFF of the stage-major
// table: from the table itself, or from a kernel's own placement of it (a
// class with a member template at<OFF, NS>(r, k)).
template <int OFF, int NS, typename T>
__device__ __forceinline__ cx<T> tw_at(const cx<T>* tw, int r, int k) {
  return tw[OFF + (r - 1) * NS + k];
}
template <int OFF, int NS, typename T>
__device__ __forceinline__ cx<T> tw_at(cx<T>* tw, int r, int k) {
  return tw[OFF + (r - 1) * NS + k];
}
template <int OFF, int NS, class TW>
__device__ __forceinline__ auto tw_at(const TW& tw, int r, int k) -> decltype(tw.template at<OFF, NS>(r, k)) {
  return tw.template at<OFF, NS>(r, k);
}

// One Stockham stage of NTR transforms of length N in LDS (element e of
// transform t at buf[t*TS + e*ES], or buf[t*TS + lds_pad(e)] with PAD)
// executed by BLOCK threads.
template <typename T, int N, int R, int NS, int ES, int TS, int NTR, int BLOCK, bool LANE_T, bool PAD, int OFF,
          class TW>
__device__ __forceinline__ void fft_stage_c(cx<T>* buf, const TW& tw, int sign) {
  constexpr int NB = N / R;
  constexpr int ITEMS = NTR * NB;
  constexpr int MAXB = (ITEMS + BLOCK - 1) / BLOCK;
  constexpr bool RD_C = !PAD || pad_offsets_const(N, R, NS, true);
  constexpr bool WR_C = !PAD || pad_offsets_const(N, R, NS, false);
  auto at = [](int e) { return PAD ? lds_pad(e) : e * ES; };
  // position of element e0 + d: a compile-time distance from e0's where that holds
  auto at_rd = [&](int e0, int d) { return RD_C ? at(e0) + (PAD ? d + (d >> 4) : d * ES) : at(e0 + d); };
  auto at_wr = [&](int e0, int d) { return WR_C ? at(e0) + (PAD ? d + (d >> 4) : d * ES) : at(e0 + d); };
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
        cx<T> a = base[at_rd(b, r * NB)];
        // k = 0 multiplies by the table's exact (1, -0): no branch needed
        if constexpr (NS > 1) {
          if (r > 0) {
            cx<T> wv = tw_at<OFF, NS>(tw, r, k);
            if (sign > 0) wv.y = -wv.y;
            a = cmul(a, wv);
          }
        }
        v[i][r] = a;
      }
      reg_dft<T, R>(v[i], sign);
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
      for (int r = 0; r < R; ++r) base[at_wr(dd[i], r * NS)] = v[i][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int S, int NS, int ES, int TS, int NTR, int BLOCK, bool LANE_T, bool PAD, int PLAN,
          class TW>
__device__ __forceinline__ void fft_stages_c(cx<T>* buf, const TW& tw, int sign) {
  if constexpr (S < fft_nstages(N, PLAN)) {
    constexpr int R = fft_radix(N, S, PLAN);
    fft_stage_c<T, N, R, NS, ES, TS, NTR, BLOCK, LANE_T, PAD, fft_tw_off(N, S, PLAN)>(buf, tw, sign);
    fft_stages_c<T, N, S + 1, NS * R, ES, TS, NTR, BLOCK, LANE_T, PAD, PLAN>(buf, tw, sign);
  }
}

// PAD (with ES = 1, TS = lds_row(N)): padded contiguous transforms
// tw: the stage-major table of N for PLAN (load_twiddles_for), or an object
// that places that table elsewhere (tw_at)
template <typename T, int N, int ES, int TS, int NTR, int BLOCK, bool LANE_T, bool PAD = false, int PLAN = 0,
          class TW = const cx<T>*>
__device__ __forceinline__ void fft_lds_c(cx<T>* buf, const TW& tw, int sign) {
  static_assert(fft_nstages(N, PLAN) > 0, "FFT length must be 2^a 3^b");
  static_assert(!PAD || (ES == 1 && TS >= lds_row(N)), "padded transforms are contiguous");
  fft_stages_c<T, N, 0, 1, ES, TS, NTR, BLOCK, LANE_T, PAD, PLAN>(buf, tw, sign);
}

// ---- one transform per wavefront ----
// A wavefront's LDS accesses execute in program order, so a transform that
// one wavefront owns needs no workgroup barrier between its passes, only a
// fence that keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Stockham stage (radix R after NS points) of one padded N-point row,
// run by the 64 lanes of one wavefront (lane b owns butterflies b, b + 64 ..).
// The same arithmetic as fft_stage_c, so the results are identical.
template <typename T, int N, int R, int NS, int OFF, class TW>
__device__ __forceinline__ void wave_stage(cx<T>* row, const TW& tw, int lane, int sign) {
  constexpr int NB = N / R, MAXB = (NB + 63) / 64;
  constexpr bool RD_C = pad_offsets_const(N, R, NS, true);
  constexpr bool WR_C = pad_offsets_const(N, R, NS, false);
  auto at_rd = [](int e0, int d) { return RD_C ? lds_pad(e0) + d + (d >> 4) : lds_pad(e0 + d); };
  auto at_wr = [](int e0, int d) { return WR_C ? lds_pad(e0) + d + (d >> 4) : lds_pad(e0 + d); };
  cx<T> v[MAXB][R];
  int dd[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = lane + 64 * i;
    dd[i] = 0;
    if (NB % 64 == 0 || b < NB) {
      const int k = b % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> a = row[at_rd(b, r * NB)];
        if constexpr (NS > 1) {
          if (r > 0) {
            cx<T> wv = tw_at<OFF, NS>(tw, r, k);
            if (sign > 0) wv.y = -wv.y;
            a = cmul(a, wv);
          }
        }
        v[i][r] = a;
      }
      reg_dft<T, R>(v[i], sign);
      dd[i] = (b / NS) * NS * R + k;
    }
  }
  wave_sync();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = lane + 64 * i;
    if (NB % 64 == 0 || b < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) row[at_wr(dd[i], r * NS)] = v[i][r];
    }
  }
  wave_sync();
}

// A stage with at most 64 butterflies, one per lane, with two options that
// leave the arithmetic unchanged: SWAP34 exchanges bits 3 and 4 of the lane
// index below 32 in the lane -> butterfly map (in the second radix-8 pass of
// a 384-point transform it puts butterfly groups whose writes fall on
// disjoint LDS banks in one 16-lane write group), and RTW takes the stage's
// twiddles, already conjugated for sign +1, from the registers rtw[r - 1].
// ZLO <= ZHI: the row is known to hold zeros at indices [ZLO, ZHI] (the
// first pass of a zero-padded spectrum), which are taken as zero without
// being read (the same arithmetic on the same values; those words need not
// be written).
template <typename T, int N, int R, int NS, int OFF, bool SWAP34, bool RTW, int ZLO = 0, int ZHI = -1, class TW>
__device__ __forceinline__ void wave_stage_lane(cx<T>* row, const TW& tw, int lane, int sign, const cx<T>* rtw) {
  constexpr int NB = N / R;
  static_assert(NB <= 64, "one butterfly per lane");
  constexpr bool RD_C = pad_offsets_const(N, R, NS, true);
  constexpr bool WR_C = pad_offsets_const(N, R, NS, false);
  auto at_rd = [](int e0, int d) { return RD_C ? lds_pad(e0) + d + (d >> 4) : lds_pad(e0 + d); };
  auto at_wr = [](int e0, int d) { return WR_C ? lds_pad(e0) + d + (d >> 4) : lds_pad(e0 + d); };
  const int b = SWAP34 && lane < 32 ? ((lane & ~24) | ((lane & 8) << 1) | ((lane & 16) >> 1)) : lane;
  cx<T> v[R];
  int dd = 0;
  if (NB == 64 || b < NB) {
    const int k = b % NS;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cx<T> a = {T(0), T(0)};
      if (ZLO > ZHI || r * NB + NB - 1 < ZLO || r * NB > ZHI) {
        a = row[at_rd(b, r * NB)];
      } else if (r * NB < ZLO || r * NB + NB - 1 > ZHI) {  // part of this read lies in the zero band
        const int e = b + r * NB;
        if (e < ZLO || e > ZHI) a = row[at_rd(b, r * NB)];
      }
      if constexpr (NS > 1) {
        if (r > 0) {
          cx<T> wv;
          if constexpr (RTW) {
            wv = rtw[r - 1];
          } else {
            wv = tw_at<OFF, NS>(tw, r, k);
            if (sign > 0) wv.y = -wv.y;
          }
          a = cmul(a, wv);
        }
      }
      v[r] = a;
    }
    reg_dft<T, R>(v, sign);
    dd = (b / NS) * NS * R + k;
  }
  wave_sync();
  if (NB == 64 || b < NB) {
#pragma unroll
    for (int r = 0; r < R; ++r) row[at_wr(dd, r * NS)] = v[r];
  }
  wave_sync();
}

// Stages S .. SEND-1 of the compile-time plan, by one wavefront.
template <typename T, int N, int PLAN, int S, int SEND, int NS = 1, class TW>
__device__ __forceinline__ void wave_fft(cx<T>* row, const TW& tw, int lane, int sign) {
  if constexpr (S < SEND) {
    constexpr int R = fft_radix(N, S, PLAN);
    wave_stage<T, N, R, NS, fft_tw_off(N, S, PLAN)>(row, tw, lane, sign);
    wave_fft<T, N, PLAN, S + 1, SEND, NS * R>(row, tw, lane, sign);
  }
}

// The last stage of the plan when it has exactly 64 butterflies: lane k
// returns its outputs k + 64 r (r < R) in v instead of writing them; with
// rtw, the twiddles come from registers (conjugated for sign +1, as in
// wave_stage_lane).
template <typename T, int N, int PLAN, class TW>
__device__ __forceinline__ void wave_last_stage(const cx<T>* row, const TW& tw, int lane, int sign,
                                                cx<T> (&v)[fft_radix(N, fft_nstages(N, PLAN) - 1, PLAN)],
                                                const cx<T>* rtw = nullptr) {
  constexpr int S = fft_nstages(N, PLAN) - 1, R = fft_radix(N, S, PLAN), NS = N / R, OFF = fft_tw_off(N, S, PLAN);
  static_assert(NS == 64, "one butterfly per lane in the last stage");
  constexpr bool RD_C = pad_offsets_const(N, R, NS, true);
  auto at_rd = [](int e0, int d) { return RD_C ? lds_pad(e0) + d + (d >> 4) : lds_pad(e0 + d); };
#pragma unroll
  for (int r = 0; r < R; ++r) {
    cx<T> a = row[at_rd(lane, r * NS)];
    if (r > 0) {
      cx<T> wv;
      if (rtw) {
        wv = rtw[r - 1];
      } else {
        wv = tw_at<OFF, NS>(tw, r, lane);
        if (sign > 0) wv.y = -wv.y;
      }
      a = cmul(a, wv);
    }
    v[r] = a;
  }
  reg_dft<T, R>(v, sign);
}

}  // namespace les

#pragma clang fp contract(fast)
