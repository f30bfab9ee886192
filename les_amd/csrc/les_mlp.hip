// les_mlp.hip -- the learned SGS closure on gfx950 (no counterpart in the
// reference code): an 8 -> 64 -> 64 -> 64 -> 2 MLP with SiLU that maps local
// strain invariants to the Lagrangian scale-dependent model's (Cs_opt2, Beta)
// (lagrange_Sdep.f90:390-421), trained on the GPU on pairs that procedure
// produces, with hand-written BF16 MFMA / FP32-accumulate kernels.
//
// MFMA mapping.  Activations are kept transposed, features x points: a
// 16 x 16 tile (16 features of 16 points) is one v_mfma_f32_16x16x32_bf16
// result, lane l = 16 g + c holding features 4g..4g+3 of point c.  That is
// exactly the lane/k assignment of the B operand of the next layer, so a
// layer's output feeds the next layer with no data movement: for a 64-wide
// layer the two K = 32 steps take output tiles (0,1) and (2,3), elements 0..3
// from the first tile and 4..7 from the second (logical k index kidx below).
// The weights are the A operands, pre-arranged per lane in that same k order
// ("packed operands", k_mlp_pack), so the order of the k sum is irrelevant.
// Weight gradients dW = dZ H^T reduce over points; a 16 x 16 LDS transpose
// per tile puts points on the k index of v_mfma_f32_16x16x16bf16_1k, and the
// bias gradients are the same products with an all-ones B operand.
//
// Rounding points (the CPU reference reproduces them): inputs, weights,
// activations, dy and dz are rounded to bf16 before each product; products
// accumulate in FP32; biases, SiLU and the optimiser run in FP32.
#include "les_xform.h"

namespace les {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MI = 8, MH = 64, MO = 2;
constexpr int OW1 = 0, OB1 = OW1 + MH * MI, OW2 = OB1 + MH, OB2 = OW2 + MH * MH, OW3 = OB2 + MH,
              OB3 = OW3 + MH * MH, OW4 = OB3 + MH, OB4 = OW4 + MO * MH, MLP_NP = OB4 + MO;
// packed A operands, 64 lanes x 16 bytes each
constexpr int OP_FW1 = 0, OP_FW2 = 4, OP_FW3 = 12, OP_FW4 = 20, OP_BW4 = 22, OP_BW3 = 26, OP_BW2 = 34,
              NOPS = 42, NOPS_FWD = 22;
constexpr int NBIAS = 3 * MH + MO;  // b1 b2 b3 b4 in LDS
constexpr int TS = 20;              // LDS transpose row (16 points + 4 pad), 40 B rows keep 8-B alignment

struct MlpNorm {  // x_n = (x - mu) * isd ; y = y_n * ysd + ymu
  float mu[MI], isd[MI], ymu[MO], ysd[MO];
};

struct MlpModel {
  float* theta = nullptr;  // FP32 master parameters (MLP_NP)
  float* m = nullptr;      // Adam moments
  float* v = nullptr;
  uint4* ops = nullptr;    // NOPS x 64 packed bf16 A operands
  float* part = nullptr;   // per-wave gradient partials
  float* lpart = nullptr;
  float* grad = nullptr;   // reduced gradient (MLP_NP)
  double* judge = nullptr; // [sum (a-b)^2, sum b^2, err, use_mlp] of the last dynamic step
  int nwaves = 0;
  long adam_t = 0;
  MlpNorm norm;
};

__device__ __forceinline__ int kidx(int s, int g, int e) { return 16 * (2 * s + (e >> 2)) + 4 * g + (e & 3); }

// SiLU with the hardware exponential (v_exp_f32 through __expf) and
// reciprocal (v_rcp_f32): a few ulp of float, far below the bf16 rounding
// the result goes through; for z < -88 the reciprocal of inf gives -0
__device__ __forceinline__ float sigmoid_f(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }
__device__ __forceinline__ float dsilu_f(float z) {
  const float s = sigmoid_f(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ short bf_bits(float x) { return __builtin_bit_cast(short, (__bf16)x); }
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 ldop(const uint4* ops, int op, int lane) {
  return __builtin_bit_cast(bf16x8, ops[op * 64 + lane]);
}

// one (operand, lane): the 8 bf16 A-operand values in kidx order
__global__ void k_mlp_pack(const float* __restrict__ th, uint4* __restrict__ ops) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= NOPS * 64) return;
  const int op = id >> 6, l = id & 63, g = l >> 4, c = l & 15;
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    float w = 0.0f;
    if (op < OP_FW2) {  // W1 (64 x 8): k = 8g + e, only g = 0 carries the 8 inputs
      if (g == 0) w = th[OW1 + (16 * op + c) * MI + e];
    } else if (op < OP_FW3) {
      const int j = op - OP_FW2, t = j >> 1, s = j & 1;
      w = th[OW2 + (16 * t + c) * MH + kidx(s, g, e)];
    } else if (op < OP_FW4) {
      const int j = op - OP_FW3, t = j >> 1, s = j & 1;
      w = th[OW3 + (16 * t + c) * MH + kidx(s, g, e)];
    } else if (op < OP_BW4) {
      const int s = op - OP_FW4;
      if (c < MO) w = th[OW4 + c * MH + kidx(s, g, e)];
    } else if (op < OP_BW3) {  // W4^T: k = output index 4g + e (rows of the dy tile)
      const int t = op - OP_BW4, o = 4 * g + e;
      if (e < 4 && o < MO) w = th[OW4 + o * MH + 16 * t + c];
    } else if (op < OP_BW2) {
      const int j = op - OP_BW3, t = j >> 1, s = j & 1;
      w = th[OW3 + kidx(s, g, e) * MH + 16 * t + c];
    } else {
      const int j = op - OP_BW2, t = j >> 1, s = j & 1;
      w = th[OW2 + kidx(s, g, e) * MH + 16 * t + c];
    }
    v[e] = (__bf16)w;
  }
  ops[id] = __builtin_bit_cast(uint4, v);
}

// activations of one 16-point tile: pre-activations z1..z3 (4 tiles of 16
// features each) and the output tile y (rows 0, 1 in lanes 0..15)
struct MlpTile {
  f32x4 z[3][4];
  f32x4 y;
};

// B operand of a 64-wide layer from the previous D tiles, step s
__device__ __forceinline__ bf16x8 act_operand(const f32x4 z[4], int s) {
  bf16x8 b;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    b[e] = (__bf16)silu_f(z[2 * s][e]);
    b[4 + e] = (__bf16)silu_f(z[2 * s + 1][e]);
  }
  return b;
}

__device__ void mlp_forward_tile(const uint4* ops, const float* sb, bf16x8 xin, MlpTile& T, int lane) {
  const int g = lane >> 4;
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 a = mfma32(ldop(ops, OP_FW1 + t, lane), xin, zero);
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r] += sb[16 * t + 4 * g + r];
    T.z[0][t] = a;
  }
#pragma unroll
  for (int l = 1; l < 3; ++l) {
    const bf16x8 h0 = act_operand(T.z[l - 1], 0), h1 = act_operand(T.z[l - 1], 1);
    const int base = l == 1 ? OP_FW2 : OP_FW3;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 a = mfma32(ldop(ops, base + 2 * t, lane), h0, zero);
      a = mfma32(ldop(ops, base + 2 * t + 1, lane), h1, a);
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] += sb[MH * l + 16 * t + 4 * g + r];
      T.z[l][t] = a;
    }
  }
  const bf16x8 h0 = act_operand(T.z[2], 0), h1 = act_operand(T.z[2], 1);
  f32x4 a = mfma32(ldop(ops, OP_FW4, lane), h0, zero);
  a = mfma32(ldop(ops, OP_FW4 + 1, lane), h1, a);
#pragma unroll
  for (int r = 0; r < 4; ++r) a[r] += (g == 0 && r < MO) ? sb[3 * MH + r] : 0.0f;
  T.y = a;
}

__device__ void load_block_consts(const uint4* __restrict__ ops_g, const float* __restrict__ th, uint4* ops,
                                  float* sb, int nops) {
  for (int i = threadIdx.x; i < nops * 64; i += blockDim.x) ops[i] = ops_g[i];
  for (int i = threadIdx.x; i < NBIAS; i += blockDim.x) {
    int j;
    if (i < MH) j = OB1 + i;
    else if (i < 2 * MH) j = OB2 + i - MH;
    else if (i < 3 * MH) j = OB3 + i - 2 * MH;
    else j = OB4 + i - 3 * MH;
    sb[i] = th[j];
  }
  __syncthreads();
}

// ---- inference: Feat(p, f[8]) gives the raw features of point p; Out(p, y0, y1)
// A wave takes 64 consecutive points: every lane computes the features of one
// point (coalesced loads), the normalised features go through LDS into four
// 16-point MFMA tiles, and the outputs come back through LDS so that every
// lane stores one point.  Trip counts are block-uniform (barriers).
template <class Feat, class Out>
__global__ void __launch_bounds__(256) k_mlp_infer(Feat feat, Out out, long P, const uint4* __restrict__ ops_g,
                                                   const float* __restrict__ th, MlpNorm nm) {
  __shared__ uint4 ops[NOPS_FWD * 64];
  __shared__ float sb[NBIAS];
  __shared__ float xs[4][64][MI + 1];
  __shared__ float ys[4][64][MO];
  load_block_consts(ops_g, th, ops, sb, NOPS_FWD);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
  const long ngroups = (P + 63) / 64, nw = (long)gridDim.x * 4, gw = (long)blockIdx.x * 4 + wave;
  const long iters = (ngroups + nw - 1) / nw;
  for (long it = 0; it < iters; ++it) {
    const long grp = it * nw + gw, p = grp * 64 + lane;
    const bool valid = grp < ngroups && p < P;
    float f[MI];
#pragma unroll
    for (int e = 0; e < MI; ++e) f[e] = 0.0f;
    if (valid) {
      feat(p, f);
#pragma unroll
      for (int e = 0; e < MI; ++e) f[e] = (f[e] - nm.mu[e]) * nm.isd[e];
    }
#pragma unroll
    for (int e = 0; e < MI; ++e) xs[wave][lane][e] = f[e];
    __syncthreads();
#pragma unroll 1
    for (int tt = 0; tt < 4; ++tt) {
      bf16x8 xin;
#pragma unroll
      for (int e = 0; e < 8; ++e) xin[e] = (__bf16)(g == 0 ? xs[wave][16 * tt + c][e] : 0.0f);
      MlpTile T;
      mlp_forward_tile(ops, sb, xin, T, lane);
      if (g == 0) {
        ys[wave][16 * tt + c][0] = T.y[0];
        ys[wave][16 * tt + c][1] = T.y[1];
      }
    }
    __syncthreads();
    if (valid) out(p, ys[wave][lane][0] * nm.ysd[0] + nm.ymu[0], ys[wave][lane][1] * nm.ysd[1] + nm.ymu[1]);
  }
}

struct FeatArray {
  const float* X;
  __device__ void operator()(long p, float* f) const {
#pragma unroll
    for (int e = 0; e < MI; ++e) f[e] = X[p * MI + e];
  }
};
struct OutArray {
  float* Y;
  __device__ void operator()(long p, float y0, float y1) const {
    Y[p * 2] = y0;
    Y[p * 2 + 1] = y1;
  }
};

// ---- training: loss 1/(2 nb) sum |y - t|^2 over the batch, in normalised units
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ s16x4 lds4(const short* row, int g) {
  return *reinterpret_cast<const s16x4*>(row + 4 * g);
}

// samples: full != 0 -> batch element b is sample b (nb <= P); otherwise a
// hash of (seed, step, b) picks a sample of the P.  Each wave accumulates its
// gradient in registers and writes one partial row part[wave][MLP_NP].
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ X, const float* __restrict__ Tt, long P,
                                                  long nb, int full, uint64_t seed, long step,
                                                  const uint4* __restrict__ ops_g, const float* __restrict__ th,
                                                  MlpNorm nm, float* __restrict__ part, float* __restrict__ lpart) {
  __shared__ uint4 ops[NOPS * 64];
  __shared__ float sb[NBIAS];
  __shared__ __align__(16) short tz[4][MH][TS];
  __shared__ __align__(16) short th_[4][MH][TS];
  __shared__ float lred[4][64];
  load_block_consts(ops_g, th, ops, sb, NOPS);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
  const long gw = (long)blockIdx.x * 4 + wave, nw = (long)gridDim.x * 4;
  const long ntiles = (nb + 15) / 16, iters = (ntiles + nw - 1) / nw;  // same trip count in every wave
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  const short one = bf_bits(1.0f);
  const s16x4 ones = {one, one, one, one};
  f32x4 gW1[4], gb1[4], gW2[16], gb2[4], gW3[16], gb3[4], gW4[4], gb4 = zero;
#pragma unroll
  for (int i = 0; i < 4; ++i) { gW1[i] = zero; gb1[i] = zero; gb2[i] = zero; gb3[i] = zero; gW4[i] = zero; }
#pragma unroll
  for (int i = 0; i < 16; ++i) { gW2[i] = zero; gW3[i] = zero; }
  float lsum = 0.0f;
  short(*TZ)[TS] = tz[wave];
  short(*TH)[TS] = th_[wave];
  for (long it = 0; it < iters; ++it) {
    const long tile = it * nw + gw;
    const long b = tile * 16 + c;
    const bool valid = tile < ntiles && b < nb;
    long q = 0;
    if (valid) q = full ? b : (long)(mix64(seed ^ mix64((uint64_t)step * 0x100000001B3ull + (uint64_t)b)) % (uint64_t)P);
    float xf[MI];  // normalised inputs of point c (lanes g == 0), else 0
    float t0 = 0.0f, t1 = 0.0f;
#pragma unroll
    for (int e = 0; e < MI; ++e) xf[e] = 0.0f;
    if (g == 0 && valid) {
#pragma unroll
      for (int e = 0; e < MI; ++e) xf[e] = (X[q * MI + e] - nm.mu[e]) * nm.isd[e];
      t0 = (Tt[q * 2] - nm.ymu[0]) / nm.ysd[0];
      t1 = (Tt[q * 2 + 1] - nm.ymu[1]) / nm.ysd[1];
    }
    bf16x8 xin;
#pragma unroll
    for (int e = 0; e < 8; ++e) xin[e] = (__bf16)xf[e];
    MlpTile T;
    mlp_forward_tile(ops, sb, xin, T, lane);
    float dy0 = 0.0f, dy1 = 0.0f;
    if (g == 0 && valid) {
      const float e0 = T.y[0] - t0, e1 = T.y[1] - t1;
      lsum += 0.5f * (e0 * e0 + e1 * e1) / (float)nb;
      dy0 = e0 / (float)nb;
      dy1 = e1 / (float)nb;
    }
    // ---- output layer: dW4 = dY h3^T, db4, dh3 = W4^T dY
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      TZ[4 * g + r][c] = bf_bits(r == 0 ? dy0 : (r == 1 ? dy1 : 0.0f));
#pragma unroll
      for (int t = 0; t < 4; ++t) TH[16 * t + 4 * g + r][c] = bf_bits(silu_f(T.z[2][t][r]));
    }
    __syncthreads();
    {
      const s16x4 A = lds4(TZ[c], g);
#pragma unroll
      for (int t2 = 0; t2 < 4; ++t2) gW4[t2] = mfma16(A, lds4(TH[16 * t2 + c], g), gW4[t2]);
      gb4 = mfma16(A, ones, gb4);
    }
    bf16x8 bdy;
#pragma unroll
    for (int e = 0; e < 8; ++e) bdy[e] = (__bf16)0.0f;
    bdy[0] = (__bf16)dy0;  // lanes g > 0 hold dy rows 4g.. = 0
    bdy[1] = (__bf16)dy1;
    f32x4 dz[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const f32x4 dh = mfma32(ldop(ops, OP_BW4 + t, lane), bdy, zero);
#pragma unroll
      for (int r = 0; r < 4; ++r) dz[t][r] = dh[r] * dsilu_f(T.z[2][t][r]);
    }
    // ---- hidden layers 3 and 2 (weights W3, W2), then layer 1 (W1)
#pragma unroll
    for (int l = 2; l >= 0; --l) {
      __syncthreads();  // the previous products have read TZ/TH
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int t = 0; t < 4; ++t) TZ[16 * t + 4 * g + r][c] = bf_bits(dz[t][r]);
      if (l > 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int t = 0; t < 4; ++t) TH[16 * t + 4 * g + r][c] = bf_bits(silu_f(T.z[l - 1][t][r]));
      } else if (g < 2) {  // rows 0..7 the normalised inputs, 8..15 zero
#pragma unroll
        for (int e = 0; e < 8; ++e) TH[8 * g + e][c] = g == 0 ? bf_bits(xf[e]) : (short)0;
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const s16x4 A = lds4(TZ[16 * t + c], g);
        if (l == 2) {
#pragma unroll
          for (int t2 = 0; t2 < 4; ++t2) gW3[4 * t + t2] = mfma16(A, lds4(TH[16 * t2 + c], g), gW3[4 * t + t2]);
          gb3[t] = mfma16(A, ones, gb3[t]);
        } else if (l == 1) {
#pragma unroll
          for (int t2 = 0; t2 < 4; ++t2) gW2[4 * t + t2] = mfma16(A, lds4(TH[16 * t2 + c], g), gW2[4 * t + t2]);
          gb2[t] = mfma16(A, ones, gb2[t]);
        } else {
          gW1[t] = mfma16(A, lds4(TH[c], g), gW1[t]);
          gb1[t] = mfma16(A, ones, gb1[t]);
        }
      }
      if (l > 0) {  // dh_{l} = W_{l+1}^T dz, then dz_{l-1}
        const int base = l == 2 ? OP_BW3 : OP_BW2;
        bf16x8 bd[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            bd[s][e] = (__bf16)dz[2 * s][e];
            bd[s][4 + e] = (__bf16)dz[2 * s + 1][e];
          }
        f32x4 nd[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          f32x4 dh = mfma32(ldop(ops, base + 2 * t, lane), bd[0], zero);
          dh = mfma32(ldop(ops, base + 2 * t + 1, lane), bd[1], dh);
#pragma unroll
          for (int r = 0; r < 4; ++r) nd[t][r] = dh[r] * dsilu_f(T.z[l - 1][t][r]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) dz[t] = nd[t];
      }
    }
    __syncthreads();
  }
  // ---- this wave's partial gradient, every parameter written once
  float* pw = part + gw * MLP_NP;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * t + 4 * g + r;
      if (c < MI) pw[OW1 + row * MI + c] = gW1[t][r];
      if (c == 0) {
        pw[OB1 + row] = gb1[t][r];
        pw[OB2 + row] = gb2[t][r];
        pw[OB3 + row] = gb3[t][r];
      }
#pragma unroll
      for (int t2 = 0; t2 < 4; ++t2) {
        pw[OW2 + row * MH + 16 * t2 + c] = gW2[4 * t + t2][r];
        pw[OW3 + row * MH + 16 * t2 + c] = gW3[4 * t + t2][r];
      }
    }
  if (g == 0)
#pragma unroll
    for (int r = 0; r < MO; ++r) {
#pragma unroll
      for (int t2 = 0; t2 < 4; ++t2) pw[OW4 + r * MH + 16 * t2 + c] = gW4[t2][r];
      if (c == 0) pw[OB4 + r] = gb4[r];
    }
  lred[wave][lane] = lsum;
  __syncthreads();
  if (lane == 0) {
    float s = 0.0f;
    for (int i = 0; i < 64; ++i) s += lred[wave][i];
    lpart[gw] = s;
  }
}

// fixed-order sum of the wave partials; optional Adam step (FP32 master)
__global__ void k_mlp_reduce_adam(const float* __restrict__ part, const float* __restrict__ lpart, int nw,
                                  float* __restrict__ grad, float* __restrict__ loss_out, float* th, float* m,
                                  float* v, int adam, float lr, float b1c, float b2c) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0 && loss_out) {
    float s = 0.0f;
    for (int w = 0; w < nw; ++w) s += lpart[w];
    *loss_out = s;
  }
  if (j >= MLP_NP) return;
  float s = 0.0f;
  for (int w = 0; w < nw; ++w) s += part[(long)w * MLP_NP + j];
  if (grad) grad[j] = s;
  if (adam) {  // Adam (Kingma & Ba 2015), beta1 0.9, beta2 0.999, eps 1e-8
    const float mj = 0.9f * m[j] + 0.1f * s, vj = 0.999f * v[j] + 0.001f * s * s;
    m[j] = mj;
    v[j] = vj;
    th[j] -= lr * (mj / b1c) / (sqrtf(vj / b2c) + 1e-8f);
  }
}

// ---- feature / label statistics: per-column mean and standard deviation
__global__ void k_mlp_moments(const float* __restrict__ A, long P, int ncol, double* part) {
  // blockIdx.y = column; NPART blocks per column write (sum, sum of squares)
  const int col = blockIdx.y;
  double s = 0.0, s2 = 0.0;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
    const double a = A[p * ncol + col];
    s += a;
    s2 += a * a;
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = s;
  sh[1][threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((long)col * gridDim.x + blockIdx.x) * 2] = sh[0][0];
    part[((long)col * gridDim.x + blockIdx.x) * 2 + 1] = sh[1][0];
  }
}

// ---- features of the Lagrangian model's filtered fields (one point per p)
// p runs over x < nx, y < ny, levels 0..nz-2 (w-nodes; level 0 the first uvp
// node, as in lagrange_Sdep.f90:131-141).  F2/F4: the 2 Delta / 4 Delta
// filtered fields of lag_filter_mom (fields 9..14 S_ij, 15..20 |S|S_ij).
template <typename T>
struct FeatLag {
  const T *F2, *F4;
  const T* S[6];
  const T* gr[9];  // dudx dudy dudz dvdx dvdy dvdz dwdx dwdy dwdz
  long n3, pl;
  int nx, ny, ld;
  float dz, idelta;
  __device__ long index(long p) const {  // 32-bit: point counts stay below 2^31
    const unsigned nxy = (unsigned)nx * ny, pp = (unsigned)p;
    const unsigned k = pp / nxy, r = pp - k * nxy;
    const unsigned y = r / (unsigned)nx, x = r - y * (unsigned)nx;
    return (long)k * pl + (long)y * ld + x;
  }
  __device__ void operator()(long p, float* f) const {
    const long i = index(p);
    const int k = (int)((unsigned)i / (unsigned)pl);
    double s[6], sb[6], sh[6], ssb[6], ssh[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      s[q] = (double)S[q][i];
      sb[q] = (double)F2[(9 + q) * n3 + i];
      sh[q] = (double)F4[(9 + q) * n3 + i];
      ssb[q] = (double)F2[(15 + q) * n3 + i];
      ssh[q] = (double)F4[(15 + q) * n3 + i];
    }
    auto dot = [](const double* a, const double* b) {
      return a[0] * b[0] + a[3] * b[3] + a[5] * b[5] + 2.0 * (a[1] * b[1] + a[2] * b[2] + a[4] * b[4]);
    };
    // rotation rate with the staggering of calc_Sij (sgs_stag.f90:392-417,475-495)
    double uy, uz, vx, vz, wx, wy;
    if (k == 0) {
      uy = gr[1][i]; uz = gr[2][i]; vx = gr[3][i]; vz = gr[5][i];
      wx = 0.5 * ((double)gr[6][i] + (double)gr[6][i + pl]);
      wy = 0.5 * ((double)gr[7][i] + (double)gr[7][i + pl]);
    } else {
      uy = 0.5 * ((double)gr[1][i] + (double)gr[1][i - pl]);
      vx = 0.5 * ((double)gr[3][i] + (double)gr[3][i - pl]);
      uz = gr[2][i]; vz = gr[5][i]; wx = gr[6][i]; wy = gr[7][i];
    }
    const double o12 = 0.5 * (uy - vx), o13 = 0.5 * (uz - wx), o23 = 0.5 * (vz - wy);
    // contractions in double, the compressing transforms in float (the
    // network sees the inputs rounded to bf16)
    f[0] = asinhf(sqrtf((float)(2.0 * dot(s, s))));
    f[1] = asinhf(sqrtf((float)(2.0 * dot(sb, sb))));
    f[2] = asinhf(sqrtf((float)(2.0 * dot(sh, sh))));
    f[3] = asinhf((float)dot(s, sb));
    f[4] = asinhf((float)dot(sb, ssb));
    f[5] = asinhf((float)dot(sh, ssh));
    f[6] = asinhf(2.0f * sqrtf((float)(o12 * o12 + o13 * o13 + o23 * o23)));
    f[7] = (k == 0 ? 0.5f * dz : (float)k * dz) * idelta;
  }
};

template <typename T>
__global__ void k_mlp_harvest(FeatLag<T> feat, const T* __restrict__ Cs, const T* __restrict__ Beta, long P,
                              float* __restrict__ X, float* __restrict__ Y) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
    float f[MI];
    feat(p, f);
#pragma unroll
    for (int e = 0; e < MI; ++e) X[p * MI + e] = f[e];
    if (Y) {
      const long i = feat.index(p);
      Y[p * 2] = (float)Cs[i];
      Y[p * 2 + 1] = (float)Beta[i];
    }
  }
}

// closure output: Cs_opt2 = max(y0, 0) and Beta on the field, only while the
// last dynamic step judged the network good enough (flag[0] != 0)
template <typename T>
struct OutCsBeta {
  T *Cs, *Beta;
  FeatLag<T> at;
  const double* flag;
  __device__ void operator()(long p, float y0, float y1) const {
    if (flag && flag[0] == 0.0) return;
    const long i = at.index(p);
    Cs[i] = (T)fmaxf(y0, 0.0f);
    Beta[i] = (T)y1;
  }
};
// the judge: network prediction (y0) against the dynamic Cs_opt2 at p
template <typename T>
struct OutJudge {
  const T* Cs;
  FeatLag<T> at;
  float* pred;  // (P) the network's Cs_opt2
  __device__ void operator()(long p, float y0, float) const { pred[p] = fmaxf(y0, 0.0f); }
};

template <typename T>
__global__ void k_mlp_err_partial(const float* __restrict__ pred, FeatLag<T> at, const T* __restrict__ Cs, long P,
                                  double* part) {
  double a = 0.0, b = 0.0;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
    const double d = (double)Cs[at.index(p)];
    const double e = (double)pred[p] - d;
    a += e * e;
    b += d * d;
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2] = sh[0][0];
    part[blockIdx.x * 2 + 1] = sh[1][0];
  }
}
// judge[0..3] = sum (pred - Cs)^2, sum Cs^2, relative L2 error, use_mlp (err <= thresh)
__global__ void k_mlp_err_final(const double* part, int nparts, double thresh, double* judge) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < nparts; ++i) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  const double err = b > 0.0 ? sqrt(a / b) : 1.0;
  judge[0] = a;
  judge[1] = b;
  judge[2] = err;
  judge[3] = err <= thresh ? 1.0 : 0.0;
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
static MlpModel* model(les_ctx* c) { return (MlpModel*)c->mlp; }

static void mlp_pack(les_ctx* c) {
  MlpModel* M = model(c);
  hipLaunchKernelGGL(k_mlp_pack, dim3((NOPS * 64 + 255) / 256), dim3(256), 0, c->stream, (const float*)M->theta,
                     M->ops);
  LCHK(c, hipGetLastError());
}

int mlp_nparam() { return MLP_NP; }

void mlp_free(les_ctx* c) {
  MlpModel* M = model(c);
  if (!M) return;
  for (void* p : {(void*)M->theta, (void*)M->m, (void*)M->v, (void*)M->ops, (void*)M->part, (void*)M->lpart,
                  (void*)M->grad, (void*)M->judge})
    if (p) (void)hipFree(p);
  delete M;
  c->mlp = nullptr;
}

void mlp_init(les_ctx* c, uint64_t seed) {
  mlp_free(c);
  MlpModel* M = new MlpModel();
  c->mlp = M;
  M->nwaves = 256 * 4;  // k_mlp_grad: 256 blocks of 4 waves
  check(c, hipMalloc(&M->theta, MLP_NP * sizeof(float)), "mlp theta");
  check(c, hipMalloc(&M->m, MLP_NP * sizeof(float)), "mlp m");
  check(c, hipMalloc(&M->v, MLP_NP * sizeof(float)), "mlp v");
  check(c, hipMalloc(&M->grad, MLP_NP * sizeof(float)), "mlp grad");
  check(c, hipMalloc(&M->ops, NOPS * 64 * sizeof(uint4)), "mlp ops");
  check(c, hipMalloc(&M->part, (size_t)M->nwaves * MLP_NP * sizeof(float)), "mlp part");
  check(c, hipMalloc(&M->lpart, (size_t)M->nwaves * sizeof(float)), "mlp lpart");
  check(c, hipMalloc(&M->judge, 4 * sizeof(double)), "mlp judge");
  if (!c->err.empty()) return;
  // He-style initialisation from a counter-based generator (reproducible)
  std::vector<float> th(MLP_NP, 0.0f);
  auto U = [&](uint64_t k) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + k + 1;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int n, int fan_in) {
    const double a = sqrt(6.0 / fan_in);
    for (int i = 0; i < n; ++i) th[off + i] = (float)(a * (2.0 * U(off + i) - 1.0));
  };
  fill(OW1, MH * MI, MI);
  fill(OW2, MH * MH, MH);
  fill(OW3, MH * MH, MH);
  fill(OW4, MO * MH, MH);
  for (int i = 0; i < MO * MH; ++i) th[OW4 + i] *= 0.1f;
  check(c, hipMemcpyAsync(M->theta, th.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, c->stream), "theta");
  std::vector<float> z(MLP_NP, 0.0f);
  check(c, hipMemcpyAsync(M->m, z.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, c->stream), "m");
  check(c, hipMemcpyAsync(M->v, z.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, c->stream), "v");
  const double j0[4] = {0.0, 0.0, 1.0, 0.0};
  check(c, hipMemcpyAsync(M->judge, j0, sizeof(j0), hipMemcpyHostToDevice, c->stream), "judge");
  check(c, hipStreamSynchronize(c->stream), "mlp init sync");
  for (int i = 0; i < MI; ++i) { M->norm.mu[i] = 0.0f; M->norm.isd[i] = 1.0f; }
  for (int i = 0; i < MO; ++i) { M->norm.ymu[i] = 0.0f; M->norm.ysd[i] = 1.0f; }
  M->adam_t = 0;
  mlp_pack(c);
}

void mlp_set_params(les_ctx* c, const float* host) {
  MlpModel* M = model(c);
  check(c, hipMemcpyAsync(M->theta, host, MLP_NP * sizeof(float), hipMemcpyHostToDevice, c->stream), "set theta");
  check(c, hipStreamSynchronize(c->stream), "set theta sync");
  mlp_pack(c);
}
void mlp_get_params(les_ctx* c, float* host) {
  MlpModel* M = model(c);
  check(c, hipMemcpyAsync(host, M->theta, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, c->stream), "get theta");
  check(c, hipStreamSynchronize(c->stream), "get theta sync");
}
void mlp_set_norm(les_ctx* c, const float* mu, const float* sd, const float* ymu, const float* ysd) {
  MlpNorm& n = model(c)->norm;
  for (int i = 0; i < MI; ++i) { n.mu[i] = mu[i]; n.isd[i] = 1.0f / sd[i]; }
  for (int i = 0; i < MO; ++i) { n.ymu[i] = ymu[i]; n.ysd[i] = ysd[i]; }
}
void mlp_get_norm(les_ctx* c, float* mu, float* sd, float* ymu, float* ysd) {
  const MlpNorm& n = model(c)->norm;
  for (int i = 0; i < MI; ++i) { mu[i] = n.mu[i]; sd[i] = 1.0f / n.isd[i]; }
  for (int i = 0; i < MO; ++i) { ymu[i] = n.ymu[i]; ysd[i] = n.ysd[i]; }
}

// mean and standard deviation of the columns of X (P x 8) and Y (P x 2)
void mlp_fit_norm(les_ctx* c, const float* X, const float* Y, long P) {
  constexpr int NB = 64;
  double* part = (double*)scratch(c, S_RED, (size_t)NB * (MI + MO) * 2 * sizeof(double));
  hipLaunchKernelGGL(k_mlp_moments, dim3(NB, MI), dim3(256), 0, c->stream, X, P, MI, part);
  hipLaunchKernelGGL(k_mlp_moments, dim3(NB, MO), dim3(256), 0, c->stream, Y, P, MO, part + NB * MI * 2);
  LCHK(c, hipGetLastError());
  std::vector<double> h((size_t)NB * (MI + MO) * 2);
  check(c, hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream), "moments");
  check(c, hipStreamSynchronize(c->stream), "moments sync");
  float mu[MI + MO], sd[MI + MO];
  for (int col = 0; col < MI + MO; ++col) {
    double s = 0.0, s2 = 0.0;
    for (int b = 0; b < NB; ++b) {
      s += h[((size_t)col * NB + b) * 2];
      s2 += h[((size_t)col * NB + b) * 2 + 1];
    }
    const double m = s / (double)P, var = s2 / (double)P - m * m;
    mu[col] = (float)m;
    sd[col] = (float)(var > 1e-20 ? sqrt(var) : 1.0);
  }
  mlp_set_norm(c, mu, sd, mu + MI, sd + MI);
}

void op_mlp_forward(les_ctx* c, const float* X, long P, float* Y) {
  OpTag _tag(c, "mlp");
  MlpModel* M = model(c);
  const long groups = (P + 63) / 64;
  const unsigned blocks = (unsigned)((groups + 3) / 4 < 2048 ? (groups + 3) / 4 : 2048);
  ProfScope _ps(c, "mlp_infer", (double)P * (MI + MO) * 4);
  hipLaunchKernelGGL((k_mlp_infer<FeatArray, OutArray>), dim3(blocks > 0 ? blocks : 1), dim3(256), 0, c->stream,
                     FeatArray{X}, OutArray{Y}, P, (const uint4*)M->ops, (const float*)M->theta, M->norm);
  LCHK(c, hipGetLastError());
}

// gradient of the batch loss (normalised units) into M->grad, loss into *loss
static void mlp_grad_step(les_ctx* c, const float* X, const float* T, long P, long nb, int full, uint64_t seed,
                          long step, float* loss, int adam, float lr) {
  MlpModel* M = model(c);
  hipLaunchKernelGGL(k_mlp_grad, dim3(M->nwaves / 4), dim3(256), 0, c->stream, X, T, P, nb, full, seed, step,
                     (const uint4*)M->ops, (const float*)M->theta, M->norm, M->part, M->lpart);
  LCHK(c, hipGetLastError());
  float b1c = 1.0f, b2c = 1.0f;
  if (adam) {
    M->adam_t += 1;
    b1c = (float)(1.0 - pow(0.9, (double)M->adam_t));
    b2c = (float)(1.0 - pow(0.999, (double)M->adam_t));
  }
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, c->stream,
                     (const float*)M->part, (const float*)M->lpart, M->nwaves, M->grad, loss, M->theta, M->m, M->v,
                     adam, lr, b1c, b2c);
  LCHK(c, hipGetLastError());
  if (adam) mlp_pack(c);
}

void op_mlp_grad(les_ctx* c, const float* X, const float* T, long P, float* grad, float* loss) {
  OpTag _tag(c, "mlp");
  mlp_grad_step(c, X, T, P, P, 1, 0, 0, loss, 0, 0.0f);
  if (grad)
    check(c, hipMemcpyAsync(grad, model(c)->grad, MLP_NP * sizeof(float), hipMemcpyDeviceToDevice, c->stream),
          "grad copy");
}

void op_mlp_train(les_ctx* c, const float* X, const float* T, long P, int batch, int steps, float lr,
                  uint64_t seed, float* loss_hist) {
  OpTag _tag(c, "mlp_train");
  for (int s = 0; s < steps; ++s)
    mlp_grad_step(c, X, T, P, batch, 0, seed, s, loss_hist ? loss_hist + s : nullptr, 1, lr);
}

template <typename T>
static FeatLag<T> feat_lag(les_ctx* c, const T* F, const T* const S[6], const T* const gr[9]) {
  const Grid& g = c->g;
  FeatLag<T> f;
  const long n3 = (long)g.ld * g.ny * g.nz;
  f.F2 = F;
  f.F4 = F + 21 * n3;
  for (int q = 0; q < 6; ++q) f.S[q] = S[q];
  for (int q = 0; q < 9; ++q) f.gr[q] = gr[q];
  f.n3 = n3;
  f.pl = (long)g.ld * g.ny;
  f.nx = g.nx;
  f.ny = g.ny;
  f.ld = g.ld;
  f.dz = (float)g.dz;
  f.idelta = (float)(1.0 / g.delta);
  return f;
}

// features (and labels) of the last dynamic step: F = its filtered fields,
// S its strain rates, gr the gradients it used
template <typename T>
void op_mlp_harvest(les_ctx* c, const T* F, const T* const S[6], const T* const gr[9], const T* Cs, const T* Beta,
                    float* X, float* Y) {
  const Grid& g = c->g;
  const long P = (long)g.nx * g.ny * (g.nz - 1);
  hipLaunchKernelGGL((k_mlp_harvest<T>), grid1((size_t)P), dim3(256), 0, c->stream, feat_lag<T>(c, F, S, gr), Cs,
                     Beta, P, X, Y);
  LCHK(c, hipGetLastError());
}

// closure steps: judge = the network against the dynamic Cs_opt2 just computed
// (sets the device flag); apply = the network's Cs_opt2/Beta on the field
// where the flag allows it
template <typename T>
void op_mlp_closure(les_ctx* c, const T* F, const T* const S[6], const T* const gr[9], T* Cs, T* Beta, int judge,
                    double thresh) {
  OpTag _tag(c, judge ? "mlp_judge" : "mlp_apply");
  MlpModel* M = model(c);
  const Grid& g = c->g;
  const long P = (long)g.nx * g.ny * (g.nz - 1);
  const FeatLag<T> fl = feat_lag<T>(c, F, S, gr);
  const long groups = (P + 63) / 64;
  const unsigned blocks = (unsigned)((groups + 3) / 4 < 2048 ? (groups + 3) / 4 : 2048);
  if (judge) {
    float* pred = (float*)scratch(c, S_MLPP, (size_t)P * sizeof(float));
    {
      ProfScope _ps(c, "mlp_infer", (double)P * (33.0 * sizeof(T) + 4));
      hipLaunchKernelGGL((k_mlp_infer<FeatLag<T>, OutJudge<T>>), dim3(blocks), dim3(256), 0, c->stream, fl,
                         OutJudge<T>{Cs, fl, pred}, P, (const uint4*)M->ops, (const float*)M->theta, M->norm);
    }
    LCHK(c, hipGetLastError());
    constexpr int NB = 256;
    double* part = (double*)scratch(c, S_RED, NB * 2 * sizeof(double));
    hipLaunchKernelGGL((k_mlp_err_partial<T>), dim3(NB), dim3(256), 0, c->stream, (const float*)pred, fl,
                       (const T*)Cs, P, part);
    hipLaunchKernelGGL(k_mlp_err_final, dim3(1), dim3(64), 0, c->stream, (const double*)part, NB, thresh,
                       M->judge);
    LCHK(c, hipGetLastError());
  } else {
    ProfScope _ps(c, "mlp_infer", (double)P * (33.0 * sizeof(T) + 2.0 * sizeof(T)));
    hipLaunchKernelGGL((k_mlp_infer<FeatLag<T>, OutCsBeta<T>>), dim3(blocks), dim3(256), 0, c->stream, fl,
                       OutCsBeta<T>{Cs, Beta, fl, (const double*)(M->judge + 3)}, P, (const uint4*)M->ops,
                       (const float*)M->theta, M->norm);
    LCHK(c, hipGetLastError());
  }
}

void mlp_judge_read(les_ctx* c, double out[4]) {
  check(c, hipMemcpyAsync(out, model(c)->judge, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "judge");
  check(c, hipStreamSynchronize(c->stream), "judge sync");
}

#define LES_MLP_INST(T)                                                                                      \
  template void op_mlp_harvest<T>(les_ctx*, const T*, const T* const[6], const T* const[9], const T*, const T*, \
                                  float*, float*);                                                            \
  template void op_mlp_closure<T>(les_ctx*, const T*, const T* const[6], const T* const[9], T*, T*, int, double);
LES_MLP_INST(double)
LES_MLP_INST(float)
#undef LES_MLP_INST

}  // namespace les
