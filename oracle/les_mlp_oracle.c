/* les_mlp_oracle.c -- CPU reference of the learned SGS closure (TEST
 * INFRASTRUCTURE ONLY: used by tests/ to check the GPU kernels; never part of
 * the product path).  The closure has no counterpart in the reference code.
 *
 * Network: 8 -> 64 -> 64 -> 64 -> 2, SiLU on the hidden layers, parameters
 * flat in the order W1(64x8) b1(64) W2(64x64) b2 W3(64x64) b3 W4(2x64) b4(2),
 * matrices row-major [out][in].  Loss over a batch of P samples:
 *   L = 1/(2P) sum_p sum_o (y_po - t_po)^2.
 *
 * With bf16 != 0 the products see the operands the GPU's bf16 MFMA sees: the
 * inputs, weights and activations of each layer rounded to bf16 (through
 * float, round-to-nearest-even), the products accumulated exactly (double
 * here, FP32 on the GPU), biases and the SiLU in full precision.  In the
 * backward pass dy, dz and the stored activations are rounded the same way,
 * and the bias gradients are sums of the rounded dz.  With bf16 == 0 the
 * whole computation is plain FP64, which the finite-difference pin checks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "les_oracle.h"

enum { NI = OR_MLP_NIN, NH = OR_MLP_NH, NO = OR_MLP_NOUT };
enum {
  OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
  OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NO * NH, NP = OB4 + NO
};

/* round to bf16 through float, round-to-nearest-even (finite inputs) */
static double rb(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}
static double q(double x, int bf16) { return bf16 ? rb(x) : x; }
static double sig(double z) { return 1.0 / (1.0 + exp(-z)); }
static double silu(double z) { return z * sig(z); }
static double dsilu(double z) { double s = sig(z); return s * (1.0 + z * (1.0 - s)); }

int or_mlp_nparam(void) { return NP; }

/* one sample: z1..z3 pre-activations, y outputs */
static void fwd1(const double *th, const double *x, int bf16, double z[3][NH], double y[NO]) {
  double in[NH], h[NH];
  for (int i = 0; i < NI; i++) in[i] = q(x[i], bf16);
  for (int o = 0; o < NH; o++) {
    double a = 0.0;
    for (int i = 0; i < NI; i++) a += q(th[OW1 + o * NI + i], bf16) * in[i];
    z[0][o] = a + th[OB1 + o];
  }
  const int ow[2] = { OW2, OW3 }, ob[2] = { OB2, OB3 };
  for (int l = 0; l < 2; l++) {
    for (int i = 0; i < NH; i++) h[i] = q(silu(z[l][i]), bf16);
    for (int o = 0; o < NH; o++) {
      double a = 0.0;
      for (int i = 0; i < NH; i++) a += q(th[ow[l] + o * NH + i], bf16) * h[i];
      z[l + 1][o] = a + th[ob[l] + o];
    }
  }
  for (int i = 0; i < NH; i++) h[i] = q(silu(z[2][i]), bf16);
  for (int o = 0; o < NO; o++) {
    double a = 0.0;
    for (int i = 0; i < NH; i++) a += q(th[OW4 + o * NH + i], bf16) * h[i];
    y[o] = a + th[OB4 + o];
  }
}

void or_mlp_forward(const double *theta, const double *x, long P, double *y, int bf16) {
  double z[3][NH];
  for (long p = 0; p < P; p++) fwd1(theta, x + NI * p, bf16, z, y + NO * p);
}

double or_mlp_loss_grad(const double *th, const double *x, const double *t, long P, double *grad, int bf16) {
  memset(grad, 0, sizeof(double) * NP);
  double loss = 0.0;
  for (long p = 0; p < P; p++) {
    double z[3][NH], y[NO], dy[NO], h[4][NH], dz[NH], dh[NH];
    fwd1(th, x + NI * p, bf16, z, y);
    for (int i = 0; i < NI; i++) h[0][i] = q(x[NI * p + i], bf16);
    for (int l = 0; l < 3; l++)
      for (int i = 0; i < NH; i++) h[l + 1][i] = q(silu(z[l][i]), bf16);
    for (int o = 0; o < NO; o++) {
      double e = y[o] - t[NO * p + o];
      loss += 0.5 * e * e / (double)P;
      dy[o] = q(e / (double)P, bf16);
    }
    /* output layer */
    for (int o = 0; o < NO; o++) {
      grad[OB4 + o] += dy[o];
      for (int i = 0; i < NH; i++) grad[OW4 + o * NH + i] += dy[o] * h[3][i];
    }
    for (int i = 0; i < NH; i++) {
      double a = 0.0;
      for (int o = 0; o < NO; o++) a += q(th[OW4 + o * NH + i], bf16) * dy[o];
      dh[i] = a;
    }
    /* hidden layers 3, 2 (weights W3, W2), then layer 1 (W1) */
    const int ow[3] = { OW1, OW2, OW3 }, ob[3] = { OB1, OB2, OB3 };
    for (int l = 2; l >= 0; l--) {
      const int nin = l == 0 ? NI : NH;
      for (int o = 0; o < NH; o++) dz[o] = q(dh[o] * dsilu(z[l][o]), bf16);
      for (int o = 0; o < NH; o++) {
        grad[ob[l] + o] += dz[o];
        for (int i = 0; i < nin; i++) grad[ow[l] + o * nin + i] += dz[o] * h[l][i];
      }
      if (l > 0)
        for (int i = 0; i < NH; i++) {
          double a = 0.0;
          for (int o = 0; o < NH; o++) a += q(th[ow[l] + o * NH + i], bf16) * dz[o];
          dh[i] = a;
        }
    }
  }
  return loss;
}
