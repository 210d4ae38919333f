/*
 * mlp_oracle.c -- CPU FP64 oracle of the pair-screening MLP (16 -> 64 -> 64 -> 1,
 * GELU erf form).  TEST INFRASTRUCTURE ONLY (see rimp2_oracle.c).
 *
 * Nothing in the reference corresponds to this network; it is pinned by
 * finite differences (FP64 mode), the closed form of GELU, and bf16
 * round-to-nearest-even on known values (tests/test_oracle.py).
 *
 * Parameter vector (float32, as on the GPU):
 *   W1[16][64] | b1[64] | W2[64][64] (W2[k][m]: input k -> output m) | b2[64] | w3[64] | b3
 * norm: mean[16], std[16], tmean, tstd.  Inputs are standardised,
 * x = (feat - mean)/std;  the network predicts the standardised residual
 * (target - feat[15] - tmean)/tstd where the target is log(-eps_ij) and
 * feature 15 is the base estimate; prediction = feat[15] + y tstd + tmean.
 * loss = mean over samples of (y - t_std)^2.
 *
 * emulate_bf16 != 0 rounds to bf16 at the points the GPU kernel feeds its
 * MFMAs: x, W1, W2, w3, h1 = GELU(z1), h2 = GELU(z2), dZ2 and dZ1 (the
 * gradient operands of the W2 / W1 products); everything else stays FP64.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NF 16
#define NH 64
#define P_W1 0
#define P_B1 (P_W1 + NF * NH)
#define P_W2 (P_B1 + NH)
#define P_B2 (P_W2 + NH * NH)
#define P_W3 (P_B2 + NH)
#define P_B3 (P_W3 + NH)
#define NPARAM (P_B3 + 1)

int64_t oracle_mlp_num_params(void) { return NPARAM; }

/* bf16 round-to-nearest-even of a float */
float oracle_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

double oracle_gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
double oracle_gelu_d(double x) { return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / sqrt(2.0 * 3.14159265358979323846); }

static inline double rb(double x, int em) { return em ? (double)oracle_bf16((float)x) : x; }

/* Returns the mean loss; grad (NPARAM doubles) and pred (n doubles, raw units)
 * may be NULL. */
double oracle_mlp_loss_grad(int64_t n, const float* feats, const float* target, const float* params,
                            const float* norm, int emulate_bf16, double* grad, double* pred) {
  const int em = emulate_bf16;
  double W1[NF][NH], W2[NH][NH], w3[NH], b1[NH], b2[NH], b3;
  for (int f = 0; f < NF; f++)
    for (int k = 0; k < NH; k++) W1[f][k] = rb(params[P_W1 + f * NH + k], em);
  for (int k = 0; k < NH; k++)
    for (int m = 0; m < NH; m++) W2[k][m] = rb(params[P_W2 + k * NH + m], em);
  for (int k = 0; k < NH; k++) {
    w3[k] = rb(params[P_W3 + k], em);
    b1[k] = params[P_B1 + k];
    b2[k] = params[P_B2 + k];
  }
  b3 = params[P_B3];
  if (grad) memset(grad, 0, sizeof(double) * NPARAM);
  double loss = 0.0;
  for (int64_t s = 0; s < n; s++) {
    double x[NF], z1[NH], h1[NH], z2[NH], h2[NH];
    for (int f = 0; f < NF; f++) {
      double z = ((double)feats[s * NF + f] - norm[f]) / norm[NF + f];
      z = z > 4.0 ? 4.0 : (z < -4.0 ? -4.0 : z); /* inputs saturate at +-4 */
      x[f] = rb(z, em);
    }
    for (int k = 0; k < NH; k++) {
      double a = 0.0;
      for (int f = 0; f < NF; f++) a += x[f] * W1[f][k];
      z1[k] = a + b1[k];
      h1[k] = rb(oracle_gelu(z1[k]), em);
    }
    for (int m = 0; m < NH; m++) {
      double a = 0.0;
      for (int k = 0; k < NH; k++) a += h1[k] * W2[k][m];
      z2[m] = a + b2[m];
      h2[m] = rb(oracle_gelu(z2[m]), em);
    }
    double y = b3;
    for (int k = 0; k < NH; k++) y += h2[k] * w3[k];
    const double base = (double)feats[s * NF + (NF - 1)]; /* residual model: base estimate = feature 15 */
    if (pred) pred[s] = base + y * norm[2 * NF + 1] + norm[2 * NF];
    if (!target) continue;
    const double d = y - ((double)target[s] - base - norm[2 * NF]) / norm[2 * NF + 1];
    loss += d * d;
    if (!grad) continue;
    const double dy = 2.0 * d / (double)n;
    double dz2[NH], d2[NH], dz1[NH], d1[NH];
    for (int k = 0; k < NH; k++) grad[P_W3 + k] += h2[k] * dy;
    grad[P_B3] += dy;
    for (int m = 0; m < NH; m++) {
      dz2[m] = dy * w3[m] * oracle_gelu_d(z2[m]);
      d2[m] = rb(dz2[m], em);
      grad[P_B2 + m] += dz2[m];
    }
    for (int k = 0; k < NH; k++)
      for (int m = 0; m < NH; m++) grad[P_W2 + k * NH + m] += h1[k] * d2[m];
    for (int k = 0; k < NH; k++) {
      double a = 0.0;
      for (int m = 0; m < NH; m++) a += d2[m] * W2[k][m];
      dz1[k] = a * oracle_gelu_d(z1[k]);
      d1[k] = rb(dz1[k], em);
      grad[P_B1 + k] += dz1[k];
    }
    for (int f = 0; f < NF; f++)
      for (int k = 0; k < NH; k++) grad[P_W1 + f * NH + k] += x[f] * d1[k];
  }
  return target ? loss / (double)n : 0.0;
}
