// Pair-screening surrogate: per-pair features, a BF16 MLP 16 -> 64 -> 64 -> 1
// (GELU, erf form) with hand-written forward / backward / Adam kernels, and the
// screened energy (pairs predicted below a threshold skip the exact kernel).
//
// Nothing here exists in the reference; the exact pair kernel has the last
// word: screened energies are always the exact kernel's energies on the kept
// pairs.
//
// Features of a unique pair (i <= j), from the FP64 B of the last transform
// (r_i[Q] = ||B[Q,:,i]||, the row norm of occupied i at auxiliary index Q;
// t = sum_Q r_i[Q] r_j[Q] >= ||V^ij||_F (Cauchy-Schwarz per Q),
// s2 = sum_Q r_i[Q]^2 r_j[Q]^2, which ||V^ij||_F^2 approaches when the
// auxiliary contributions add incoherently; gbar = 2 mean(e_a) - e_i - e_j):
//   0 log(||B_i|| ||B_j|| / t)     1 log(||B_i|| / ||B_j||)
//   2..9  log(f_s / t), f_s = sum_{Q in slab s} r_i[Q] r_j[Q]  (8 equal Q-slabs)
//   10 log(t^2 / s2)  (coherence: 0 .. log naux)
//   11 e_i + e_j     12 2 min(e_a) - e_i - e_j     13 gbar
//   14 [i == j]      15 log(s2 / gbar)  (the base estimate of -eps_ij)
// All but the energies and the base estimate are dimensionless, so a network
// trained on one fragment shape applies to others.  The network predicts the
// residual log(-eps_ij) - feature 15: prediction = feature 15 + output, so an
// untrained or out-of-range network falls back to the base estimate.
// Features and the residual target are standardised with the statistics set
// by rimp2_mlp_set_norm (written into the MLP state); standardised features
// saturate at +-4.
//
// Network: z1 = x W1 + b1, h1 = GELU(z1), z2 = h1 W2 + b2, h2 = GELU(z2),
// y = h2 w3 + b3.  Weights and activations feed the MFMAs as BF16
// (v_mfma_f32_32x32x16_bf16 / 16x16x32_bf16), accumulation is FP32, the master
// weights and Adam moments are FP32.  One workgroup handles 64 samples and
// writes its gradient partial; a fixed-order reduction and the Adam update
// follow (deterministic).
#include <cmath>
#include <vector>

#include "rimp2_common.h"

namespace rimp2 {

constexpr int NF = RIMP2_NFEAT;  // 16
constexpr int NH = 64;
// parameter vector layout (float32)
constexpr int P_W1 = 0;                  // [NF][NH]
constexpr int P_B1 = P_W1 + NF * NH;     // [NH]
constexpr int P_W2 = P_B1 + NH;          // [NH][NH]  (W2[k][m]: input k -> output m)
constexpr int P_B2 = P_W2 + NH * NH;     // [NH]
constexpr int P_W3 = P_B2 + NH;          // [NH]
constexpr int P_B3 = P_W3 + NH;          // [1]
constexpr int NPARAM = P_B3 + 1;         // 5313
// standardised inputs saturate here: a fragment whose features leave the
// training range cannot drive the network into wild extrapolation
constexpr float XCLIP = 4.0f;

struct MlpState {
  bool init = false;
  int64_t step = 0;
  DevBuf params, m, v, grad, partial, feats, target, pred, loss, norm, rowcap, pairs_keep, pair_list;
  float norm_h[2 * NF + 2];
  bool have_norm = false;
  ~MlpState() {
    for (DevBuf* b : {&params, &m, &v, &grad, &partial, &feats, &target, &pred, &loss, &norm, &rowcap, &pairs_keep,
                      &pair_list})
      b->release();
  }
};

void mlp_destroy(MlpState* m) { delete m; }

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// bf16 round-to-nearest-even
__device__ inline __bf16 tobf(float x) { return (__bf16)x; }

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
// r[i*naux + Q] = ||B[Q,:,i]||_2 ; one thread per (Q, i)
__global__ void rownorm_kernel(const double* B, int64_t naux, int64_t nvirt, int64_t nocc, float* r) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= naux * nocc) return;
  const int64_t q = t % naux, i = t / naux;
  double s = 0.0;
  for (int64_t a = 0; a < nvirt; a++) {
    const double x = B[q + naux * (a + nvirt * i)];
    s += x * x;
  }
  r[i * naux + q] = (float)sqrt(s);
}

// one workgroup per unique pair; 8 slabs; raw (unstandardised) features
__global__ __launch_bounds__(256) void pair_features_kernel(const float* r, int64_t naux, int64_t /*nocc*/,
                                                            const double* eo, const double* ev, int64_t nvirt,
                                                            float* F) {
  const int64_t p = blockIdx.x;
  int64_t j = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (j * (j + 1) / 2 > p) j--;
  while ((j + 1) * (j + 2) / 2 <= p) j++;
  const int64_t i = p - j * (j + 1) / 2;
  __shared__ double sl[8][256];
  __shared__ double ni[256], nj[256], nc[256], evs[2][256];
  double f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double a2 = 0, b2 = 0, c2 = 0;
  for (int64_t q = threadIdx.x; q < naux; q += 256) {
    const double x = r[i * naux + q], y = r[j * naux + q];
    f[(q * 8) / naux] += x * y;
    a2 += x * x;
    b2 += y * y;
    c2 += (x * y) * (x * y);
  }
  double emin = INFINITY, esum = 0.0;
  for (int64_t a = threadIdx.x; a < nvirt; a += 256) {
    emin = fmin(emin, ev[a]);
    esum += ev[a];
  }
  for (int s = 0; s < 8; s++) sl[s][threadIdx.x] = f[s];
  ni[threadIdx.x] = a2;
  nj[threadIdx.x] = b2;
  nc[threadIdx.x] = c2;
  evs[0][threadIdx.x] = emin;
  evs[1][threadIdx.x] = esum;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int s = 0; s < 8; s++) sl[s][threadIdx.x] += sl[s][threadIdx.x + w];
      ni[threadIdx.x] += ni[threadIdx.x + w];
      nj[threadIdx.x] += nj[threadIdx.x + w];
      nc[threadIdx.x] += nc[threadIdx.x + w];
      evs[0][threadIdx.x] = fmin(evs[0][threadIdx.x], evs[0][threadIdx.x + w]);
      evs[1][threadIdx.x] += evs[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double tiny = 1e-300;
    float* o = F + p * NF;
    double tot = 0.0;
    for (int s = 0; s < 8; s++) tot += sl[s][0];
    const double ltot = log(tot + tiny), lni = 0.5 * log(ni[0] + tiny), lnj = 0.5 * log(nj[0] + tiny);
    const double ls2 = log(nc[0] + tiny);
    o[0] = (float)(lni + lnj - ltot);
    o[1] = (float)(lni - lnj);
    for (int s = 0; s < 8; s++) o[2 + s] = (float)(log(sl[s][0] + tiny) - ltot);
    o[10] = (float)(2.0 * ltot - ls2);
    const double eij = eo[i] + eo[j];
    const double gmin = 2.0 * evs[0][0] - eij, gmean = 2.0 * evs[1][0] / (double)nvirt - eij;
    o[11] = (float)eij;
    o[12] = (float)gmin;
    o[13] = (float)gmean;
    o[14] = (i == j) ? 1.0f : 0.0f;
    o[15] = (float)(ls2 - log(gmean));
  }
}

// ---------------------------------------------------------------------------
// MLP forward (+ backward) on 64 samples per workgroup
// ---------------------------------------------------------------------------
struct MlpArgs {
  const float* feats;   // [n][NF] raw
  const float* target;  // [n] raw (log(-eps)) or null
  const float* params;  // [NPARAM] fp32 master
  const float* norm;    // [2*NF + 2]: mean[NF], std[NF], tmean, tstd
  float* pred;          // [n] raw prediction or null
  float* partial;       // [nwg][NPARAM] gradient partials (train) or null
  float* loss_part;     // [nwg] sum of squared errors (normalised units) or null
  int64_t n;
  float inv_n;          // 1 / n for the mean loss
};

template <bool TRAIN>
__global__ __launch_bounds__(256) void mlp_kernel(MlpArgs a) {
  __shared__ __align__(16) __bf16 Xs[64][NF];     // samples x features
  __shared__ __align__(16) __bf16 XT[NF][64];     // features x samples
  __shared__ __align__(16) __bf16 W1T[NH][NF];    // unit x feature
  __shared__ __align__(16) __bf16 W2T[NH][NH];    // out m x in k
  __shared__ __align__(16) __bf16 W2R[NH][NH];    // in k x out m
  __shared__ __align__(16) __bf16 H1[64][NH];     // sample x unit
  __shared__ __align__(16) __bf16 H1T[NH][64];    // unit x sample
  __shared__ float Z1[64][NH + 1];
  __shared__ float Z2[64][NH + 1];
  __shared__ float H2[64][NH + 1];
  __shared__ __align__(16) __bf16 D2[64][NH];     // dZ2 sample x unit
  __shared__ __align__(16) __bf16 D2T[NH][64];    // dZ2 unit x sample
  __shared__ __align__(16) __bf16 D1T[NH][64];    // dZ1 unit x sample
  __shared__ float dy[64], yv[64];
  __shared__ float D1f[64][NH + 1];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t s0 = (int64_t)blockIdx.x * 64;
  const float* P = a.params;

  // load + standardise the samples (padding samples are zero and masked)
  for (int e = tid; e < 64 * NF; e += 256) {
    const int s = e / NF, f = e % NF;
    float x = 0.f;
    if (s0 + s < a.n) x = fminf(XCLIP, fmaxf(-XCLIP, (a.feats[(s0 + s) * NF + f] - a.norm[f]) / a.norm[NF + f]));
    Xs[s][f] = tobf(x);
    XT[f][s] = tobf(x);
  }
  for (int e = tid; e < NF * NH; e += 256) {
    const int f = e / NH, k = e % NH;
    W1T[k][f] = tobf(P[P_W1 + f * NH + k]);
  }
  for (int e = tid; e < NH * NH; e += 256) {
    const int k = e / NH, m = e % NH;
    const __bf16 w = tobf(P[P_W2 + k * NH + m]);
    W2T[m][k] = w;
    W2R[k][m] = w;
  }
  __syncthreads();

  const int rb = (wave >> 1) * 32, cb = (wave & 1) * 32;  // this wave's 32x32 block
  // ---- layer 1: Z1 = X W1 + b1 (K = 16: one MFMA) ----
  {
    v16f acc = {};
    const v8bf A = *reinterpret_cast<const v8bf*>(&Xs[rb + (lane & 31)][8 * (lane >> 5)]);
    const v8bf Bv = *reinterpret_cast<const v8bf*>(&W1T[cb + (lane & 31)][8 * (lane >> 5)]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bv, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int row = rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), col = cb + (lane & 31);
      const float z = acc[r] + P[P_B1 + col];
      Z1[row][col] = z;
      const __bf16 h = tobf(gelu(z));
      H1[row][col] = h;
      H1T[col][row] = h;
    }
  }
  __syncthreads();
  // ---- layer 2: Z2 = H1 W2 + b2 (K = 64: four MFMAs) ----
  {
    v16f acc = {};
#pragma unroll
    for (int k0 = 0; k0 < NH; k0 += 16) {
      const v8bf A = *reinterpret_cast<const v8bf*>(&H1[rb + (lane & 31)][k0 + 8 * (lane >> 5)]);
      const v8bf Bv = *reinterpret_cast<const v8bf*>(&W2T[cb + (lane & 31)][k0 + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int row = rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), col = cb + (lane & 31);
      const float z = acc[r] + P[P_B2 + col];
      Z2[row][col] = z;
      H2[row][col] = tobf(gelu(z));  // the activation is BF16 like h1
    }
  }
  __syncthreads();
  // ---- layer 3: y = h2 w3 + b3 (4 threads per sample) ----
  {
    const int s = tid >> 2, part = tid & 3;
    float acc = 0.f;
    for (int k = part * 16; k < part * 16 + 16; k++) acc += H2[s][k] * (float)tobf(P[P_W3 + k]);
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (part == 0) {
      const float y = acc + P[P_B3];
      yv[s] = y;
      const float base = s0 + s < a.n ? a.feats[(s0 + s) * NF + (NF - 1)] : 0.f;  // residual model
      if (a.pred && s0 + s < a.n) a.pred[s0 + s] = base + y * a.norm[2 * NF + 1] + a.norm[2 * NF];
      float d = 0.f;
      if (TRAIN && s0 + s < a.n) d = y - (a.target[s0 + s] - base - a.norm[2 * NF]) / a.norm[2 * NF + 1];
      dy[s] = d;  // residual; the loss gradient is 2 d / n
    }
  }
  __syncthreads();
  if (!TRAIN) return;

  float* G = a.partial + (int64_t)blockIdx.x * NPARAM;
  if (tid == 0) {
    float l = 0.f;
    for (int s = 0; s < 64; s++) l += dy[s] * dy[s];
    a.loss_part[blockIdx.x] = l;
  }
  __syncthreads();
  if (tid < 64) dy[tid] *= 2.0f * a.inv_n;
  __syncthreads();
  // grads of w3, b3
  if (tid < NH) {
    float g = 0.f;
    for (int s = 0; s < 64; s++) g += H2[s][tid] * dy[s];
    G[P_W3 + tid] = g;
  }
  if (tid == 64) {
    float g = 0.f;
    for (int s = 0; s < 64; s++) g += dy[s];
    G[P_B3] = g;
  }
  __syncthreads();  // H2 is overwritten with dZ2 below
  // dZ2 = dy w3^T * GELU'(Z2)
  for (int e = tid; e < 64 * NH; e += 256) {
    const int s = e / NH, k = e % NH;
    const float d = dy[s] * (float)tobf(P[P_W3 + k]) * gelu_d(Z2[s][k]);
    const __bf16 db = tobf(d);
    D2[s][k] = db;
    D2T[k][s] = db;
    H2[s][k] = d;  // keep the FP32 dZ2 for the bias gradient
  }
  __syncthreads();
  if (tid < NH) {
    float g = 0.f;
    for (int s = 0; s < 64; s++) g += H2[s][tid];
    G[P_B2 + tid] = g;
  }
  // gW2[k][m] = sum_s H1[s][k] dZ2[s][m]   (rows k from H1T, K = samples)
  {
    v16f acc = {};
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += 16) {
      const v8bf A = *reinterpret_cast<const v8bf*>(&H1T[rb + (lane & 31)][k0 + 8 * (lane >> 5)]);
      const v8bf Bv = *reinterpret_cast<const v8bf*>(&D2T[cb + (lane & 31)][k0 + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int k = rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), m = cb + (lane & 31);
      G[P_W2 + k * NH + m] = acc[r];
    }
  }
  // dH1[s][k] = sum_m dZ2[s][m] W2[k][m];  dZ1 = dH1 * GELU'(Z1)
  {
    v16f acc = {};
#pragma unroll
    for (int m0 = 0; m0 < NH; m0 += 16) {
      const v8bf A = *reinterpret_cast<const v8bf*>(&D2[rb + (lane & 31)][m0 + 8 * (lane >> 5)]);
      const v8bf Bv = *reinterpret_cast<const v8bf*>(&W2R[cb + (lane & 31)][m0 + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int s = rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), k = cb + (lane & 31);
      const float d = acc[r] * gelu_d(Z1[s][k]);
      D1f[s][k] = d;
      D1T[k][s] = tobf(d);
    }
  }
  __syncthreads();
  if (tid < NH) {
    float g = 0.f;
    for (int s = 0; s < 64; s++) g += D1f[s][tid];
    G[P_B1 + tid] = g;
  }
  // gW1[f][k] = sum_s X[s][f] dZ1[s][k]: 16 x 64, one 16x16 block per wave,
  // K = 64 samples as two 16x16x32 MFMAs
  {
    v4f acc = {};
    const int kb = wave * 16;
#pragma unroll
    for (int s0k = 0; s0k < 64; s0k += 32) {
      const v8bf A = *reinterpret_cast<const v8bf*>(&XT[lane & 15][s0k + 8 * (lane >> 4)]);
      const v8bf Bv = *reinterpret_cast<const v8bf*>(&D1T[kb + (lane & 15)][s0k + 8 * (lane >> 4)]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int f = 4 * (lane >> 4) + r, k = kb + (lane & 15);
      G[P_W1 + f * NH + k] = acc[r];
    }
  }
}

// grad[p] = sum over workgroups (fixed order); loss = sum / n
__global__ void mlp_reduce_kernel(const float* partial, const float* loss_part, int64_t nwg, float inv_n,
                                  float* grad, float* loss) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) {
    float g = 0.f;
    for (int64_t w = 0; w < nwg; w++) g += partial[w * NPARAM + p];
    grad[p] = g;
  }
  if (p == 0) {
    float l = 0.f;
    for (int64_t w = 0; w < nwg; w++) l += loss_part[w];
    *loss = l * inv_n;
  }
}

__global__ void adam_kernel(float* params, float* m, float* v, const float* grad, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// keep[p] = 1 if |predicted eps| >= thresh; pred holds log(-eps)
__global__ void screen_kernel(const float* pred, int64_t n, double log_thresh, int* keep) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) keep[p] = ((double)pred[p] >= log_thresh) ? 1 : 0;
}

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

MlpState* state(Ctx* c) {
  if (!c->mlp) c->mlp = new MlpState();
  return c->mlp;
}

void require_init(MlpState* s) {
  if (!s->init) fail(RIMP2_ERR_STATE, "rimp2_mlp_init must run first");
  if (!s->have_norm) fail(RIMP2_ERR_STATE, "rimp2_mlp_set_norm must run first");
}

// runs the MLP kernel on n samples already on the device (s->feats / s->target)
void run_mlp(Ctx* c, MlpState* s, int64_t n, bool train, float* dpred) {
  const int64_t nwg = ceil_div(n, 64);
  MlpArgs a;
  a.feats = s->feats.as<float>();
  a.target = train ? s->target.as<float>() : nullptr;
  a.params = s->params.as<float>();
  a.norm = s->norm.as<float>();
  a.pred = dpred;
  a.partial = train ? static_cast<float*>(s->partial.ensure((size_t)(nwg * NPARAM) * sizeof(float))) : nullptr;
  a.loss_part = train ? static_cast<float*>(s->loss.ensure((size_t)(nwg + 1) * sizeof(float))) + 1 : nullptr;
  a.n = n;
  a.inv_n = 1.0f / (float)n;
  ScopedTimer t(c, RIMP2_K_MLP);
  if (train) {
    hipLaunchKernelGGL(mlp_kernel<true>, dim3((unsigned)nwg), dim3(256), 0, c->stream, a);
    RIMP2_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((unsigned)ceil_div(NPARAM, 256)), dim3(256), 0, c->stream,
                       a.partial, a.loss_part, nwg, a.inv_n, s->grad.as<float>(), s->loss.as<float>());
  } else {
    hipLaunchKernelGGL(mlp_kernel<false>, dim3((unsigned)nwg), dim3(256), 0, c->stream, a);
  }
  RIMP2_HIP_CHECK(hipGetLastError());
}

void upload_samples(Ctx* c, MlpState* s, int64_t n, const float* feats, const float* target) {
  if (n <= 0 || !feats) fail(RIMP2_ERR_ARG, "empty or null feature array");
  float* df = static_cast<float*>(s->feats.ensure((size_t)(n * NF) * sizeof(float)));
  RIMP2_HIP_CHECK(hipMemcpyAsync(df, feats, (size_t)(n * NF) * sizeof(float), hipMemcpyHostToDevice, c->stream));
  if (target) {
    float* dt = static_cast<float*>(s->target.ensure((size_t)n * sizeof(float)));
    RIMP2_HIP_CHECK(hipMemcpyAsync(dt, target, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  }
}

}  // namespace

// features of the context's transformed B into s->feats (device) -> n pairs
static int64_t features_device(Ctx* c, int64_t nocc, int64_t nvirt, const double* d_eo, const double* d_ev) {
  MlpState* s = state(c);
  const int64_t naux = c->naux;
  float* r = static_cast<float*>(s->rowcap.ensure((size_t)(naux * nocc) * sizeof(float)));
  const int64_t np = nocc * (nocc + 1) / 2;
  float* F = static_cast<float*>(s->feats.ensure((size_t)(np * NF) * sizeof(float)));
  ScopedTimer t(c, RIMP2_K_FEATURES);
  hipLaunchKernelGGL(rownorm_kernel, dim3((unsigned)ceil_div(naux * nocc, 256)), dim3(256), 0, c->stream,
                     c->B64.as<double>(), naux, nvirt, nocc, r);
  RIMP2_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(pair_features_kernel, dim3((unsigned)np), dim3(256), 0, c->stream, r, naux, nocc, d_eo, d_ev,
                     nvirt, F);
  RIMP2_HIP_CHECK(hipGetLastError());
  return np;
}

static void check_B(Ctx* c, int64_t nocc, int64_t nvirt) {
  check_pending(c);
  if (!c->have_B) fail(RIMP2_ERR_STATE, "rimp2_transform must run first");
  if (c->nocc != nocc || c->nvirt != nvirt)
    fail(RIMP2_ERR_STATE, "features requested for (nocc=%lld, nvirt=%lld) but the transform holds (%lld, %lld)",
         (long long)nocc, (long long)nvirt, (long long)c->nocc, (long long)c->nvirt);
}

static void upload_eps(Ctx* c, int64_t nocc, int64_t nvirt, const double* eo, const double* ev) {
  if (!eo || !ev) fail(RIMP2_ERR_ARG, "null orbital energies");
  double* d = static_cast<double*>(c->eps_o.ensure((size_t)(nocc + nvirt) * sizeof(double)));
  RIMP2_HIP_CHECK(hipMemcpyAsync(d, eo, (size_t)nocc * sizeof(double), hipMemcpyHostToDevice, c->stream));
  RIMP2_HIP_CHECK(hipMemcpyAsync(d + nocc, ev, (size_t)nvirt * sizeof(double), hipMemcpyHostToDevice, c->stream));
}

}  // namespace rimp2

using namespace rimp2;

extern "C" {

int64_t rimp2_mlp_num_params(void) { return NPARAM; }

int rimp2_pair_features(void* ctx, int64_t nocc, int64_t nvirt, const double* eps_o, const double* eps_v,
                        float* feats) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (!feats) fail(RIMP2_ERR_ARG, "null feats");
  check_B(c, nocc, nvirt);
  upload_eps(c, nocc, nvirt, eps_o, eps_v);
  const double* d = c->eps_o.as<double>();
  const int64_t np = features_device(c, nocc, nvirt, d, d + nocc);
  RIMP2_HIP_CHECK(hipMemcpyAsync(feats, state(c)->feats.ptr, (size_t)(np * NF) * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_mlp_init(void* ctx, uint64_t seed) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  MlpState* s = state(c);
  std::vector<float> p(NPARAM, 0.f);
  // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights from a counter-based stream; zero biases
  auto U = [&](uint64_t k) { return (double)(mix64(mix64(seed ^ 0x5EEDULL) + (k + 1) * 0x9E3779B97F4A7C15ULL) >> 11) * 0x1.0p-53; };
  for (int e = 0; e < NF * NH; e++) p[P_W1 + e] = (float)((2.0 * U(e) - 1.0) / std::sqrt((double)NF));
  for (int e = 0; e < NH * NH; e++) p[P_W2 + e] = (float)((2.0 * U(10000 + e) - 1.0) / std::sqrt((double)NH));
  for (int e = 0; e < NH; e++) p[P_W3 + e] = (float)((2.0 * U(20000 + e) - 1.0) / std::sqrt((double)NH));
  float* dp = static_cast<float*>(s->params.ensure(NPARAM * sizeof(float)));
  RIMP2_HIP_CHECK(hipMemcpyAsync(dp, p.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, c->stream));
  RIMP2_HIP_CHECK(hipMemsetAsync(s->m.ensure(NPARAM * sizeof(float)), 0, NPARAM * sizeof(float), c->stream));
  RIMP2_HIP_CHECK(hipMemsetAsync(s->v.ensure(NPARAM * sizeof(float)), 0, NPARAM * sizeof(float), c->stream));
  s->grad.ensure(NPARAM * sizeof(float));
  if (!s->have_norm) {
    for (int f = 0; f < NF; f++) {
      s->norm_h[f] = 0.f;
      s->norm_h[NF + f] = 1.f;
    }
    s->norm_h[2 * NF] = 0.f;
    s->norm_h[2 * NF + 1] = 1.f;
    RIMP2_HIP_CHECK(hipMemcpyAsync(s->norm.ensure(sizeof(s->norm_h)), s->norm_h, sizeof(s->norm_h),
                                   hipMemcpyHostToDevice, c->stream));
    s->have_norm = true;
  }
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  s->step = 0;
  s->init = true;
  RIMP2_API_END(ctx)
}

int rimp2_mlp_set_norm(void* ctx, const float* mean, const float* stdev, float tmean, float tstd) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (!mean || !stdev) fail(RIMP2_ERR_ARG, "null normalisation arrays");
  MlpState* s = state(c);
  for (int f = 0; f < NF; f++) {
    if (!(stdev[f] > 0.f) || !std::isfinite(mean[f])) fail(RIMP2_ERR_ARG, "feature %d: bad mean/std", f);
    s->norm_h[f] = mean[f];
    s->norm_h[NF + f] = stdev[f];
  }
  if (!(tstd > 0.f) || !std::isfinite(tmean)) fail(RIMP2_ERR_ARG, "bad target mean/std");
  s->norm_h[2 * NF] = tmean;
  s->norm_h[2 * NF + 1] = tstd;
  RIMP2_HIP_CHECK(hipMemcpyAsync(s->norm.ensure(sizeof(s->norm_h)), s->norm_h, sizeof(s->norm_h),
                                 hipMemcpyHostToDevice, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  s->have_norm = true;
  RIMP2_API_END(ctx)
}

int rimp2_mlp_grad(void* ctx, int64_t n, const float* feats, const float* target, float* loss, float* grad) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  MlpState* s = state(c);
  require_init(s);
  if (!target || !loss || !grad) fail(RIMP2_ERR_ARG, "null target / loss / grad");
  upload_samples(c, s, n, feats, target);
  run_mlp(c, s, n, true, nullptr);
  RIMP2_HIP_CHECK(hipMemcpyAsync(grad, s->grad.ptr, NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipMemcpyAsync(loss, s->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_mlp_train_step(void* ctx, int64_t n, const float* feats, const float* target, float lr, float* loss) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  MlpState* s = state(c);
  require_init(s);
  if (!target) fail(RIMP2_ERR_ARG, "null target");
  if (!(lr > 0.f)) fail(RIMP2_ERR_ARG, "learning rate must be positive");
  upload_samples(c, s, n, feats, target);
  run_mlp(c, s, n, true, nullptr);
  s->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)s->step), c2 = 1.f - std::pow(b2, (float)s->step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)ceil_div(NPARAM, 256)), dim3(256), 0, c->stream,
                     s->params.as<float>(), s->m.as<float>(), s->v.as<float>(), s->grad.as<float>(), lr, b1, b2,
                     1e-8f, c1, c2);
  RIMP2_HIP_CHECK(hipGetLastError());
  if (loss) RIMP2_HIP_CHECK(hipMemcpyAsync(loss, s->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_mlp_predict(void* ctx, int64_t n, const float* feats, float* pred) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  MlpState* s = state(c);
  require_init(s);
  if (!pred) fail(RIMP2_ERR_ARG, "null pred");
  upload_samples(c, s, n, feats, nullptr);
  float* dp = static_cast<float*>(s->pred.ensure((size_t)n * sizeof(float)));
  run_mlp(c, s, n, false, dp);
  RIMP2_HIP_CHECK(hipMemcpyAsync(pred, dp, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_mlp_get_params(void* ctx, float* params, int64_t n) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  MlpState* s = state(c);
  if (!s->init) fail(RIMP2_ERR_STATE, "rimp2_mlp_init must run first");
  if (!params || n != NPARAM) fail(RIMP2_ERR_ARG, "parameter vector must hold %d floats", NPARAM);
  RIMP2_HIP_CHECK(hipMemcpyAsync(params, s->params.ptr, NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_mlp_set_params(void* ctx, const float* params, int64_t n) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  MlpState* s = state(c);
  if (!s->init) fail(RIMP2_ERR_STATE, "rimp2_mlp_init must run first");
  if (!params || n != NPARAM) fail(RIMP2_ERR_ARG, "parameter vector must hold %d floats", NPARAM);
  for (int64_t k = 0; k < n; k++)
    if (!std::isfinite(params[k])) fail(RIMP2_ERR_NUMERIC, "parameter %lld not finite", (long long)k);
  RIMP2_HIP_CHECK(hipMemcpyAsync(s->params.ptr, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_screened_energy(void* ctx, int64_t nocc, int64_t nvirt, const double* eps_o, const double* eps_v, int mode,
                          double thresh, double* eps_pair, double* e_tot, int64_t* nskipped) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  MlpState* s = state(c);
  require_init(s);
  if (!e_tot) fail(RIMP2_ERR_ARG, "null e_tot");
  if (!(thresh >= 0.0)) fail(RIMP2_ERR_ARG, "threshold must be >= 0");
  check_B(c, nocc, nvirt);
  if (c->mode != mode) fail(RIMP2_ERR_STATE, "screened energy in mode %d but the transform holds mode %d", mode, c->mode);
  for (int64_t i = 0; i < nocc; i++)
    if (!std::isfinite(eps_o[i])) fail(RIMP2_ERR_NUMERIC, "eps_o not finite");
  for (int64_t a = 0; a < nvirt; a++)
    if (!std::isfinite(eps_v[a])) fail(RIMP2_ERR_NUMERIC, "eps_v not finite");
  double mo = -INFINITY, mv = INFINITY;
  for (int64_t i = 0; i < nocc; i++) mo = std::fmax(mo, eps_o[i]);
  for (int64_t a = 0; a < nvirt; a++) mv = std::fmin(mv, eps_v[a]);
  if (!(2.0 * mo - 2.0 * mv < 0.0)) fail(RIMP2_ERR_NUMERIC, "orbital-energy denominator >= 0");
  upload_eps(c, nocc, nvirt, eps_o, eps_v);
  const double* deo = c->eps_o.as<double>();
  const int64_t np = features_device(c, nocc, nvirt, deo, deo + nocc);
  float* dp = static_cast<float*>(s->pred.ensure((size_t)np * sizeof(float)));
  run_mlp(c, s, np, false, dp);
  int* keep = static_cast<int*>(s->pairs_keep.ensure((size_t)np * sizeof(int)));
  const double lt = thresh > 0.0 ? std::log(thresh) : -INFINITY;
  hipLaunchKernelGGL(screen_kernel, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, c->stream, dp, np, lt, keep);
  RIMP2_HIP_CHECK(hipGetLastError());
  std::vector<int> hk(np);
  RIMP2_HIP_CHECK(hipMemcpyAsync(hk.data(), keep, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<int64_t> list;
  list.reserve(np);
  for (int64_t p = 0; p < np; p++)
    if (hk[p]) list.push_back(p);
  int64_t* dl = static_cast<int64_t*>(s->pair_list.ensure(std::max<size_t>(list.size(), 1) * sizeof(int64_t)));
  if (!list.empty())
    RIMP2_HIP_CHECK(hipMemcpyAsync(dl, list.data(), list.size() * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  double* dpe = eps_pair ? static_cast<double*>(c->eps_pair.ensure((size_t)(nocc * nocc) * sizeof(double))) : nullptr;
  double* det = static_cast<double*>(c->etot.ensure(sizeof(double)));
  pair_energies_device(c, c->geom, c->packed.ptr, c->scales.as<double>(), nocc, nvirt, deo, deo + nocc, dpe, det, dl,
                       (int64_t)list.size());
  if (eps_pair)
    RIMP2_HIP_CHECK(hipMemcpyAsync(eps_pair, dpe, (size_t)(nocc * nocc) * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
  RIMP2_HIP_CHECK(hipMemcpyAsync(e_tot, det, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (nskipped) *nskipped = np - (int64_t)list.size();
  RIMP2_API_END(ctx)
}

}  // extern "C"
