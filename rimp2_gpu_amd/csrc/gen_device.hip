// Seeded fragment generator on the GPU: the same counter-based streams and
// formulas as gen.cpp (standard and locality-decay variants), evaluated with
// the device math library, so values agree with the host generator to a few
// ulps (not bit for bit).  V = A A^T / naux + I is formed on the FP64 MFMA GEMM.
// Used to synthesise large fragments and surrogate training sets without a
// host round trip.
#include "rimp2_common.h"

namespace rimp2 {
namespace {

constexpr uint64_t kPhi = 0x9E3779B97F4A7C15ULL;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed * kPhi + (stream + 1) * 0xD1B54A32D192ED03ULL);
}
__device__ inline double uniform53(uint64_t key, uint64_t n) {
  return (double)(mix64(key + (n + 1) * kPhi) >> 11) * 0x1.0p-53;
}
__device__ inline double normal(uint64_t key, uint64_t n) {
  const uint64_t k = n >> 1;
  const double u1 = 1.0 - uniform53(key, 2 * k);
  const double u2 = uniform53(key, 2 * k + 1);
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586476925286766559 * u2;
  return (n & 1) ? r * sin(th) : r * cos(th);
}

__global__ void gen_x_kernel(uint64_t kx, int64_t naux, int64_t nvirt, int64_t nocc, double lam, double* X) {
  RIMP2_GRID_STRIDE(t, naux * nvirt * nocc) {
    const int64_t q = t % naux, c = t / naux, a = c % nvirt, i = c / nvirt;
    double v = normal(kx, (uint64_t)t) * (0.1 / sqrt(1.0 + (double)a / 16.0));
    if (lam > 0.0) {
      const double xa = ((double)a + 0.5) * (double)nocc / (double)nvirt;
      const double xq = ((double)q + 0.5) * (double)nocc / (double)naux;
      v *= exp(-fabs(xa - (double)i) / lam) * exp(-fabs(xq - (double)i) / lam);
    }
    X[t] = v;
  }
}

__global__ void gen_a_kernel(uint64_t ka, int64_t naux, int64_t nocc, double lam, double* A) {
  RIMP2_GRID_STRIDE(t, naux * naux) {
    const int64_t p = t % naux, k = t / naux;
    double v = normal(ka, (uint64_t)t);
    if (lam > 0.0) v *= exp(-fabs((double)(p - k)) * ((double)nocc / (double)naux) / lam);
    A[t] = v;
  }
}

__global__ void add_identity_kernel(double* V, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) V[p + n * p] += 1.0;
}

__global__ void gen_eps_kernel(uint64_t ko, uint64_t kv, int64_t nocc, int64_t nvirt, double* eo, double* ev) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (eo && t < nocc) eo[t] = -2.0 + 1.7 * uniform53(ko, (uint64_t)t);
  if (ev && t < nvirt) ev[t] = 0.1 + 2.9 * uniform53(kv, (uint64_t)t);
}

}  // namespace
}  // namespace rimp2

using namespace rimp2;

extern "C" int rimp2_gen_fragment_device(void* ctx, uint64_t seed, int64_t naux, int64_t nvirt, int64_t nocc,
                                         double lam, double* d_Pia, double* d_V, double* d_eps_o, double* d_eps_v) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (naux <= 0 || nvirt <= 0 || nocc <= 0) fail(RIMP2_ERR_ARG, "dimensions must be positive");
  if (!(lam == lam)) fail(RIMP2_ERR_ARG, "locality length is NaN");
  hipStream_t s = c->stream;
  if (d_Pia) {
    const int64_t n = naux * nvirt * nocc;
    hipLaunchKernelGGL(gen_x_kernel, dim3(grid_blocks(n, 256)), dim3(256), 0, s, stream_key(seed, 0), naux,
                       nvirt, nocc, lam, d_Pia);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  if (d_V) {
    double* A = static_cast<double*>(c->work.ensure((size_t)(naux * naux) * sizeof(double)));
    hipLaunchKernelGGL(gen_a_kernel, dim3(grid_blocks(naux * naux, 256)), dim3(256), 0, s, stream_key(seed, 1),
                       naux, nocc, lam, A);
    RIMP2_HIP_CHECK(hipGetLastError());
    gemm_nt_f64(s, naux, naux, naux, 1.0 / (double)naux, A, naux, A, naux, 0.0, d_V, naux);
    hipLaunchKernelGGL(add_identity_kernel, dim3((unsigned)ceil_div(naux, 256)), dim3(256), 0, s, d_V, naux);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  if (d_eps_o || d_eps_v) {
    const int64_t m = nocc > nvirt ? nocc : nvirt;
    hipLaunchKernelGGL(gen_eps_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, stream_key(seed, 2),
                       stream_key(seed, 3), nocc, nvirt, d_eps_o, d_eps_v);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  RIMP2_API_END(ctx)
}
