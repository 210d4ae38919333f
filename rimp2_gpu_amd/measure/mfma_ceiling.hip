// Sustained rate of the int8 MFMA that the exact-slice pair kernels execute
// (v_mfma_i32_32x32x32_i8), measured on the card the benchmark runs on.
//
// The published INT8 peak (5.0 POP/s dense) assumes the full engine clock.
// Under a sustained matrix load the card lowers its clock by how much power
// the operands draw: operands with few toggling bits keep it near 2.35 GHz,
// random operands (what the slices of real data are) pull it far lower.  This
// probe gives the rate the card holds with random operands, in three forms:
//   0 registers only: each wave repeats the pair kernel's per-k-step pattern
//     (5 row slices x 5 column slices, the 15 products with s + t <= 4 into
//     five int32 accumulators) on operands that never leave its registers;
//   1 with LDS reads: the same, with the operands of the next step read from
//     a 40 KB LDS region each step (10 x 1 KB per wave per 15 MFMAs, the
//     pair kernel's ratio);
//   2 the pair kernel's whole operand flow: as 1, plus per step a workgroup
//     barrier and 40 KB of global -> LDS DMA (buffer_load ... lds, five 1-KB
//     transfers per wave) into a four-stage LDS ring from a 10 MB buffer of
//     random bytes that stays in the caches, the ring read one stage behind
//     the DMA.
// One workgroup of eight waves per CU, two waves per SIMD, as the pair kernel.
// Used by the benchmark to put the pair kernel's achieved rate beside the
// rate the card sustains; no part of the energy path calls it.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int THREADS = 512;
constexpr int LDS_BYTES = 40 * 1024;  // 40 fragments of 1 KB (64 lanes x 16 bytes)
constexpr int FRAGS = 10;             // 5 row slices + 5 column slices
constexpr int PRODUCTS = 15;          // slice products with s + t <= 4

// 32-bit integer hash (Wellons' "lowbias32"): random operand bytes from an index.
__device__ inline uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

template <bool LDS>
__global__ __launch_bounds__(THREADS) void i8_mfma_loop(int64_t iters, uint32_t seed, int* sink) {
  __shared__ __align__(16) unsigned char lds[LDS ? LDS_BYTES : 16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t gtid = blockIdx.x * THREADS + threadIdx.x;
  if (LDS) {
    for (int o = threadIdx.x; o < LDS_BYTES / 4; o += THREADS)
      reinterpret_cast<uint32_t*>(lds)[o] = mix(seed ^ (uint32_t)(blockIdx.x * (LDS_BYTES / 4) + o) * 0x9E3779B9u);
    __syncthreads();
  }
  v4i f[2][FRAGS];
  for (int q = 0; q < FRAGS; q++)
    for (int e = 0; e < 4; e++) f[0][q][e] = f[1][q][e] = (int)mix(seed + gtid * (4 * FRAGS) + 4 * q + e);
  v16i acc[5];
  for (int d = 0; d < 5; d++)
    for (int e = 0; e < 16; e++) acc[d][e] = 0;
  for (int64_t it = 0; it < iters; it += 2) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (LDS) {
        // next step's operands: fragments base .. base + 9 of the 40 (in bounds: base + 9 < 40)
        const int base = (int)(((it + h) * 3 + wave) % 30);
#pragma unroll
        for (int q = 0; q < FRAGS; q++)
          f[h ^ 1][q] = *reinterpret_cast<const v4i*>(lds + (base + q) * 1024 + lane * 16);
      }
#pragma unroll
      for (int s = 0; s < 5; s++)
#pragma unroll
        for (int t = 0; t < 5; t++)
          if (s + t <= 4) acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f[h][s], f[h][5 + t], acc[s + t], 0, 0, 0);
    }
  }
  int r = 0;
  for (int d = 0; d < 5; d++)
    for (int e = 0; e < 16; e++) r += acc[d][e];
  sink[gtid] = r;  // keeps the accumulators live
}

// Form 2: the flow of one pair-kernel stage per step (see the header).
constexpr int RING = 4, STAGE = 40 * 1024, NI = STAGE / (8 * 1024);
constexpr int SRC_BLOCKS = 64;  // source window: 64 x ring bytes = 10 MB
#define PROBE_WAIT_VMCNT_LGKM0(n) __builtin_amdgcn_s_waitcnt((((n) & 15) | (((n) >> 4) << 14)) | 0x0070)
typedef __attribute__((address_space(3))) void* lds_ptr_t;

__global__ __launch_bounds__(THREADS) void i8_flow_loop(int64_t iters, const unsigned char* src, int* sink) {
  __shared__ __align__(16) unsigned char lds[RING * STAGE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t gtid = blockIdx.x * THREADS + threadIdx.x;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void*)src, (short)0, SRC_BLOCKS * RING * STAGE, 0x00020000);
  for (int o = threadIdx.x; o < RING * STAGE / 4; o += THREADS)
    reinterpret_cast<uint32_t*>(lds)[o] = mix(12345u ^ (uint32_t)(blockIdx.x * (RING * STAGE / 4) + o));
  __syncthreads();
  v4i f[2][FRAGS];
  for (int q = 0; q < FRAGS; q++)
    for (int e = 0; e < 4; e++) f[0][q][e] = f[1][q][e] = (int)mix(gtid * (4 * FRAGS) + 4 * q + e);
  v16i acc[5];
  for (int d = 0; d < 5; d++)
    for (int e = 0; e < 16; e++) acc[d][e] = 0;
  const uint32_t window = (uint32_t)(blockIdx.x % SRC_BLOCKS) * (uint32_t)(RING * STAGE);
  for (int64_t it = 0; it < iters; it += 2) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int64_t st = it + h;
      PROBE_WAIT_VMCNT_LGKM0(2 * NI);  // this wave's DMA of two stages ago and its LDS reads are done
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const int wslot = (int)((st + 3) % RING), rslot = (int)((st + 1) % RING);
      unsigned char* dst = lds + wslot * STAGE;
      const unsigned char* rd = lds + rslot * STAGE;
      const uint32_t so = window + (uint32_t)(wslot * STAGE);
#pragma unroll
      for (int n = 0; n < NI; n++)  // transfer n: 1 KB at (n * 8 + wave) KB of the stage
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_ptr_t)(dst + (n * 8 + wave) * 1024), 16,
                                                 (uint32_t)(lane * 16 + (n * 8 + wave) * 1024), so, 0, 0);
#pragma unroll
      for (int q = 0; q < FRAGS; q++)  // fragments (3 wave + q) % 40 of the stage: in bounds
        f[h ^ 1][q] = *reinterpret_cast<const v4i*>(rd + ((wave * 3 + q) % 40) * 1024 + lane * 16);
#pragma unroll
      for (int s = 0; s < 5; s++)
#pragma unroll
        for (int t = 0; t < 5; t++)
          if (s + t <= 4) acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f[h][s], f[h][5 + t], acc[s + t], 0, 0, 0);
#pragma unroll
      for (int g = 0; g < NI; g++) {  // as the pair kernel: one MFMA, LDS reads, two MFMAs, one DMA
        __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x0100, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x0008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);
      }
    }
  }
  PROBE_WAIT_VMCNT_LGKM0(0);  // no DMA in flight when the wave ends
  int r = 0;
  for (int d = 0; d < 5; d++)
    for (int e = 0; e < 16; e++) r += acc[d][e];
  sink[gtid] = r;
}

// Random operand bytes for the DMA source (constant bytes would keep the
// multipliers idle and the clock high).
__global__ void fill_random(uint32_t* p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    p[k] = mix((uint32_t)k * 0x9E3779B9u ^ 0x2545F491u);
}

int check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "rimp2_measure_i8_mfma: %s: %s\n", what, hipGetErrorString(e));
    return 1;
  }
  return 0;
}

}  // namespace

// Times one launch of `iters` steps (rounded up to even; at least 2) after a
// short warm-up launch on `device`.  lds_reads = 0: operands in registers
// only; 1: the next step's operands read from LDS each step; 2: the pair
// kernel's whole operand flow (LDS reads, barrier, global -> LDS DMA).  Outputs:
// seconds of the timed launch, int8 operations it executed (2 * 32^3 per
// MFMA) and MFMAs each SIMD issued.  Returns 0, or 1 on a HIP error (printed
// to stderr).
extern "C" int rimp2_measure_i8_mfma(int device, int64_t iters, int lds_reads, double* seconds, double* ops,
                                     double* mfma_per_simd) {
  if (iters < 2) iters = 2;
  iters += iters & 1;
  if (!seconds || !ops || !mfma_per_simd) return 1;
  if (check(hipSetDevice(device), "hipSetDevice")) return 1;
  int ncu = 0;
  if (check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device), "CU count")) return 1;
  if (lds_reads < 0 || lds_reads > 2) return 1;
  int* sink = nullptr;
  unsigned char* src = nullptr;
  if (check(hipMalloc(&sink, (size_t)ncu * THREADS * sizeof(int)), "hipMalloc")) return 1;
  if (lds_reads == 2) {
    const size_t sb = (size_t)SRC_BLOCKS * RING * STAGE;
    int rc0 = check(hipMalloc(&src, sb), "hipMalloc");
    if (!rc0) {
      hipLaunchKernelGGL(fill_random, dim3(1024), dim3(256), 0, 0, reinterpret_cast<uint32_t*>(src),
                         (int64_t)(sb / 4));
      rc0 = check(hipGetLastError(), "fill launch");
    }
    if (rc0) {
      (void)hipFree(sink);
      if (src) (void)hipFree(src);
      return 1;
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = check(hipEventCreate(&e0), "event") || check(hipEventCreate(&e1), "event");
  float ms = 0.f;
  for (int pass = 0; pass < 2 && !rc; pass++) {
    const int64_t n = pass == 0 ? 20000 : iters;
    rc = check(hipEventRecord(e0, 0), "event record");
    if (!rc) {
      if (lds_reads == 2)
        hipLaunchKernelGGL(i8_flow_loop, dim3(ncu), dim3(THREADS), 0, 0, n, src, sink);
      else if (lds_reads)
        hipLaunchKernelGGL(i8_mfma_loop<true>, dim3(ncu), dim3(THREADS), 0, 0, n, 12345u, sink);
      else
        hipLaunchKernelGGL(i8_mfma_loop<false>, dim3(ncu), dim3(THREADS), 0, 0, n, 12345u, sink);
      rc = check(hipGetLastError(), "launch") || check(hipEventRecord(e1, 0), "event record") ||
           check(hipEventSynchronize(e1), "synchronize") || check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(sink);
  if (src) (void)hipFree(src);
  if (rc) return 1;
  const double waves = (double)ncu * (THREADS / 64);
  *seconds = ms * 1e-3;
  *ops = waves * (double)iters * PRODUCTS * 65536.0;
  *mfma_per_simd = (double)iters * PRODUCTS * (THREADS / 64) / 4.0;
  return 0;
}
