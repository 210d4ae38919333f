// Fused RI-MP2 pair kernel: shared device code (policies + one work unit).
//
// For every unique occupied pair i <= j and every unordered pair of virtual
// blocks {A, B} (A <= B, block width TM) one workgroup forms, over LDS-staged
// auxiliary (Q) slabs on MFMA,
//     tile 1:  V_ab    = sum_Q B^Q_ia B^Q_jb      rows B_i[A] x cols B_j[B]
//     tile 2:  V_ba    = sum_Q B^Q_ja B^Q_ib      rows B_j[A] x cols B_i[B]
// Tile 2 is the exchange partner of tile 1 element by element: swapping the
// roles of the two LDS operand slabs (B_j[A] as rows, B_i[B] as columns)
// produces the (b,a) transpose in the same register layout as tile 1, so the
// epilogue pairs V_ab with V_ba without any data movement.  The epilogue then
// applies the amplitude/energy formula in FP64 registers:
//     eps_ij += sum_{a in A, b in B} t_ab (2 V_ab - V_ba) (+ the (B,A) part),
//     t_ab = V_ab / (e_i + e_j - e_a - e_b)
// which is RIMP2_calc_toccA (RIMP2.F90:399-404: gmocont, deltaEPS, t = g/Delta)
// combined with RIMP2_calc_gocc (RIMP2.F90:577-581) at U = identity.  For
// i == j the matrix V^{ii} is symmetric and tile 2 equals tile 1 (skipped).
// Per-lane Kahan sums are reduced by wave shuffles and LDS to one FP64 partial
// per workgroup; the partials are summed per pair and then over pairs in a
// fixed order with Neumaier compensation (deterministic).
#pragma once
#include "rimp2_common.h"

namespace rimp2 {
namespace pairk {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Precision policies.  RB rows and KK k-values per fragment, P planes,
// LB bytes per lane per fragment, ND accumulators per output fragment.
// ---------------------------------------------------------------------------
struct PolF64 {
  static constexpr int RB = 16, KK = 4, P = 1, LB = 8, ND = 1, NE = 4, KSTEPS = 4;
  using frag_t = double;
  using acc_t = v4d;
  __device__ static void zero(acc_t& a) { a = v4d{0, 0, 0, 0}; }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], acc[0], 0, 0, 0);
  }
  __device__ static double value(const acc_t (&acc)[ND], int r) { return acc[0][r]; }
  __device__ static int row(int lane, int r) { return 4 * r + (lane >> 4); }
  __device__ static int col(int lane, int r) { return lane & 15; }
};

struct PolF32 {
  static constexpr int RB = 32, KK = 2, P = 1, LB = 4, ND = 1, NE = 16, KSTEPS = 8;
  using frag_t = float;
  using acc_t = v16f;
  __device__ static void zero(acc_t& a) {
#pragma unroll
    for (int r = 0; r < 16; r++) a[r] = 0.f;
  }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0], b[0], acc[0], 0, 0, 0);
  }
  __device__ static double value(const acc_t (&acc)[ND], int r) { return (double)acc[0][r]; }
  __device__ static int row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
  __device__ static int col(int lane, int r) { return lane & 31; }
};

struct PolBF16x3 {
  static constexpr int RB = 32, KK = 16, P = 2, LB = 16, ND = 1, NE = 16, KSTEPS = 1;
  using frag_t = v8bf;
  using acc_t = v16f;
  __device__ static void zero(acc_t& a) { PolF32::zero(a); }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
    // small terms first: hi*lo + lo*hi, then hi*hi
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[1], acc[0], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[0], acc[0], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[0], acc[0], 0, 0, 0);
  }
  __device__ static double value(const acc_t (&acc)[ND], int r) { return (double)acc[0][r]; }
  __device__ static int row(int lane, int r) { return PolF32::row(lane, r); }
  __device__ static int col(int lane, int r) { return lane & 31; }
};

// Exact-slice int8: S slices, diagonals d = s + t <= DMAX accumulate into
// acc[d] (int32, exact: |sum| <= 5 * 127^2 * naux < 2^31 for naux <= 16384).
template <int S, int DMAX>
struct PolI8 {
  static constexpr int RB = 32, KK = 32, P = S, LB = 16, ND = DMAX + 1, NE = 16, KSTEPS = 1;
  using frag_t = v4i;
  using acc_t = v16i;
  __device__ static void zero(acc_t& a) {
#pragma unroll
    for (int r = 0; r < 16; r++) a[r] = 0;
  }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
#pragma unroll
    for (int d = DMAX; d >= 0; d--)
#pragma unroll
      for (int s = 0; s < S; s++) {
        const int t = d - s;
        if (t >= 0 && t < S) acc[d] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s], b[t], acc[d], 0, 0, 0);
      }
  }
  // unscaled: sum_d acc[d] * 254^{-d}  (row/col scales applied by the caller)
  __device__ static double value(const acc_t (&acc)[ND], int r) {
    constexpr double w = 1.0 / 254.0;
    double v = (double)acc[DMAX][r];
#pragma unroll
    for (int d = DMAX - 1; d >= 0; d--) v = v * w + (double)acc[d][r];
    return v * (1.0 / (127.0 * 127.0));
  }
  __device__ static int row(int lane, int r) { return PolF32::row(lane, r); }
  __device__ static int col(int lane, int r) { return lane & 31; }
};

template <class Pol>
struct is_i8 { static constexpr bool value = (Pol::LB == 16 && Pol::KK == 32); };

// j*(j+1)/2 + i  ->  (i, j),  i <= j
__device__ inline void tri_decode(int64_t p, int64_t& i, int64_t& j) {
  int64_t jj = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (jj * (jj + 1) / 2 > p) jj--;
  while ((jj + 1) * (jj + 2) / 2 <= p) jj++;
  j = jj;
  i = p - jj * (jj + 1) / 2;
}

__device__ inline void kahan_add(double& s, double& c, double x) {
  double y = x - c;
  double t = s + y;
  c = (t - s) - y;
  s = t;
}


// Operand view of one fragment's packed B (single fragment or one member of
// a batch).
struct FragView {
  const unsigned char* packed;  // packed B (layout in pack.hip)
  const double* sigma;          // int8 row scales [nocc][nvp] (null otherwise)
  const double* eps_o;          // [nocc]
  const double* eps_v;          // [nvp], padding entries = 1.0
  int64_t nvp, nks, nrb;
};

// Bytes of LDS one stage of the operand pipeline occupies (4 operand slabs).
template <class Pol, int TM>
constexpr int pair_stage_bytes() {
  return 4 * Pol::KSTEPS * (TM / Pol::RB) * Pol::P * 64 * Pol::LB;
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Workgroup barrier for the operand pipeline.  __syncthreads() carries a
// workgroup release/acquire fence, and the release half makes the compiler wait
// for every LDS DMA still in flight (vmcnt(0)), which would drain the ring at
// each stage.  The stage data is instead covered by the explicit vmcnt waits
// below (a wave's DMA writes are in LDS once its count drops), so only the
// execution barrier is needed; the signal fences keep the compiler from moving
// LDS accesses across it.
__device__ inline void pipeline_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// s_waitcnt on the vector-memory counter only (expcnt / lgkmcnt left at their
// maxima): returns once at most n of this wave's vector loads are in flight.
#define RIMP2_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt((((n) & 15) | (((n) >> 4) << 14)) | 0x0F70)

// One work unit: occupied pair (i <= j), virtual blocks A <= Bk.  All 256
// threads call it; the return value (the unit's energy partial) is valid in
// thread 0.  Ends with a barrier, so the stage buffers may be reused at once.
//
// Operand pipeline: three LDS stage buffers; the slabs of stages s+1 and s+2
// are in flight HBM/L2 -> LDS (global_load_lds_dwordx4, each wave instruction
// moves 1 KB, no staging registers) while the MFMAs consume stage s.  The
// buffers are separate LDS objects so the compiler can see that the DMA
// target never aliases the buffer being read.  Because the packed
// layout stores every fragment in lane order, the DMA destination of a wave is
// one contiguous KB and the MFMA operand reads (ds_read of 64 x lane_bytes)
// are contiguous and bank-conflict free.
template <class Pol, int TM>
__device__ double pair_tile(const FragView& f, int64_t i, int64_t j, int64_t A, int64_t Bk,
                            unsigned char* __restrict__ buf0, unsigned char* __restrict__ buf1,
                            unsigned char* __restrict__ buf2) {
  constexpr int FPB = TM / Pol::RB;           // fragments per block side
  constexpr int WT = TM / 2;                  // wave tile side
  constexpr int FW = WT / Pol::RB;            // fragments per wave side
  constexpr int FB = 64 * Pol::LB;            // bytes per fragment
  constexpr int CHUNK = FPB * Pol::P * FB;    // contiguous bytes per operand per k-step
  constexpr int OPB = Pol::KSTEPS * CHUNK;    // bytes per operand per stage
  constexpr int SB = 4 * OPB;                 // bytes per stage
  constexpr int NI = SB / 4096;               // 1-KB DMA instructions per wave per stage
  static_assert(FW >= 1, "wave tile smaller than one MFMA fragment");
  static_assert(SB % 4096 == 0 && OPB % 1024 == 0, "stage must split into whole 1-KB wave transfers");
  const bool same_occ = (i == j), same_blk = (A == Bk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;

  typename Pol::acc_t acc1[FW][FW][Pol::ND], acc2[FW][FW][Pol::ND];
#pragma unroll
  for (int x = 0; x < FW; x++)
#pragma unroll
    for (int y = 0; y < FW; y++)
#pragma unroll
      for (int d = 0; d < Pol::ND; d++) {
        Pol::zero(acc1[x][y][d]);
        Pol::zero(acc2[x][y][d]);
      }

  // operand slabs: 0 = B_i[A] (rows), 1 = B_j[B] (cols), 2 = B_j[A], 3 = B_i[B]
  const size_t per_ks = (size_t)f.nrb * Pol::P * FB;
  const size_t per_occ = (size_t)f.nks * per_ks;
  const unsigned char* src[4];
  src[0] = f.packed + i * per_occ + (size_t)A * CHUNK;
  src[1] = f.packed + j * per_occ + (size_t)Bk * CHUNK;
  src[2] = f.packed + j * per_occ + (size_t)A * CHUNK;
  src[3] = f.packed + i * per_occ + (size_t)Bk * CHUNK;
  const int64_t nst = f.nks / Pol::KSTEPS;

  // stage st -> LDS buffer dst (wave-uniform destinations, per-lane sources).
  // Every wave issues exactly NI transfers per stage (all four slabs, also for
  // i == j) so that the wait counts below are compile-time constants.
  auto issue = [&](int64_t st, unsigned char* dst) {
#pragma unroll
    for (int n = 0; n < NI; n++) {
      const int wb = (n * 4 + wave) * 1024;
      const int op = wb / OPB;
      const int rem = wb - op * OPB + lane * 16;
      const int ks = rem / CHUNK, r = rem % CHUNK;
      const unsigned char* g = src[op] + (size_t)(st * Pol::KSTEPS + ks) * per_ks + r;
      __builtin_amdgcn_global_load_lds(g, (lds_ptr_t)(dst + wb), 16, 0, 0);
    }
  };
  auto compute = [&](const unsigned char* stage) {
#pragma unroll
    for (int ks = 0; ks < Pol::KSTEPS; ks++) {
      typename Pol::frag_t ra[FW][Pol::P], cb[FW][Pol::P];
      const unsigned char* base = stage + ks * CHUNK;
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int pl = 0; pl < Pol::P; pl++) {
          ra[x][pl] = *reinterpret_cast<const typename Pol::frag_t*>(
              base + 0 * OPB + ((wr * FW + x) * Pol::P + pl) * FB + lane * Pol::LB);
          cb[x][pl] = *reinterpret_cast<const typename Pol::frag_t*>(
              base + 1 * OPB + ((wc * FW + x) * Pol::P + pl) * FB + lane * Pol::LB);
        }
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int y = 0; y < FW; y++) Pol::mma(acc1[x][y], ra[x], cb[y]);
      if (!same_occ) {
#pragma unroll
        for (int x = 0; x < FW; x++)
#pragma unroll
          for (int pl = 0; pl < Pol::P; pl++) {
            ra[x][pl] = *reinterpret_cast<const typename Pol::frag_t*>(
                base + 2 * OPB + ((wr * FW + x) * Pol::P + pl) * FB + lane * Pol::LB);
            cb[x][pl] = *reinterpret_cast<const typename Pol::frag_t*>(
                base + 3 * OPB + ((wc * FW + x) * Pol::P + pl) * FB + lane * Pol::LB);
          }
#pragma unroll
        for (int x = 0; x < FW; x++)
#pragma unroll
          for (int y = 0; y < FW; y++) Pol::mma(acc2[x][y], ra[x], cb[y]);
      }
    }
  };

  // step(st): wait for stage st (stage st+1 stays in flight), barrier (stage
  // st visible to all waves; the buffer of stage st-1 is free), issue stage
  // st+2 into that buffer, consume stage st.  The issue is unconditional (the
  // last two steps re-fetch stage nst-1 into a free buffer): a branch around
  // the DMA makes the compiler lose track of which buffer is in flight and
  // wait for all DMA before every LDS read.
  auto step = [&](int64_t st, const unsigned char* cur, unsigned char* next2) {
    RIMP2_WAIT_VMCNT(NI);
    pipeline_barrier();
    issue(st + 2 < nst ? st + 2 : nst - 1, next2);
    compute(cur);
  };
  issue(0, buf0);
  issue(1, buf1);
  // nst is a multiple of 3 (the auxiliary padding guarantees it), so the
  // three-step body needs no early exits.
  for (int64_t st = 0; st < nst; st += 3) {
    step(st, buf0, buf2);
    step(st + 1, buf1, buf0);
    step(st + 2, buf2, buf1);
  }
  RIMP2_WAIT_VMCNT(0);  // the trailing re-fetches have landed before LDS is reused

  // ---- epilogue: energy terms in FP64 ----
  const double ei = f.eps_o[i], ej = f.eps_o[j];
  double s = 0.0, c = 0.0;
#pragma unroll
  for (int x = 0; x < FW; x++)
#pragma unroll
    for (int y = 0; y < FW; y++)
#pragma unroll
      for (int r = 0; r < Pol::NE; r++) {
        const int64_t a = A * TM + wr * WT + x * Pol::RB + Pol::row(lane, r);
        const int64_t b = Bk * TM + wc * WT + y * Pol::RB + Pol::col(lane, r);
        double v1 = Pol::value(acc1[x][y], r);
        double v2 = same_occ ? v1 : Pol::value(acc2[x][y], r);
        if (is_i8<Pol>::value) {
          v1 *= f.sigma[i * f.nvp + a] * f.sigma[j * f.nvp + b];
          if (!same_occ) v2 *= f.sigma[j * f.nvp + a] * f.sigma[i * f.nvp + b];
          else v2 = v1;
        }
        const double inv = 1.0 / (ei + ej - f.eps_v[a] - f.eps_v[b]);
        double e;
        if (same_blk) e = v1 * (2.0 * v1 - v2);
        else e = 2.0 * (v1 * v1 + v2 * v2 - v1 * v2);
        kahan_add(s, c, e * inv);
      }
  double tot = s - c;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
  // the wave sums go through buf0 once every wave is done with the stages
  // (keeps the workgroup's LDS at exactly the stage buffers)
  __syncthreads();
  double* wsum = reinterpret_cast<double*>(buf0);
  if (lane == 0) wsum[wave] = tot;
  __syncthreads();
  const double res = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
  __syncthreads();
  return res;
}

// Declares the three stage buffers of a pair kernel instantiation.
#define RIMP2_PAIR_STAGE_BUFFERS(Pol, TM)                                            \
  __shared__ __align__(16) unsigned char rimp2_buf0[pair_stage_bytes<Pol, TM>()];  \
  __shared__ __align__(16) unsigned char rimp2_buf1[pair_stage_bytes<Pol, TM>()];  \
  __shared__ __align__(16) unsigned char rimp2_buf2[pair_stage_bytes<Pol, TM>()]

}  // namespace pairk
}  // namespace rimp2
