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
// Per-lane Kahan sums are reduced by DPP lane moves and LDS to one FP64 partial
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
  static constexpr int RB = 16, KK = 4, P = 1, LB = 8, ND = 1, NE = 4, KSTEPS = 4, MMA_PER_CALL = 1;
  using frag_t = double;
  using acc_t = v4d;
  using xpose_t = double;  // element type of the diagonal unit's LDS transpose (holds V exactly)
  __device__ static void zero(acc_t& a) { a = v4d{0, 0, 0, 0}; }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], acc[0], 0, 0, 0);
  }
  __device__ static double value(const acc_t (&acc)[ND], int r) { return acc[0][r]; }
  __device__ static int row(int lane, int r) { return 4 * r + (lane >> 4); }
  __device__ static int col(int lane, int /*r*/) { return lane & 15; }
};

struct PolF32 {
  static constexpr int RB = 32, KK = 2, P = 1, LB = 4, ND = 1, NE = 16, KSTEPS = 8, MMA_PER_CALL = 1;
  using frag_t = float;
  using acc_t = v16f;
  using xpose_t = float;
  __device__ static void zero(acc_t& a) {
#pragma unroll
    for (int r = 0; r < 16; r++) a[r] = 0.f;
  }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0], b[0], acc[0], 0, 0, 0);
  }
  __device__ static double value(const acc_t (&acc)[ND], int r) { return (double)acc[0][r]; }
  __device__ static int row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
  __device__ static int col(int lane, int /*r*/) { return lane & 31; }
};

// FP32 on 16x16x4 MFMAs (the 32-wide tile): operands as PolF64 (lane l holds
// row l % 16, k l / 16), result D[4 (l / 16) + r][l % 16] (the f32 layout,
// not the f64 interleave).
struct PolF32R16 {
  static constexpr int RB = 16, KK = 4, P = 1, LB = 4, ND = 1, NE = 4, KSTEPS = 4, MMA_PER_CALL = 1;
  using frag_t = float;
  using acc_t = v4f;
  using xpose_t = float;
  __device__ static void zero(acc_t& a) { a = v4f{0.f, 0.f, 0.f, 0.f}; }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
    acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], acc[0], 0, 0, 0);
  }
  __device__ static double value(const acc_t (&acc)[ND], int r) { return (double)acc[0][r]; }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static int col(int lane, int /*r*/) { return lane & 15; }
};

struct PolBF16x3 {
  static constexpr int RB = 32, KK = 16, P = 2, LB = 16, ND = 1, NE = 16, KSTEPS = 1, MMA_PER_CALL = 3;
  using frag_t = v8bf;
  using acc_t = v16f;
  using xpose_t = float;
  __device__ static void zero(acc_t& a) { PolF32::zero(a); }
  __device__ static void mma(acc_t (&acc)[ND], const frag_t (&a)[P], const frag_t (&b)[P]) {
    // small terms first: hi*lo + lo*hi, then hi*hi
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[1], acc[0], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[0], acc[0], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[0], acc[0], 0, 0, 0);
  }
  __device__ static double value(const acc_t (&acc)[ND], int r) { return (double)acc[0][r]; }
  __device__ static int row(int lane, int r) { return PolF32::row(lane, r); }
  __device__ static int col(int lane, int /*r*/) { return lane & 31; }
};

// Exact-slice int8: S slices, diagonals d = s + t <= DMAX accumulate into
// acc[d] (int32, exact: |sum| <= 5 * 127^2 * naux < 2^31 for naux <= 16384).
template <int S, int DMAX>
struct PolI8 {
  static constexpr int RB = 32, KK = 32, P = S, LB = 16, ND = DMAX + 1, NE = 16, KSTEPS = 1;
  // products s + t = d <= DMAX with 0 <= s, t < S
  static constexpr int MMA_PER_CALL = (DMAX + 1) * (DMAX + 2) / 2 - (DMAX >= S ? (DMAX - S + 1) * (DMAX - S + 2) : 0);
  using frag_t = v4i;
  using acc_t = v16i;
  using xpose_t = double;  // scaled slice sums are not float-exact
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
  __device__ static int col(int lane, int /*r*/) { return lane & 31; }
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

// Sum of v over the 64 lanes of a wave, in a fixed order, without LDS round
// trips: two quad permutes and two row rotates (DPP) leave every lane of a
// 16-lane row holding the row sum, and the four row sums are read from lanes
// 0, 16, 32 and 48.  The result is wave-uniform.
template <int CTRL>
__device__ inline double dpp_mov(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ inline double lane_value(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ inline double wave_sum(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x128>(v);  // row_ror:8
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
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
  int64_t nvirt;                // virtuals before the padding (0: treat all nvp as real)
};

// Output of the amplitude form of the pair kernel: for the unordered
// occupied pair i <= j of a chunk i in [i0, i0 + nchunk), every element
//   T[a + nv b + nv^2 (j - i0) + nv^2 nj (i - i0)] = w_ij V^{ij}_ab / (e_i + e_j - e_b - e_a)
// with w_ij = 1 (i < j), 1/2 (i = j), a, b < nv (padding rows are not written).
struct AmplOut {
  double* T;
  const double* eo;  // [nocc]
  const double* ev;  // [nv]
  int64_t nv, nj, i0;
};

// Bytes of LDS one stage of the operand pipeline occupies (4 operand slabs).
template <class Pol, int TM>
constexpr int pair_stage_bytes() {
  return 4 * Pol::KSTEPS * (TM / Pol::RB) * Pol::P * 64 * Pol::LB;
}

// A diagonal block's exchange tile V^{ji}[A,A] is the transpose of V^{ij}[A,A],
// so where the tile fits one stage buffer (for the epilogue's LDS transpose;
// FP64 elements, or FP32 for the modes that accumulate in FP32, whose values
// FP32 holds exactly) two diagonal blocks share one work unit and no MFMA
// forms the transpose.  Units per occupied pair over nb virtual blocks: the nb(nb-1)/2
// off-diagonal tile pairs A < B, then ceil(nb/2) units of two diagonal
// blocks; without packing, the nb(nb+1)/2 tile pairs A <= B.
template <class Pol, int TM>
constexpr bool packs_diagonal() {
  return TM * TM * (int)sizeof(typename Pol::xpose_t) <= pair_stage_bytes<Pol, TM>();
}

__host__ __device__ inline int64_t pair_units(int64_t nb, bool packed) {
  return packed ? nb * (nb - 1) / 2 + (nb + 1) / 2 : nb * (nb + 1) / 2;
}

enum UnitKind { UNIT_TILE_PAIR = 0, UNIT_TWO_DIAG = 1, UNIT_ONE_DIAG = 2 };

// Unit u of a triangle of nb blocks in flat order -> blocks (A, B) and its
// kind: a tile pair (A < B, or A <= B without packing), two diagonal blocks A
// and B = A + 1, or the last diagonal block A alone (odd nb; B = A, its second
// tile unused).
__device__ inline int unit_decode_flat(int64_t u, int64_t nb, bool packed, int64_t& A, int64_t& B) {
  if (!packed) {
    tri_decode(u, A, B);
    return UNIT_TILE_PAIR;
  }
  const int64_t noff = nb * (nb - 1) / 2;
  if (u < noff) {
    tri_decode(u, A, B);
    B += 1;
    return UNIT_TILE_PAIR;
  }
  A = 2 * (u - noff);
  B = A + 1;
  if (B >= nb) {
    B = A;
    return UNIT_ONE_DIAG;
  }
  return UNIT_TWO_DIAG;
}

// Super-tile edge (in blocks) of the unit order below.
constexpr int64_t UNIT_SUPER = 8;

// Unit u of an occupied pair -> blocks (A, B) and its kind.  The triangle of
// nb x nb blocks is cut into super-tiles of UNIT_SUPER x UNIT_SUPER blocks,
// taken column by column (super-column GB: the off-diagonal super-tiles
// GA < GB, then the diagonal one).  An off-diagonal super-tile of 8 x gB
// blocks lists its units row by row (A outer, B inner); a diagonal one uses
// the flat order over its own gB blocks.  Workgroups that run at the same time
// on one XCD take consecutive units; a run of 32 of them then reads 16 or 24
// distinct operand slabs (46 on average in the flat order at nb = 32), so
// more of the operand stream is served from that XCD's L2.
// For nb <= UNIT_SUPER the order is the flat one.
__device__ inline int unit_decode(int64_t u64, int64_t nb64, bool packed, int64_t& A, int64_t& B) {
  if (nb64 <= UNIT_SUPER) return unit_decode_flat(u64, nb64, packed, A, B);
  // 32-bit from here: units per pair < 2^31 (nb < 2^15)
  constexpr uint32_t G = (uint32_t)UNIT_SUPER;
  const uint32_t u = (uint32_t)u64, nb = (uint32_t)nb64;
  const uint32_t ng = (nb + G - 1) / G;
  // units before full super-column g: sum_{c<g} (G*G*c + dG) = (G*G/2) g(g-1) + dG g
  const uint32_t dG = packed ? G * (G - 1) / 2 + G / 2 : G * (G + 1) / 2;
  auto before = [&](uint32_t g) { return (G * G / 2) * g * (g - 1) + dG * g; };
  const float b = (float)dG - (float)(G * G / 2);
  int32_t gi = (int32_t)((-b + sqrtf(b * b + 2.0f * (float)(G * G) * (float)u)) / (float)(G * G));
  uint32_t gb = gi < 0 ? 0u : (uint32_t)gi;
  if (gb > ng - 1) gb = ng - 1;
  while (gb > 0 && before(gb) > u) gb--;
  while (gb + 1 < ng && before(gb + 1) <= u) gb++;
  uint32_t v = u - before(gb);
  const uint32_t gB = (nb - G * gb < G) ? nb - G * gb : G;
  if (v < gb * G * gB) {
    const uint32_t ga = v / (G * gB), w = v - ga * (G * gB);
    const uint32_t wa = w / gB;
    A = (int64_t)(G * ga + wa);
    B = (int64_t)(G * gb + (w - wa * gB));
    return UNIT_TILE_PAIR;
  }
  v -= gb * G * gB;
  const int kind = unit_decode_flat(v, gB, packed, A, B);
  A += G * gb;
  B += G * gb;
  return kind;
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Load through a global-address-space pointer: a global load counts on the
// vector-memory counter only and cannot alias LDS, where a load through a
// generic pointer (a pointer read from memory, as in the batch's fragment
// table) is a flat load that the compiler must order against every LDS access.
template <class T>
__device__ inline T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// Wave-uniform copies (scalar registers) of values the compiler cannot prove
// uniform (read from memory at a uniform address, or decoded from a uniform
// unit index in vector arithmetic): a buffer resource built from a value in
// vector registers costs a readfirstlane loop around every DMA.
__device__ inline int uni_int(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uni(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <class T>
__device__ inline T* uni(T* p) {
  return reinterpret_cast<T*>(uni((int64_t)reinterpret_cast<uintptr_t>(p)));
}

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

// Raw buffer resource (stride 0) over [base, base + bytes); gfx9 word 3.
__device__ inline __amdgpu_buffer_rsrc_t slab_rsrc(const unsigned char* base, size_t bytes) {
  const int n = bytes > 0xFFFFFFFFull ? -1 : (int)(uint32_t)bytes;
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, n, 0x00020000);
}

// s_waitcnt on the vector-memory counter only (expcnt / lgkmcnt left at their
// maxima): returns once at most n of this wave's vector loads are in flight.
#define RIMP2_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt((((n) & 15) | (((n) >> 4) << 14)) | 0x0F70)
// Same, and additionally every LDS read of this wave has returned (lgkmcnt 0).
#define RIMP2_WAIT_VMCNT_LGKM0(n) __builtin_amdgcn_s_waitcnt((((n) & 15) | (((n) >> 4) << 14)) | 0x0070)

// Threads per pair workgroup: eight waves, four per tile.
constexpr int PAIR_THREADS = 512;

// Waves per SIMD the pair kernels are compiled for: as many workgroups as
// the four stage buffers leave room for in the CU's 160 KB of LDS (two per
// CU below 80 KB of stages), each with two waves per SIMD.  The register
// budget per wave follows (512 / waves), so a second workgroup that the LDS
// admits is not turned away by the register count.
template <class Pol, int TM>
constexpr int pair_waves_per_eu() {
  return 4 * pair_stage_bytes<Pol, TM>() <= 80 * 1024 ? 4 : 2;
}

// One work unit of occupied pair (i <= j) (blocks A, Bk and the unit kind from
// unit_decode).  All PAIR_THREADS threads call it; the return value (the
// unit's energy partial) is valid in thread 0.  Ends with a barrier, so the
// stage buffers may be reused at once.
//
// Waves 0-3 form tile 1, waves 4-7 tile 2, each group on a 2 x 2 grid of
// wave tiles.  Splitting the tiles over two wave groups halves the
// accumulators per wave (the exact-slice modes hold five int32 diagonals per
// output) so that two waves share each SIMD: when one waits at a barrier or
// behind a full memory queue, the other keeps the matrix pipe busy.
//
// Operand pipeline: four LDS stage buffers fed by buffer_load_dwordx4 ... lds
// (each wave instruction moves 1 KB, no staging registers) and two register
// sets per wave.  Step st:
//   wait   own DMA of stage st+1 landed; own LDS reads of stage st finished
//   barrier  stage st+1 is in LDS for every wave; buffer st%4 is free
//   issue  stage st+4 -> buffer st%4           (three steps of latency cover)
//   read   stage st+1 -> register set (st+1)&1 (LDS latency hidden ...)
//   MFMA   stage st from register set st&1      (... behind these)
// The body is unrolled four times (4 buffers, 2 register sets) so every
// buffer and register set is named statically, and the last nst % 4 stages
// follow as straight-line steps (the auxiliary axis is padded to one stage;
// to four in the exact-slice modes, which have no tail steps).  The
// prologue's four issues may run past stage nst - 1 into the next orbital's
// block or beyond the slab record (zeros); those buffers are never read.  The
// issue is unconditional (the last steps re-fetch stage nst-1 into a free
// buffer, the last read fetches a valid but unused buffer): a branch around
// the DMA makes the compiler lose track of which buffer is in flight and wait
// for all DMA before every LDS read.  Because the packed layout stores every
// fragment in lane order, a wave's DMA destination is one contiguous KB and
// the MFMA operand reads (ds_read of 64 x lane_bytes) are contiguous and
// bank-conflict free.
//
// DMA addressing: the slab resource and the stage offset of each transfer are
// wave-uniform (SGPRs) and the per-lane part is a loop-invariant 32-bit
// offset, so each transfer is one buffer load plus the M0 write.
// AMPL: the amplitude form (AmplOut above) instead of the energy epilogue.
// SKIP: waves whose block is all virtual padding skip their MFMAs (the batch,
// where most fragments' nvirt is not a multiple of TM; off elsewhere, where the
// branch measured 0.3-0.5% slower on 256/2048/6144 with nothing to skip).
template <class Pol, int TM, bool AMPL = false, bool SKIP = false>
__device__ double pair_tile(const FragView& f, int64_t i, int64_t j, int64_t A, int64_t Bk, int kind,
                            unsigned char* __restrict__ buf0, unsigned char* __restrict__ buf1,
                            unsigned char* __restrict__ buf2, unsigned char* __restrict__ buf3,
                            const AmplOut* ao = nullptr) {
  constexpr int NW = PAIR_THREADS / 64;       // waves
  constexpr int FPB = TM / Pol::RB;           // fragments per block side
  constexpr int WT = TM / 2;                  // wave tile side
  constexpr int FW = WT / Pol::RB;            // fragments per wave side
  constexpr int FB = 64 * Pol::LB;            // bytes per fragment
  constexpr int CHUNK = FPB * Pol::P * FB;    // contiguous bytes per operand per k-step
  constexpr int OPB = Pol::KSTEPS * CHUNK;    // bytes per operand per stage
  constexpr int SB = 4 * OPB;                 // bytes per stage
  constexpr int NI = SB / (NW * 1024);        // 1-KB DMA instructions per wave per stage
  constexpr int KS = Pol::KSTEPS;
  using frag_t = typename Pol::frag_t;
  static_assert(FW >= 1, "wave tile smaller than one MFMA fragment");
  static_assert(SB % (NW * 1024) == 0 && OPB % 1024 == 0, "stage must split into whole 1-KB wave transfers");
  static_assert(WT * WT * 8 <= SB, "the exchange hand-over needs one FP64 wave tile per stage buffer");
  const bool same_occ = (i == j), same_blk = (A == Bk);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int grp = wave >> 2;  // 0: tile 1, 1: tile 2
  // Wave tile (wr, wc) = block k = 2 wr + wc of the group's 2 x 2.  Tile 1's
  // waves take block k = wave.  With SKIP, a wave whose block lies wholly in
  // the virtual padding (the last block of a fragment whose nvirt is not a
  // multiple of TM) skips its MFMAs and LDS reads, and in a unit with a
  // padding block tile 2's waves take block 3 - (wave - 4): with waves w and
  // w + 4 on one SIMD, each SIMD then holds block k of tile 1 and block 3 - k
  // of tile 2, so when the second row half or the second column half of both
  // tiles is padding every SIMD keeps one busy wave.  Otherwise tile 2's waves
  // take block wave - 4.
  const int64_t nv = f.nvirt > 0 ? f.nvirt : f.nvp;
  const bool padded = SKIP && ((Bk + 1) * TM > nv || (A + 1) * TM > nv || kind == UNIT_ONE_DIAG);
  const int wq = (grp && padded) ? 3 - (wave & 3) : (wave & 3);
  const int wr = wq >> 1, wc = wq & 1;

  typename Pol::acc_t acc[FW][FW][Pol::ND];
#pragma unroll
  for (int x = 0; x < FW; x++)
#pragma unroll
    for (int y = 0; y < FW; y++)
#pragma unroll
      for (int d = 0; d < Pol::ND; d++) Pol::zero(acc[x][y][d]);

  // operand slabs (rows, cols of tile 1; rows, cols of tile 2):
  //   tile pair:       B_i[A], B_j[B]; B_j[A], B_i[B]   (V^{ij}[A,B], V^{ji}[A,B])
  //   diagonal blocks: B_i[A], B_j[A]; B_i[B], B_j[B]   (V^{ij}[A,A], V^{ij}[B,B])
  const bool tile_pair = (kind == UNIT_TILE_PAIR);
  const int64_t o1 = tile_pair ? j : i, o2 = tile_pair ? i : j;  // orbitals of slabs 2, 3
  const int64_t b1 = tile_pair ? Bk : A, b2 = tile_pair ? A : Bk;  // blocks of slabs 1, 2
  const size_t per_ks = (size_t)f.nrb * Pol::P * FB;
  const size_t per_occ = (size_t)f.nks * per_ks;
  const size_t stage_stride = (size_t)KS * per_ks;
  const unsigned char* s0 = f.packed + i * per_occ + (size_t)A * CHUNK;
  const unsigned char* s1 = f.packed + j * per_occ + (size_t)b1 * CHUNK;
  const unsigned char* s2 = f.packed + o1 * per_occ + (size_t)b2 * CHUNK;
  const unsigned char* s3 = f.packed + o2 * per_occ + (size_t)Bk * CHUNK;
  const int64_t nst = f.nks / KS;

  // One buffer resource per slab, covering the rest of that occupied
  // orbital's packed block (per_occ < 4 GiB is checked on the host; a transfer
  // outside the record range would read zeros, never other memory).
  // Transfer n of this wave: LDS bytes [(NW n + wave) KB, +1 KB) of the stage,
  // from slab op (resource rs[n]) at per-lane offset voff[n] plus the
  // wave-uniform stage offset.
  const __amdgpu_buffer_rsrc_t r0 = slab_rsrc(s0, per_occ - (size_t)A * CHUNK);
  const __amdgpu_buffer_rsrc_t r1 = slab_rsrc(s1, per_occ - (size_t)b1 * CHUNK);
  const __amdgpu_buffer_rsrc_t r2 = slab_rsrc(s2, per_occ - (size_t)b2 * CHUNK);
  const __amdgpu_buffer_rsrc_t r3 = slab_rsrc(s3, per_occ - (size_t)Bk * CHUNK);
  __amdgpu_buffer_rsrc_t rs[NI];
  uint32_t voff[NI];
#pragma unroll
  for (int n = 0; n < NI; n++) {
    const int wb = (n * NW + wave) * 1024;
    const int op = wb / OPB;
    const int rem = wb - op * OPB + lane * 16;
    voff[n] = (uint32_t)((rem / CHUNK) * per_ks + (rem % CHUNK));
    rs[n] = op == 0 ? r0 : op == 1 ? r1 : op == 2 ? r2 : r3;
  }
  auto issue = [&](int64_t st, unsigned char* dst) {
    const uint32_t so = (uint32_t)((size_t)st * stage_stride);
#pragma unroll
    for (int n = 0; n < NI; n++)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs[n], (lds_ptr_t)(dst + (n * NW + wave) * 1024), 16, voff[n], so, 0,
                                               0);
  };

  struct Ops {
    frag_t r[KS][FW][Pol::P], c[KS][FW][Pol::P];
  };
  // this wave's operands: rows from slab 2 grp, columns from slab 2 grp + 1
  const int slab_r = 2 * grp, slab_c = 2 * grp + 1;
  auto frag = [&](const unsigned char* stage, int ks, int slab, int blk, int pl) {
    return *reinterpret_cast<const frag_t*>(stage + ks * CHUNK + slab * OPB + (blk * Pol::P + pl) * FB +
                                            lane * Pol::LB);
  };
  auto load = [&](const unsigned char* stage, Ops& o) {
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int pl = 0; pl < Pol::P; pl++) {
          o.r[ks][x][pl] = frag(stage, ks, slab_r, wr * FW + x, pl);
          o.c[ks][x][pl] = frag(stage, ks, slab_c, wc * FW + x, pl);
        }
  };
  auto mma = [&](const Ops& o) {
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int y = 0; y < FW; y++) Pol::mma(acc[x][y], o.r[ks][x], o.c[ks][y]);
  };
  // Issue is in order per wave: a DMA or LDS read that waits for a full queue
  // also holds back every MFMA behind it.  Interleaving one DMA and a share of
  // the LDS reads between groups of MFMAs keeps both queues short and the
  // matrix pipe fed while the stage transfer proceeds.  Group order: one MFMA,
  // the group's LDS reads, its other MFMAs, then the DMA (measured against
  // six other orders on one box: 2% faster for the exact-slice kernel than
  // DMA-reads-MFMAs, no worse for FP64; front-loading all LDS reads or all
  // DMAs was slower).
  constexpr int NLDS = 2 * KS * FW * Pol::P;
  constexpr int NMMA = KS * FW * FW * Pol::MMA_PER_CALL;
  constexpr int GL = (NLDS + NI - 1) / NI, GM = (NMMA + NI - 1) / NI;
  auto step = [&](int64_t st, Ops& cur, Ops& nxt, unsigned char* freed, const unsigned char* next) {
    RIMP2_WAIT_VMCNT_LGKM0(2 * NI);
    pipeline_barrier();
    issue(st + 4 < nst ? st + 4 : nst - 1, freed);
    load(next, nxt);
    mma(cur);
#pragma unroll
    for (int g = 0; g < NI; g++) {
      __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);       // MFMA
      __builtin_amdgcn_sched_group_barrier(0x0100, GL, 0);      // LDS reads
      __builtin_amdgcn_sched_group_barrier(0x0008, GM - 1, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);       // VMEM read (the DMA)
    }
  };

  // Epilogue inputs: the virtual energies and exact-slice row scales of the
  // rows and columns this lane's output elements sit on (rows A and cols Bk
  // for a tile pair, with the scales sigma_i x sigma_j in tile 1 and sigma_j x
  // sigma_i in tile 2; rows and cols the group's own block for a diagonal
  // unit, A for group 0 and Bk for group 1, Bk = A for a lone block, with
  // sigma_i x sigma_j).  They are loaded after the main loop, all at once
  // behind a scheduling barrier, so their latency is paid once per unit and
  // not once per element.
  const int64_t rowblk = (tile_pair || grp == 0) ? A : Bk;
  const int64_t colblk = tile_pair ? Bk : rowblk;
  const int64_t orbR = (tile_pair && grp == 1) ? j : i, orbC = (tile_pair && grp == 1) ? i : j;
  double eRow[FW][Pol::NE], sRow[FW][Pol::NE], eCol[FW], sCol[FW], ei = 0.0, ej = 0.0;

  // a stage of a wave with nothing to compute: its DMA share and the barrier
  auto idle = [&](int64_t st, unsigned char* freed) {
    RIMP2_WAIT_VMCNT_LGKM0(2 * NI);
    pipeline_barrier();
    issue(st + 4 < nst ? st + 4 : nst - 1, freed);
  };
  // this wave's block has real rows and columns (a lone diagonal block leaves
  // group 1 without a tile)
  const int64_t rblk = (tile_pair || grp == 0) ? A : Bk, cblk = tile_pair ? Bk : rblk;
  const bool busy =
      !SKIP || (rblk * TM + wr * WT < nv && cblk * TM + wc * WT < nv && !(kind == UNIT_ONE_DIAG && grp == 1));

  Ops ra, rb;
  issue(0, buf0);
  issue(1, buf1);
  issue(2, buf2);
  issue(3, buf3);
  RIMP2_WAIT_VMCNT(3 * NI);
  pipeline_barrier();
  // Whole groups of four stages (one per stage buffer), then the last 0-3
  // stages, each a straight-line step with its buffers named statically.  Both
  // branches meet every barrier (busy is uniform over the wave).  The
  // exact-slice policies pad the auxiliary axis to whole groups instead
  // (rest = 0): with the tail steps compiled in, their kernels measured 1.3%
  // slower at 64/512/2048 and no faster on C3.
  const int64_t nst4 = is_i8<Pol>::value ? nst : (nst & ~(int64_t)3);
  const int rest = is_i8<Pol>::value ? 0 : (int)(nst - nst4);
  if (busy) {
    load(buf0, ra);
    for (int64_t st = 0; st < nst4; st += 4) {
      step(st, ra, rb, buf0, buf1);
      step(st + 1, rb, ra, buf1, buf2);
      step(st + 2, ra, rb, buf2, buf3);
      step(st + 3, rb, ra, buf3, buf0);
    }
    if (rest >= 1) {
      step(nst4, ra, rb, buf0, buf1);
      if (rest >= 2) {
        step(nst4 + 1, rb, ra, buf1, buf2);
        if (rest >= 3) step(nst4 + 2, ra, rb, buf2, buf3);
      }
    }
  } else {
    for (int64_t st = 0; st < nst4; st += 4) {
      idle(st, buf0);
      idle(st + 1, buf1);
      idle(st + 2, buf2);
      idle(st + 3, buf3);
    }
    if (rest >= 1) {
      idle(nst4, buf0);
      if (rest >= 2) {
        idle(nst4 + 1, buf1);
        if (rest >= 3) idle(nst4 + 2, buf2);
      }
    }
  }
  // the accumulators as FP64 values first: the loads below then reuse their
  // registers
  double val[FW][FW][Pol::NE];
  if constexpr (!AMPL) {
#pragma unroll
    for (int x = 0; x < FW; x++)
#pragma unroll
      for (int y = 0; y < FW; y++)
#pragma unroll
        for (int r = 0; r < Pol::NE; r++) val[x][y][r] = Pol::value(acc[x][y], r);
    __builtin_amdgcn_sched_barrier(0);
    const double* evR = f.eps_v + rowblk * TM + wr * WT;
    const double* evC = f.eps_v + colblk * TM + wc * WT;
#pragma unroll
    for (int x = 0; x < FW; x++) {
#pragma unroll
      for (int r = 0; r < Pol::NE; r++) eRow[x][r] = gload(evR + x * Pol::RB + Pol::row(lane, r));
      eCol[x] = gload(evC + x * Pol::RB + Pol::col(lane, 0));
    }
    if constexpr (is_i8<Pol>::value) {
      const double* sgR = f.sigma + orbR * f.nvp + rowblk * TM + wr * WT;
      const double* sgC = f.sigma + orbC * f.nvp + colblk * TM + wc * WT;
#pragma unroll
      for (int x = 0; x < FW; x++) {
#pragma unroll
        for (int r = 0; r < Pol::NE; r++) sRow[x][r] = gload(sgR + x * Pol::RB + Pol::row(lane, r));
        sCol[x] = gload(sgC + x * Pol::RB + Pol::col(lane, 0));
      }
    }
    ei = gload(f.eps_o + i);
    ej = gload(f.eps_o + j);
    __builtin_amdgcn_sched_barrier(0);
  }
  RIMP2_WAIT_VMCNT_LGKM0(0);  // trailing re-fetches and the unused last read are done before LDS is reused

  if constexpr (AMPL) {
    // ---- epilogue: amplitudes to global memory (registers only) ----
    // tile pair: tile 1 holds V^{ij}[A,B] (a, b); tile 2 holds V^{ji}[A,B]
    // (a, b) = V^{ij}(b, a), written transposed.  Diagonal units: each wave
    // group holds V^{ij}[X,X] of its block X (group 1 idle for a lone block).
    const double w = (i == j) ? 0.5 : 1.0;
    const double eij = ao->eo[i] + ao->eo[j];
    double* __restrict__ T = ao->T + ao->nv * ao->nv * ((j - ao->i0) + ao->nj * (i - ao->i0));
    auto put = [&](int64_t a, int64_t b, double v) {
      if (a < ao->nv && b < ao->nv) T[a + ao->nv * b] = w * (v / ((eij - ao->ev[b]) - ao->ev[a]));
    };
    const bool active = tile_pair || !(grp == 1 && kind == UNIT_ONE_DIAG);
    const int64_t X = grp ? Bk : A;
    if (active) {
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int y = 0; y < FW; y++)
#pragma unroll
          for (int r = 0; r < Pol::NE; r++) {
            const int64_t al = wr * WT + x * Pol::RB + Pol::row(lane, r);
            const int64_t bl = wc * WT + y * Pol::RB + Pol::col(lane, r);
            double v = Pol::value(acc[x][y], r);
            if (!tile_pair) {
              if (is_i8<Pol>::value) v *= f.sigma[i * f.nvp + X * TM + al] * f.sigma[j * f.nvp + X * TM + bl];
              put(X * TM + al, X * TM + bl, v);
            } else if (grp == 0) {
              if (is_i8<Pol>::value) v *= f.sigma[i * f.nvp + A * TM + al] * f.sigma[j * f.nvp + Bk * TM + bl];
              put(A * TM + al, Bk * TM + bl, v);
            } else {
              if (is_i8<Pol>::value) v *= f.sigma[j * f.nvp + A * TM + al] * f.sigma[i * f.nvp + Bk * TM + bl];
              put(Bk * TM + bl, A * TM + al, v);
            }
          }
    }
    __syncthreads();  // every wave's LDS reads are done before the next unit's DMA
    return 0.0;
  }

  // ---- epilogue: energy terms in FP64 ----
  double s = 0.0, c = 0.0;
  auto buf_of = [&](int k) { return k == 0 ? buf0 : k == 1 ? buf1 : k == 2 ? buf2 : buf3; };
  __syncthreads();  // every wave's stage DMA has landed and its reads are done
  if (tile_pair) {
    // tile 2's wave (wr, wc) hands its scaled V^{ji} values to tile 1's wave
    // (wr, wc) through stage buffer 2 wr + wc, in register order (lane-
    // contiguous, conflict free); tile 1's wave forms the energy terms
    double* X = reinterpret_cast<double*>(buf_of(2 * wr + wc));
    if (grp == 1 && !same_occ) {
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int y = 0; y < FW; y++)
#pragma unroll
          for (int r = 0; r < Pol::NE; r++) {
            double v2 = val[x][y][r];
            if (is_i8<Pol>::value) v2 *= sRow[x][r] * sCol[y];  // sigma_j[a] sigma_i[b]
            X[((x * FW + y) * Pol::NE + r) * 64 + lane] = v2;
          }
    }
    __syncthreads();
    if (grp == 0) {
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int y = 0; y < FW; y++)
#pragma unroll
          for (int r = 0; r < Pol::NE; r++) {
            double v1 = val[x][y][r];
            if (is_i8<Pol>::value) v1 *= sRow[x][r] * sCol[y];  // sigma_i[a] sigma_j[b]
            const double v2 = same_occ ? v1 : X[((x * FW + y) * Pol::NE + r) * 64 + lane];
            const double inv = 1.0 / (ei + ej - eRow[x][r] - eCol[y]);
            double e;
            if (same_blk) e = v1 * (2.0 * v1 - v2);
            else e = 2.0 * (v1 * v1 + v2 * v2 - v1 * v2);
            kahan_add(s, c, e * inv);
          }
    }
  } else if constexpr (packs_diagonal<Pol, TM>()) {
    // each wave group holds one diagonal block X (group 1: block Bk, idle for
    // a lone block): eps += sum_ab V_ab (2 V_ab - V_ba) / Delta_ab with V_ba
    // from an LDS copy of the tile in stage buffer grp (XOR-swizzled columns:
    // the row writes and the transposed reads are both bank-conflict free)
    const bool active = !(grp == 1 && kind == UNIT_ONE_DIAG);
    using xt = typename Pol::xpose_t;
    xt* T = reinterpret_cast<xt*>(grp ? buf1 : buf0);
    auto swz = [](int a, int b) { return a * TM + (b ^ (a & 31)); };
    double v[FW][FW][Pol::NE];
    if (active) {
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int y = 0; y < FW; y++)
#pragma unroll
          for (int r = 0; r < Pol::NE; r++) {
            const int al = wr * WT + x * Pol::RB + Pol::row(lane, r);
            const int bl = wc * WT + y * Pol::RB + Pol::col(lane, r);
            double w = val[x][y][r];
            if (is_i8<Pol>::value) w *= sRow[x][r] * sCol[y];  // sigma_i[a] sigma_j[b]
            v[x][y][r] = w;
            T[swz(al, bl)] = (xt)w;
          }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int x = 0; x < FW; x++)
#pragma unroll
        for (int y = 0; y < FW; y++)
#pragma unroll
          for (int r = 0; r < Pol::NE; r++) {
            const int al = wr * WT + x * Pol::RB + Pol::row(lane, r);
            const int bl = wc * WT + y * Pol::RB + Pol::col(lane, r);
            const double vab = v[x][y][r], vba = (double)T[swz(bl, al)];
            const double inv = 1.0 / (ei + ej - eRow[x][r] - eCol[y]);
            kahan_add(s, c, vab * (2.0 * vab - vba) * inv);
          }
    }
  }
  const double tot = wave_sum(s - c);
  // the wave sums go through the first bytes of buf0: after this barrier no
  // wave reads the hand-over or transpose data any more (every stage buffer
  // holds at least 8 doubles; the final barrier protects them from the next
  // unit's DMA)
  __syncthreads();
  double* wsum = reinterpret_cast<double*>(buf0);
  if (lane == 0) wsum[grp * 4 + wq] = tot;  // by block: the sum below runs in block order
  __syncthreads();
  double res = 0.0;
#pragma unroll
  for (int w = 0; w < NW; w++) res += wsum[w];
  __syncthreads();
  return res;
}

// Declares the four stage buffers of a pair kernel instantiation.
#define RIMP2_PAIR_STAGE_BUFFERS(Pol, TM)                                            \
  __shared__ __align__(16) unsigned char rimp2_buf0[pair_stage_bytes<Pol, TM>()];  \
  __shared__ __align__(16) unsigned char rimp2_buf1[pair_stage_bytes<Pol, TM>()];  \
  __shared__ __align__(16) unsigned char rimp2_buf2[pair_stage_bytes<Pol, TM>()];  \
  __shared__ __align__(16) unsigned char rimp2_buf3[pair_stage_bytes<Pol, TM>()]

}  // namespace pairk
}  // namespace rimp2
