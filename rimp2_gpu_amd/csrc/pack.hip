// Packing of the transformed 3-index tensor B[naux, nvirt, nocc] (FP64) into
// the pair kernel's operand layout, one layout per precision mode.
//
// Layout ("MFMA-operand order"): for occupied i, k-step ks (kk auxiliary
// values), row block rb (rb virtuals) and plane p, 64 lanes x lane_bytes are
// stored in lane order, so that a wave's operand fragment is one contiguous
// 64*lane_bytes chunk both in HBM and in LDS:
//     packed[i][ks][rb][p][lane][lane_bytes]
// The same fragment serves as the row (A) and the column (B) operand because
// the MFMA A[m][k] and B[k][n] lane maps coincide (m <-> n):
//   fp64 16x16x4 : lane l holds (row l%16, k l/16)              8 bytes
//   fp32 32x32x2 : lane l holds (row l%32, k l/32)              4 bytes
//   fp32 16x16x4 : lane l holds (row l%16, k l/16)              4 bytes (32-wide tile)
//   bf16 32x32x16: lane l holds (row l%32, k 8*(l/32)+e, e<8)  16 bytes
//   i8   32x32x32: lane l holds (row l%32, k 16*(l/32)+e, e<16) 16 bytes
// Padding rows (virtual a >= nvirt) and padding k (Q >= naux) are zero.
//
// Exact-slice (int8) modes: per row (a,i) the scale sigma = max_Q |B[Q,a,i]|,
// x = 127 B / sigma in [-127, 127];  s_1 = rint(x), r_1 = x - s_1,
// s_k = rint(254 r_{k-1}), r_k = 254 r_{k-1} - s_k  (|r| <= 1/2 keeps every
// slice in [-127, 127]).  Then B = sigma/127 * sum_k s_k 254^{-(k-1)} + O(254^{-S}).
#include "rimp2_common.h"

namespace rimp2 {

int mode_slices(int mode) {
  switch (mode) {
    case RIMP2_MODE_I8S4: return 4;
    case RIMP2_MODE_I8S5: return 5;
    default: return 0;
  }
}

// Auxiliary padding quantum per mode: one LDS stage of the pair kernel (16
// auxiliary values; its loop runs whole groups of four stages and then the
// last 0-3), and in the exact-slice modes four stages of 32 (their kernels
// run whole groups only).
static int64_t q_quantum(int mode) {
  switch (mode) {
    case RIMP2_MODE_FP64: return 16;
    case RIMP2_MODE_FP32: return 16;
    case RIMP2_MODE_BF16X3: return 16;
    default: return 4 * 32;
  }
}

PackGeom make_geom(int mode, int64_t naux, int64_t nvirt, int tile) {
  PackGeom g;
  g.mode = mode;
  switch (mode) {
    case RIMP2_MODE_FP64: g.rb = 16; g.kk = 4; g.planes = 1; g.lane_bytes = 8; break;
    case RIMP2_MODE_FP32:  // 32-wide tile: 16x16x4 operands; else 32x32x2
      if (tile == 32) { g.rb = 16; g.kk = 4; } else { g.rb = 32; g.kk = 2; }
      g.planes = 1; g.lane_bytes = 4; break;
    case RIMP2_MODE_BF16X3: g.rb = 32; g.kk = 16; g.planes = 2; g.lane_bytes = 16; break;
    case RIMP2_MODE_I8S4:
    case RIMP2_MODE_I8S5: g.rb = 32; g.kk = 32; g.planes = mode_slices(mode); g.lane_bytes = 16; break;
    default: fail(RIMP2_ERR_ARG, "unknown precision mode %d", mode);
  }
  g.nvp = round_up(nvirt, tile);
  g.nqp = round_up(naux, q_quantum(mode));
  g.nks = g.nqp / g.kk;
  g.nrb = g.nvp / g.rb;
  // the pair kernel addresses one occupied orbital's block with 32-bit offsets
  if (g.bytes_per_occ() > 0xFFFFFFFFull)
    fail(RIMP2_ERR_ARG, "packed block per occupied orbital (%zu bytes) exceeds 4 GiB (naux=%lld nvirt=%lld)",
         g.bytes_per_occ(), (long long)naux, (long long)nvirt);
  return g;
}

namespace {

// This lane's share of max_q |col[q]| over q = lane, lane + 64, ... < n: four
// independent maxima, so four loads of the wave are in flight at once (one
// chain of load-then-max per lane leaves the pass latency-bound).
__device__ inline double col_absmax(const double* __restrict__ col, int64_t n, int lane) {
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  int64_t q = lane;
  for (; q + 192 < n; q += 256) {
    m0 = fmax(m0, fabs(col[q]));
    m1 = fmax(m1, fabs(col[q + 64]));
    m2 = fmax(m2, fabs(col[q + 128]));
    m3 = fmax(m3, fabs(col[q + 192]));
  }
  for (; q < n; q += 64) m0 = fmax(m0, fabs(col[q]));
  return fmax(fmax(m0, m1), fmax(m2, m3));
}

// The same, and z = NaN when some value is NaN or +-Inf (fma(x, 0, z) stays a
// signed zero for finite x; fmax alone drops NaNs).
__device__ inline double col_absmax_z(const double* __restrict__ col, int64_t n, int lane, double& z) {
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, z0 = 0.0, z1 = 0.0;
  int64_t q = lane;
  for (; q + 192 < n; q += 256) {
    const double a = col[q], b = col[q + 64], c = col[q + 128], d = col[q + 192];
    m0 = fmax(m0, fabs(a));
    m1 = fmax(m1, fabs(b));
    m2 = fmax(m2, fabs(c));
    m3 = fmax(m3, fabs(d));
    z0 = fma(a, 0.0, fma(b, 0.0, z0));
    z1 = fma(c, 0.0, fma(d, 0.0, z1));
  }
  for (; q < n; q += 64) {
    m0 = fmax(m0, fabs(col[q]));
    z0 = fma(col[q], 0.0, z0);
  }
  z = z0 + z1;
  return fmax(fmax(m0, m1), fmax(m2, m3));
}

// sigma[i*nvp + a] = max_Q |B[Q,a,i]| (0 for padding rows); one wave per row.
__global__ __launch_bounds__(256) void rowmax_kernel(const double* __restrict__ B, int64_t naux, int64_t nvirt,
                                                     int64_t nocc, int64_t nvp, double* __restrict__ sigma) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nvp * nocc) return;
  const int64_t i = row / nvp, a = row % nvp;
  double m = 0.0;
  if (a < nvirt) m = col_absmax(B + naux * (a + nvirt * i), naux, lane);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if (lane == 0) sigma[row] = m;
}

struct PackArgs {
  const double* B;
  int64_t naux, nvirt, nocc;
  int64_t nks, nrb;
  const double* sigma;
  void* dst;
};

__device__ inline double bget(const PackArgs& p, int64_t q, int64_t a, int64_t i) {
  return (q < p.naux && a < p.nvirt) ? p.B[q + p.naux * (a + p.nvirt * i)] : 0.0;
}

// FP32 in 16x16x4 operand order (the fp64 fragment shape with 4-byte lanes):
// packer id beside the mode numbers.
constexpr int PACK_F32R16 = 16;
// Six int8 slices in the exact-slice layout: the operands of the int8
// metric transform (transform_i8.hip).
constexpr int PACK_I8S6 = 17;

// bf16 round-to-nearest-even of a float (no NaN inputs reach here).
__device__ inline uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Packing by tiles: one workgroup per (occupied orbital i, 32 virtual rows,
// TQ auxiliary values).  The tile is read with coalesced loads (each wave
// instruction reads 64 consecutive auxiliary values of one row) into LDS; each
// thread then forms whole lanes of the MFMA fragments the tile covers and
// writes them in fragment order, where the 64 lanes of a fragment and the
// fragments of consecutive row blocks are contiguous, so a wave's writes are
// contiguous too.  Packed element of lane l of fragment f = (i nks + ks) nrb +
// rb, plane k (P planes of 64 lanes x lane bytes per fragment):
//   fp64 / fp32 16x16x4: row 16 rb + l % 16, aux 4 ks + l / 16
//   fp32 32x32x2:        row 32 rb + l % 32, aux 2 ks + l / 32
//   bf16x3 32x32x16:     row 32 rb + l % 32, aux 16 ks + 8 (l / 32) + e, e < 8;
//                        plane 0 = hi = bf16(b), plane 1 = lo = bf16(b - hi)
//   int8 32x32x32:       row 32 rb + l % 32, aux 32 ks + 16 (l / 32) + e, e < 16;
//                        plane k = slice k of 127 b / sigma (s_k = rint(254 r_{k-1}))
// Tiles at the padded edges may hold fewer k-steps or row blocks; values past
// naux or nvirt read as zero.
template <int MODE>
struct PackT;
template <>
struct PackT<RIMP2_MODE_FP64> { static constexpr int RB = 16, KK = 4, P = 1, TQ = 64; };
template <>
struct PackT<PACK_F32R16> { static constexpr int RB = 16, KK = 4, P = 1, TQ = 64; };
template <>
struct PackT<RIMP2_MODE_FP32> { static constexpr int RB = 32, KK = 2, P = 1, TQ = 64; };
template <>
struct PackT<RIMP2_MODE_BF16X3> { static constexpr int RB = 32, KK = 16, P = 2, TQ = 64; };
template <>
struct PackT<RIMP2_MODE_I8S4> { static constexpr int RB = 32, KK = 32, P = 4, TQ = 128; };
template <>
struct PackT<RIMP2_MODE_I8S5> { static constexpr int RB = 32, KK = 32, P = 5, TQ = 128; };
template <>
struct PackT<PACK_I8S6> { static constexpr int RB = 32, KK = 32, P = 6, TQ = 128; };

constexpr int TILE_ROWS = 32;

// Tiles of one fragment's packing: nocc x (tiles of k-steps) x (tiles of row blocks).
template <int MODE>
__host__ __device__ inline int64_t pack_tiles(const PackArgs& p) {
  constexpr int KT = PackT<MODE>::TQ / PackT<MODE>::KK, RT = TILE_ROWS / PackT<MODE>::RB;
  return p.nocc * ((p.nks + KT - 1) / KT) * ((p.nrb + RT - 1) / RT);
}

// Exact-slice planes of one lane: 16 values x (127 / sigma) split into S int8
// slices, s_0 = rint(x), s_k = rint(r_k) with r_k = 254 (r_{k-1} - s_{k-1}).
// No slice leaves [-127, 127]: |x| <= 127 (1 + 2^-53)^2 because sigma is the
// row's largest |B|, and every later |r_k| <= 254 / 2.  rint(r) is formed as
// (r + 1.5 2^52) - 1.5 2^52 (exact for |r| < 2^51 under round-to-nearest-even,
// the rounding of rint); the low 32 bits of r + 1.5 2^52 then hold rint(r) in
// two's complement, so its low byte is the slice.
template <int S>
__device__ inline void slice_lane(const double* v, double inv, uint4 (&out)[S]) {
#pragma clang fp contract(off)  // each product and sum rounded on its own, as written
  constexpr double M = 6755399441055744.0;  // 1.5 * 2^52
  uint32_t w[S][4];
#pragma unroll
  for (int k = 0; k < S; k++)
#pragma unroll
    for (int d = 0; d < 4; d++) w[k][d] = 0;
#pragma unroll
  for (int e = 0; e < 16; e++) {
    double x = v[e] * inv;
#pragma unroll
    for (int k = 0; k < S; k++) {
      const double t = x + M;
      const uint32_t lo = (uint32_t)__builtin_bit_cast(uint64_t, t);
      w[k][e >> 2] |= (lo & 0xFFu) << (8 * (e & 3));
      if (k + 1 < S) x = (x - (t - M)) * 254.0;
    }
  }
#pragma unroll
  for (int k = 0; k < S; k++) out[k] = make_uint4(w[k][0], w[k][1], w[k][2], w[k][3]);
}

// Tile u of p's packing; all 256 threads call it.  tile: LDS [TILE_ROWS][TQ + 1].
template <int MODE>
__device__ void pack_tile(const PackArgs& p, int64_t nvp, int64_t u, double* tile) {
  using T = PackT<MODE>;
  constexpr int TQ = T::TQ, PITCH = TQ + 1, KT = TQ / T::KK, RT = TILE_ROWS / T::RB;
  static_assert(T::P >= 1 && TQ % 64 == 0 && TQ % T::KK == 0 && TILE_ROWS % T::RB == 0,
                "a tile holds whole k-steps and row blocks and loads 64 values per wave instruction");
  const int64_t nat = (p.nrb + RT - 1) / RT, nqt = (p.nks + KT - 1) / KT;
  const int64_t at = u % nat, qt = (u / nat) % nqt, i = u / (nat * nqt);
  const int64_t a0 = at * TILE_ROWS, q0 = qt * TQ;
  const int tid = threadIdx.x;
  {  // wave w loads rows w, w + 4, ...: 64 consecutive auxiliary values per instruction,
     // all of a thread's loads issued before its LDS stores
    const int w = tid >> 6, l = tid & 63;
    double v[TILE_ROWS / 4][TQ / 64];
#pragma unroll
    for (int rr = 0; rr < TILE_ROWS / 4; rr++)
#pragma unroll
      for (int h = 0; h < TQ / 64; h++) v[rr][h] = bget(p, q0 + h * 64 + l, a0 + w + 4 * rr, i);
#pragma unroll
    for (int rr = 0; rr < TILE_ROWS / 4; rr++)
#pragma unroll
      for (int h = 0; h < TQ / 64; h++) tile[(w + 4 * rr) * PITCH + h * 64 + l] = v[rr][h];
  }
  __syncthreads();
  const int nks_t = (int)min((int64_t)KT, p.nks - qt * KT), nrb_t = (int)min((int64_t)RT, p.nrb - at * RT);
  const int items = nks_t * nrb_t * 64;
  for (int it = tid; it < items; it += 256) {
    const int lane = it & 63, fl = it >> 6;
    const int rbl = fl % nrb_t, ksl = fl / nrb_t;
    const int64_t f = (i * p.nks + qt * KT + ksl) * p.nrb + at * RT + rbl;
    if constexpr (MODE == RIMP2_MODE_FP64 || MODE == PACK_F32R16) {
      const double v = tile[(rbl * 16 + (lane & 15)) * PITCH + ksl * 4 + (lane >> 4)];
      if constexpr (MODE == RIMP2_MODE_FP64) static_cast<double*>(p.dst)[f * 64 + lane] = v;
      else static_cast<float*>(p.dst)[f * 64 + lane] = (float)v;
    } else if constexpr (MODE == RIMP2_MODE_FP32) {
      const double v = tile[(lane & 31) * PITCH + ksl * 2 + (lane >> 5)];
      static_cast<float*>(p.dst)[f * 64 + lane] = (float)v;
    } else if constexpr (MODE == RIMP2_MODE_BF16X3) {
      const double* v = tile + (lane & 31) * PITCH + ksl * 16 + 8 * (lane >> 5);
      uint16_t hi[8], lo[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float b = (float)v[e];
        hi[e] = f2bf(b);
        lo[e] = f2bf(b - bf2f(hi[e]));
      }
      uint4 vh, vl;
      vh.x = hi[0] | ((uint32_t)hi[1] << 16); vh.y = hi[2] | ((uint32_t)hi[3] << 16);
      vh.z = hi[4] | ((uint32_t)hi[5] << 16); vh.w = hi[6] | ((uint32_t)hi[7] << 16);
      vl.x = lo[0] | ((uint32_t)lo[1] << 16); vl.y = lo[2] | ((uint32_t)lo[3] << 16);
      vl.z = lo[4] | ((uint32_t)lo[5] << 16); vl.w = lo[6] | ((uint32_t)lo[7] << 16);
      uint4* dst = static_cast<uint4*>(p.dst);
      dst[(f * 2 + 0) * 64 + lane] = vh;
      dst[(f * 2 + 1) * 64 + lane] = vl;
    } else {
      constexpr int S = T::P;
      const double sg = p.sigma[i * nvp + a0 + (lane & 31)];
      const double inv = (sg > 0.0) ? 127.0 / sg : 0.0;
      uint4 out[S];
      slice_lane<S>(tile + (lane & 31) * PITCH + ksl * 32 + 16 * (lane >> 5), inv, out);
      uint4* dst = static_cast<uint4*>(p.dst);
#pragma unroll
      for (int k = 0; k < S; k++) dst[(f * S + k) * 64 + lane] = out[k];
    }
  }
  __syncthreads();  // the tile is reused by the next one
}

template <int MODE>
__global__ __launch_bounds__(256) void pack_tile_kernel(PackArgs p, int64_t nvp) {
  __shared__ double tile[TILE_ROWS * (PackT<MODE>::TQ + 1)];
  const int64_t n = pack_tiles<MODE>(p);
  for (int64_t u = blockIdx.x; u < n; u += gridDim.x) pack_tile<MODE>(p, nvp, u, tile);
}

// Grouped packing of many fragments (one launch for a batch).  Job j owns the
// global tiles [first, next job's first); a workgroup finds the job of its
// tile by bisection (uniform).
struct PackJob {
  PackArgs p;
  double* sigma;  // row scales written by rowmax_group_kernel (p.sigma reads them)
  int64_t nvp;
  int64_t first;
  int* status;    // CHECK rowmax: bit 4 set on a NaN / Inf in the job's rows (or null)
};

// Padded virtual energies of many fragments: out[a] = ev[a] (a < nvirt), 1.0
// on padding rows (their B rows are zero); jobs own elements [first, next).
struct PadJob {
  const double* ev;
  double* out;
  int64_t nvirt;
  int64_t first;
};

template <class Job>
__device__ inline int pack_job_find(const Job* __restrict__ jobs, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The job of item t for items taken in increasing order: found by bisection
// once, then stepped forward.  (A bisection per item is a chain of dependent
// global loads, several microseconds per tile or row.)
template <class Job>
struct JobWalk {
  const Job* __restrict__ jobs;
  int n, j;
  int64_t next;  // first item of job j + 1
  __device__ JobWalk(const Job* __restrict__ js, int nj, int64_t t0) : jobs(js), n(nj) {
    j = pack_job_find(jobs, n, t0);
    next = j + 1 < n ? jobs[j + 1].first : INT64_MAX;
  }
  __device__ int at(int64_t t) {
    while (t >= next) {
      j++;
      next = j + 1 < n ? jobs[j + 1].first : INT64_MAX;
    }
    return j;
  }
};

// One tile per workgroup (grid-stride beyond the grid cap): many short-lived
// workgroups overlap their load phases better than long runs of tiles per
// workgroup (measured: fp64 1.38 against 1.61 ms at 256 DEC fragments).
template <int MODE>
__global__ __launch_bounds__(256) void pack_tile_group_kernel(const PackJob* __restrict__ jobs, int njobs,
                                                              int64_t total) {
  __shared__ double tile[TILE_ROWS * (PackT<MODE>::TQ + 1)];
  for (int64_t u = blockIdx.x; u < total; u += gridDim.x) {
    const int j = pack_job_find(jobs, njobs, u);
    pack_tile<MODE>(jobs[j].p, jobs[j].nvp, u - jobs[j].first, tile);
  }
}

// Row scales of the exact-slice modes for many fragments: one wave per row,
// each wave taking rows [w chunk, (w + 1) chunk) in order; jobs own rows
// [first, next first) (row = i * nvp + a).  CHECK: a row holding a NaN or
// Inf sets bit 4 of its job's status.
template <bool CHECK>
__global__ __launch_bounds__(256) void rowmax_group_kernel(const PackJob* __restrict__ jobs, int njobs, int64_t total,
                                                           int64_t chunk) {
  const int lane = threadIdx.x & 63;
  const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * chunk;
  const int64_t r1 = r0 + chunk < total ? r0 + chunk : total;
  if (r0 >= r1) return;
  JobWalk<PackJob> w(jobs, njobs, r0);
  for (int64_t row = r0; row < r1; row++) {
    const PackJob& J = jobs[w.at(row)];
    const int64_t r = row - J.first, i = r / J.nvp, a = r % J.nvp;
    double m = 0.0;
    if constexpr (CHECK) {
      double z = 0.0;
      if (a < J.p.nvirt) m = col_absmax_z(J.p.B + J.p.naux * (a + J.p.nvirt * i), J.p.naux, lane, z);
      if (__any(z != z) && lane == 0 && J.status) atomicOr(J.status, 4);
    } else {
      if (a < J.p.nvirt) m = col_absmax(J.p.B + J.p.naux * (a + J.p.nvirt * i), J.p.naux, lane);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (lane == 0) J.sigma[r] = m;
  }
}

__global__ __launch_bounds__(256) void pad_group_kernel(const PadJob* __restrict__ jobs, int njobs, int64_t total) {
  for (int64_t base = (int64_t)blockIdx.x * 256; base < total; base += (int64_t)gridDim.x * 256) {
    int j = pack_job_find(jobs, njobs, base);
    const int64_t t = base + threadIdx.x;
    if (t >= total) continue;
    while (j + 1 < njobs && jobs[j + 1].first <= t) j++;
    const int64_t a = t - jobs[j].first;
    jobs[j].out[a] = (a < jobs[j].nvirt) ? jobs[j].ev[a] : 1.0;
  }
}

}  // namespace

template <int MODE>
static void launch_pack_tiles(const PackArgs& p, int64_t nvp, hipStream_t s, unsigned max_blocks) {
  unsigned blocks = grid_blocks(pack_tiles<MODE>(p), 1);
  if (max_blocks && blocks > max_blocks) blocks = max_blocks;  // the kernel loops over its tiles
  hipLaunchKernelGGL(pack_tile_kernel<MODE>, dim3(blocks), dim3(256), 0, s, p, nvp);
}

void pack_B(Ctx* c, const double* dB, int64_t naux, int64_t nvirt, int64_t nocc, const PackGeom& g, void* dst,
            double* dscales, hipStream_t s, unsigned max_blocks) {
  ScopedTimer tm(c, RIMP2_K_PACK);
  PackArgs p{dB, naux, nvirt, nocc, g.nks, g.nrb, dscales, dst};
  switch (g.mode) {
    case RIMP2_MODE_FP64: launch_pack_tiles<RIMP2_MODE_FP64>(p, g.nvp, s, max_blocks); break;
    case RIMP2_MODE_FP32:
      if (g.rb == 16) launch_pack_tiles<PACK_F32R16>(p, g.nvp, s, max_blocks);
      else launch_pack_tiles<RIMP2_MODE_FP32>(p, g.nvp, s, max_blocks);
      break;
    case RIMP2_MODE_BF16X3: launch_pack_tiles<RIMP2_MODE_BF16X3>(p, g.nvp, s, max_blocks); break;
    case RIMP2_MODE_I8S4:
    case RIMP2_MODE_I8S5:
      hipLaunchKernelGGL(rowmax_kernel, dim3((unsigned)ceil_div(g.nvp * nocc, 4)), dim3(256), 0, s, dB, naux, nvirt,
                         nocc, g.nvp, dscales);
      RIMP2_HIP_CHECK(hipGetLastError());
      if (g.mode == RIMP2_MODE_I8S4) launch_pack_tiles<RIMP2_MODE_I8S4>(p, g.nvp, s, max_blocks);
      else launch_pack_tiles<RIMP2_MODE_I8S5>(p, g.nvp, s, max_blocks);
      break;
    default: fail(RIMP2_ERR_ARG, "unknown precision mode %d", g.mode);
  }
  RIMP2_HIP_CHECK(hipGetLastError());
}

void pack_slices6(hipStream_t s, const double* src, int64_t nq, int64_t nrows, int64_t norb, int64_t nks,
                  int64_t nrp, double* sigma, void* dst, unsigned max_blocks) {
  if (nks * 32 < nq || nrp < nrows || nrp % 32) fail(RIMP2_ERR_ARG, "six-slice packing geometry");
  hipLaunchKernelGGL(rowmax_kernel, dim3((unsigned)ceil_div(nrp * norb, 4)), dim3(256), 0, s, src, nq, nrows, norb, nrp,
                     sigma);
  RIMP2_HIP_CHECK(hipGetLastError());
  PackArgs p{src, nq, nrows, norb, nks, nrp / 32, sigma, dst};
  launch_pack_tiles<PACK_I8S6>(p, nrp, s, max_blocks);
  RIMP2_HIP_CHECK(hipGetLastError());
}

// Workgroups of the grouped row-maximum launch: each wave takes a contiguous
// run of rows.
constexpr int64_t kGroupBlocks = 2048;

// tiles of one item's packing (the packer of mode and fragment shape rb)
static int64_t tiles_of(int mode, int rb, const PackArgs& p) {
  switch (mode) {
    case RIMP2_MODE_FP64: return pack_tiles<RIMP2_MODE_FP64>(p);
    case RIMP2_MODE_FP32: return rb == 16 ? pack_tiles<PACK_F32R16>(p) : pack_tiles<RIMP2_MODE_FP32>(p);
    case RIMP2_MODE_BF16X3: return pack_tiles<RIMP2_MODE_BF16X3>(p);
    case RIMP2_MODE_I8S4: return pack_tiles<RIMP2_MODE_I8S4>(p);
    case RIMP2_MODE_I8S5: return pack_tiles<RIMP2_MODE_I8S5>(p);
  }
  fail(RIMP2_ERR_ARG, "unknown precision mode %d", mode);
  return 0;
}

void pack_B_group(const std::vector<PackItem>& items, int mode, hipStream_t s, DevBuf& jobs,
                  std::vector<unsigned char>& blob) {
  if (items.empty()) return;
  std::vector<PackJob> pj;
  std::vector<PadJob> dj;
  int64_t ne = 0, nr = 0, np = 0;
  for (const PackItem& it : items) {
    if (it.g.mode != mode || it.g.rb != items[0].g.rb)
      fail(RIMP2_ERR_ARG, "grouped packing: geometries of another mode or tile");
    PackJob j;
    j.p = PackArgs{it.B, it.naux, it.nvirt, it.nocc, it.g.nks, it.g.nrb, it.sigma, it.dst};
    j.sigma = it.sigma;
    j.nvp = it.g.nvp;
    j.first = ne;
    j.status = nullptr;
    pj.push_back(j);
    ne += tiles_of(mode, it.g.rb, j.p);
    dj.push_back(PadJob{it.ev, it.ev_padded, it.nvirt, np});
    np += it.g.nvp;
  }
  // rowmax reuses the pack jobs with first counted in rows (i * nvp + a)
  std::vector<PackJob> rj = pj;
  for (size_t k = 0; k < rj.size(); k++) {
    rj[k].first = nr;
    nr += items[k].nocc * items[k].g.nvp;
  }
  blob.clear();
  auto put = [&](const void* p, size_t bytes) {
    const size_t off = (size_t)round_up((int64_t)blob.size(), 256);
    blob.resize(off + bytes);
    memcpy(blob.data() + off, p, bytes);
    return off;
  };
  const size_t opj = put(pj.data(), pj.size() * sizeof(PackJob));
  const size_t orj = put(rj.data(), rj.size() * sizeof(PackJob));
  const size_t odj = put(dj.data(), dj.size() * sizeof(PadJob));
  unsigned char* d = static_cast<unsigned char*>(jobs.ensure(blob.size()));
  RIMP2_HIP_CHECK(hipMemcpyAsync(d, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
  const int nj = (int)items.size();
  if (mode_slices(mode)) {
    const int64_t rwaves = std::min<int64_t>(nr, kGroupBlocks * 4), rchunk = ceil_div(nr, rwaves);
    hipLaunchKernelGGL(rowmax_group_kernel<false>, dim3((unsigned)ceil_div(rwaves, 4)), dim3(256), 0, s,
                       (const PackJob*)(d + orj), nj, nr, rchunk);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  const dim3 gb(grid_blocks(ne, 1));  // one workgroup per tile (capped; the kernel loops)
  const PackJob* jp = (const PackJob*)(d + opj);
  switch (mode) {
    case RIMP2_MODE_FP64: hipLaunchKernelGGL(pack_tile_group_kernel<RIMP2_MODE_FP64>, gb, dim3(256), 0, s, jp, nj, ne); break;
    case RIMP2_MODE_FP32:
      if (items[0].g.rb == 16) hipLaunchKernelGGL(pack_tile_group_kernel<PACK_F32R16>, gb, dim3(256), 0, s, jp, nj, ne);
      else hipLaunchKernelGGL(pack_tile_group_kernel<RIMP2_MODE_FP32>, gb, dim3(256), 0, s, jp, nj, ne);
      break;
    case RIMP2_MODE_BF16X3:
      hipLaunchKernelGGL(pack_tile_group_kernel<RIMP2_MODE_BF16X3>, gb, dim3(256), 0, s, jp, nj, ne);
      break;
    case RIMP2_MODE_I8S4: hipLaunchKernelGGL(pack_tile_group_kernel<RIMP2_MODE_I8S4>, gb, dim3(256), 0, s, jp, nj, ne); break;
    case RIMP2_MODE_I8S5: hipLaunchKernelGGL(pack_tile_group_kernel<RIMP2_MODE_I8S5>, gb, dim3(256), 0, s, jp, nj, ne); break;
    default: fail(RIMP2_ERR_ARG, "unknown precision mode %d", mode);
  }
  RIMP2_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(pad_group_kernel, dim3(grid_blocks(np, 256)), dim3(256), 0, s, (const PadJob*)(d + odj), nj, np);
  RIMP2_HIP_CHECK(hipGetLastError());
}

void pack_slices6_group(const std::vector<Slice6Item>& items, hipStream_t s, DevBuf& jobs,
                        std::vector<unsigned char>& blob) {
  blob.clear();
  if (items.empty()) return;
  std::vector<PackJob> pj, rj;
  int64_t ne = 0, nr = 0;
  for (const Slice6Item& it : items) {
    if (it.nks * 32 < it.nq || it.nrp < it.nrows || it.nrp % 32 || it.nks % 4)
      fail(RIMP2_ERR_ARG, "grouped six-slice packing geometry");
    PackJob j;
    j.p = PackArgs{it.src, it.nq, it.nrows, it.norb, it.nks, it.nrp / 32, it.sigma, it.dst};
    j.sigma = it.sigma;
    j.nvp = it.nrp;
    j.status = it.status;
    j.first = ne;
    pj.push_back(j);
    ne += pack_tiles<PACK_I8S6>(j.p);
    j.first = nr;
    rj.push_back(j);
    nr += it.norb * it.nrp;
  }
  auto put = [&](const void* p, size_t bytes) {
    const size_t off = (size_t)round_up((int64_t)blob.size(), 256);
    blob.resize(off + bytes);
    memcpy(blob.data() + off, p, bytes);
    return off;
  };
  const size_t opj = put(pj.data(), pj.size() * sizeof(PackJob));
  const size_t orj = put(rj.data(), rj.size() * sizeof(PackJob));
  unsigned char* d = static_cast<unsigned char*>(jobs.ensure(blob.size()));
  RIMP2_HIP_CHECK(hipMemcpyAsync(d, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
  const int nj = (int)items.size();
  const int64_t rwaves = std::min<int64_t>(nr, kGroupBlocks * 4), rchunk = ceil_div(nr, rwaves);
  hipLaunchKernelGGL(rowmax_group_kernel<true>, dim3((unsigned)ceil_div(rwaves, 4)), dim3(256), 0, s,
                     (const PackJob*)(d + orj), nj, nr, rchunk);
  RIMP2_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(pack_tile_group_kernel<PACK_I8S6>, dim3(grid_blocks(ne, 1)), dim3(256), 0, s,
                     (const PackJob*)(d + opj), nj, ne);
  RIMP2_HIP_CHECK(hipGetLastError());
}

}  // namespace rimp2
