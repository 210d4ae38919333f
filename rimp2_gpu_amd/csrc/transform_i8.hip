// Metric transform B = L^{-1} X on int8 MFMA, for the exact-slice modes.
//
// Both operands are split into six int8 slices per element, each column
// scaled by its own largest |value|: X by (auxiliary) column (i, a) over P,
// L^{-T} by column Q over P (a row of L^{-1}).  Every slice product with
// s + t <= 5 (21 of the 36) accumulates exactly in int32, one accumulator per
// diagonal d = s + t (at most 6 products a diagonal: |acc| <= 6 x 127^2 x naux
// < 2^31 for naux <= 16384, the bound every exact-slice entry point checks);
// the dropped products sit below 254^{-6} of the scale.
// The FP64 epilogue forms sum_d acc_d 254^{-d} / 127^2 x sigma_Q x sigma_ia.
// With five slices the transform's own error in E would be ~10x the pair
// kernel's; with six it is ~1/30 of it (numpy model at 2048/512/8 and
// 512/128/16), and B's relative error ~4e-13.
//
// One workgroup of four waves per 64 (virtual rows a of one occupied i) x 64
// (auxiliary Q) tile of B^T: X is the MFMA row operand and L^{-T} the column
// operand, so each lane's output runs along Q and the stores of B (Q fastest)
// are coalesced.  Wave (wr, wc) holds the 32 x 32 block (a rows wr, Q cols
// wc).  A 32-deep k-step (one stage) brings 12 KB of X and 12 KB of L^{-T}
// into LDS through buffer_load ... lds, six 1-KB transfers per wave, in a
// three-stage ring; the K loop of a Q tile stops at its last row (L^{-1} is
// lower triangular; at least three stages, where the operands are zero).
#include <algorithm>

#include "pair_core.h"

namespace rimp2 {
namespace {

using pairk::lds_ptr_t;
using pairk::pipeline_barrier;
using pairk::slab_rsrc;
using Pol = pairk::PolI8<6, 5>;

constexpr int TI_THREADS = 256;
constexpr int TI_NW = TI_THREADS / 64;  // waves
constexpr int TI_SL = 6;                // slices
constexpr int TI_FB = 1024;             // bytes per fragment (64 lanes x 16)
constexpr int TI_CHUNK = 2 * TI_SL * TI_FB;  // one operand's two row blocks per k-step: 12 KB
constexpr int TI_SB = 2 * TI_CHUNK;          // stage: X then L^{-T}, 24 KB
constexpr int TI_NI = TI_SB / (TI_NW * 1024);  // 1-KB transfers per wave per stage
static_assert(TI_NI == 6, "transfers 0-2 of a wave carry X, 3-5 carry L^-T");

struct TransformI8Args {
  const unsigned char* Xp;  // X slices: [orb][ks][rb (32 rows of a)][slice][lane x 16 B]
  const unsigned char* Ap;  // L^{-T} slices: [ks][rb (32 rows of Q)][slice][lane x 16 B]
  const double* sX;         // [orb][nvp] column scales of X
  const double* sA;         // [nqp] column scales of L^{-T}
  double* B;                // [n, nvirt, nocc] out; orbital i0 + orb
  int64_t n, nvirt, nvp, nqp, nks;
  int64_t i0, norb;
  int64_t MT, NT;           // Q tiles (64), virtual tiles (64) per orbital
};

// Tile u of p (all 256 threads; the arguments and u wave-uniform).  COLS: tiles
// column by column (orbital, virtual tile), the Q tiles of a column together,
// longest K loop first: they read the same X slab, which the launch keeps on
// one XCD (xcd_unit), so what each column re-reads is L^{-T}, small enough for
// the Infinity Cache, where Q tile by Q tile every X slab came from HBM once
// per Q tile (64/512/2048: MFMA busy 0.56 -> 0.71, 1.39 -> 1.22 ms).  Else Q
// tile by Q tile (longest first),
// the columns inside: the grouped launch of many small fragments, whose X
// slabs stay in the caches, measured 0.4 ms faster per 256-fragment prepare
// in that order.
template <bool COLS>
__device__ __attribute__((always_inline)) inline void transform_i8_tile(const TransformI8Args& p, int64_t u,
                                                                        unsigned char* __restrict__ buf0,
                                                                        unsigned char* __restrict__ buf1,
                                                                        unsigned char* __restrict__ buf2) {
  const int64_t ncol = p.norb * p.NT;
  const int64_t col = COLS ? u / p.MT : u % ncol, mt = p.MT - 1 - (COLS ? u % p.MT : u / ncol);
  const int64_t io = col / p.NT, nt = col % p.NT;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 1, wc = wave & 1;
  // K stages: up to the tile's last Q row (at least the three the prologue issues)
  int64_t nst = 2 * (mt + 1) > 3 ? 2 * (mt + 1) : 3;
  if (nst > p.nks) nst = p.nks;
  const size_t x_ks = (size_t)(p.nvp / 32) * TI_SL * TI_FB, a_ks = (size_t)(p.nqp / 32) * TI_SL * TI_FB;
  const size_t x_orb = (size_t)p.nks * x_ks;
  const unsigned char* xs = p.Xp + (size_t)io * x_orb + (size_t)(2 * nt) * TI_SL * TI_FB;
  const unsigned char* as = p.Ap + (size_t)(2 * mt) * TI_SL * TI_FB;
  const __amdgpu_buffer_rsrc_t rx = slab_rsrc(xs, x_orb - (size_t)(2 * nt) * TI_SL * TI_FB);
  const __amdgpu_buffer_rsrc_t ra = slab_rsrc(as, (size_t)p.nks * a_ks - (size_t)(2 * mt) * TI_SL * TI_FB);
  // transfer n of this wave: KB (4 n + wave) of the stage; 0-11 X, 12-23 L^{-T}
  uint32_t voff[TI_NI];
#pragma unroll
  for (int k = 0; k < TI_NI; k++) voff[k] = (uint32_t)((((4 * k + wave) % 12) * 1024) + lane * 16);
  auto issue = [&](int64_t st, unsigned char* dst) {
    const uint32_t sx = (uint32_t)((size_t)st * x_ks), sa = (uint32_t)((size_t)st * a_ks);
#pragma unroll
    for (int k = 0; k < TI_NI; k++)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(k < 3 ? rx : ra, (lds_ptr_t)(dst + (4 * k + wave) * 1024), 16,
                                               voff[k], k < 3 ? sx : sa, 0, 0);
  };
  struct Ops {
    Pol::frag_t r[TI_SL], c[TI_SL];
  };
  auto load = [&](const unsigned char* stg, Ops& o) {
#pragma unroll
    for (int s = 0; s < TI_SL; s++) {
      o.r[s] = *reinterpret_cast<const Pol::frag_t*>(stg + (wr * TI_SL + s) * TI_FB + lane * 16);
      o.c[s] = *reinterpret_cast<const Pol::frag_t*>(stg + TI_CHUNK + (wc * TI_SL + s) * TI_FB + lane * 16);
    }
  };
  Pol::acc_t acc[Pol::ND];
#pragma unroll
  for (int d = 0; d < Pol::ND; d++) Pol::zero(acc[d]);
  static_assert(Pol::MMA_PER_CALL == 21, "six slices, diagonals up to 5");
  // step st: stage st + 1 has landed (stage st + 2 may be in flight), every
  // wave has read stage st into registers; the freed buffer takes stage st + 3
  // (re-fetches of the last stage past the end keep the counts uniform)
  auto step = [&](int64_t st, Ops& cur, Ops& nxt, unsigned char* freed, const unsigned char* next) {
    RIMP2_WAIT_VMCNT_LGKM0(TI_NI);
    pipeline_barrier();
    issue(st + 3 < nst ? st + 3 : nst - 1, freed);
    load(next, nxt);
    Pol::mma(acc, cur.r, cur.c);
    // 21 MFMAs, 12 LDS reads and 6 DMAs interleaved
#pragma unroll
    for (int g = 0; g < TI_NI; g++) {
      __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x0100, 2, 0);  // LDS reads
      __builtin_amdgcn_sched_group_barrier(0x0008, 2, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);  // VMEM read (the DMA)
    }
    __builtin_amdgcn_sched_group_barrier(0x0008, Pol::MMA_PER_CALL - 3 * TI_NI, 0);
  };
  Ops ra_, rb_;
  issue(0, buf0);
  issue(1, buf1);
  issue(2, buf2);
  RIMP2_WAIT_VMCNT(2 * TI_NI);
  pipeline_barrier();
  load(buf0, ra_);
  // three buffers and two register sets: the pattern repeats every six steps;
  // the last 0-5 steps run straight-line
  int64_t st = 0;
  for (; st + 6 <= nst; st += 6) {
    step(st, ra_, rb_, buf0, buf1);
    step(st + 1, rb_, ra_, buf1, buf2);
    step(st + 2, ra_, rb_, buf2, buf0);
    step(st + 3, rb_, ra_, buf0, buf1);
    step(st + 4, ra_, rb_, buf1, buf2);
    step(st + 5, rb_, ra_, buf2, buf0);
  }
  const int64_t rem = nst - st;
  if (rem > 0) step(st, ra_, rb_, buf0, buf1);
  if (rem > 1) step(st + 1, rb_, ra_, buf1, buf2);
  if (rem > 2) step(st + 2, ra_, rb_, buf2, buf0);
  if (rem > 3) step(st + 3, rb_, ra_, buf0, buf1);
  if (rem > 4) step(st + 4, ra_, rb_, buf1, buf2);
  RIMP2_WAIT_VMCNT_LGKM0(0);  // trailing re-fetches are done before the workgroup ends
  // epilogue: element (a row, Q col) of the wave's block
  const int64_t i = p.i0 + io;
  const int64_t q = mt * 64 + wc * 32 + Pol::col(lane, 0);
  const double sq = q < p.n ? p.sA[q] : 0.0;
  double* Bi = p.B + (size_t)p.n * (size_t)p.nvirt * (size_t)i;
#pragma unroll
  for (int r = 0; r < Pol::NE; r++) {
    const int64_t a = nt * 64 + wr * 32 + Pol::row(lane, r);
    if (a < p.nvirt && q < p.n) Bi[q + p.n * a] = Pol::value(acc, r) * (p.sX[io * p.nvp + a] * sq);
  }
}

// Three 24-KB stage buffers: two workgroups per CU (144 of its 160 KB), so one
// workgroup's pipeline fill and epilogue run under the other's MFMAs.
// Tile of workgroup slot s on XCD x when each XCD takes a contiguous range of
// the total tiles (the hardware places workgroup b on XCD b % 8); >= total for
// the spare slots of the last range.
__device__ inline int64_t xcd_unit(int64_t x, int64_t s, int64_t total) { return x * ((total + 7) / 8) + s; }

#define RIMP2_TI_LDS                                   \
  __shared__ __align__(16) unsigned char buf0[TI_SB]; \
  __shared__ __align__(16) unsigned char buf1[TI_SB]; \
  __shared__ __align__(16) unsigned char buf2[TI_SB]

__global__ __launch_bounds__(TI_THREADS, 2) void transform_i8_kernel(TransformI8Args p, int64_t total) {
  RIMP2_TI_LDS;
  const int64_t u = xcd_unit(blockIdx.x % 8, blockIdx.x / 8, total);
  if (u < total) transform_i8_tile<true>(p, u, buf0, buf1, buf2);
}

// Grouped form: job j owns the tiles [first, next job's first).
struct TransformI8Job {
  TransformI8Args a;
  int64_t first;
};

template <class Job>
__device__ inline int find_job(const Job* __restrict__ jobs, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Tiles of all jobs, grid-stride beyond the grid; a barrier between tiles keeps
// every wave's last reads of the ring ahead of the next tile's DMA.
__global__ __launch_bounds__(TI_THREADS, 2) void transform_i8_group_kernel(const TransformI8Job* __restrict__ jobs,
                                                                        int njobs, int64_t total) {
  RIMP2_TI_LDS;
  using pairk::uni;
  for (int64_t u = blockIdx.x; u < total; u += gridDim.x) {
    const int j = pairk::uni_int(find_job(jobs, njobs, u));
    const TransformI8Args& J = jobs[j].a;
    TransformI8Args a;  // wave-uniform copies: the buffer resources stay in scalar registers
    a.Xp = uni(J.Xp);
    a.Ap = uni(J.Ap);
    a.sX = uni(J.sX);
    a.sA = uni(J.sA);
    a.B = uni(J.B);
    a.n = uni(J.n);
    a.nvirt = uni(J.nvirt);
    a.nvp = uni(J.nvp);
    a.nqp = uni(J.nqp);
    a.nks = uni(J.nks);
    a.i0 = 0;
    a.norb = uni(J.norb);
    a.MT = uni(J.MT);
    a.NT = uni(J.NT);
    transform_i8_tile<false>(a, u - uni(jobs[j].first), buf0, buf1, buf2);
    pipeline_barrier();
  }
}

// L^{-T}[P + n Q] = L^{-1}[Q + n P] by 64 x 64 tiles through LDS (tile b).
__device__ inline void transpose_tile(const double* __restrict__ Li, int64_t n, double* __restrict__ Lt, int64_t b,
                                      double (&t)[64][65]) {
  const int64_t tiles = (n + 63) / 64;
  const int64_t bq = b % tiles, bp = b / tiles;
  const int tid = threadIdx.x;
#pragma unroll 4
  for (int r = 0; r < 16; r++) {
    const int c = (tid >> 6) + 4 * r;  // column P of L^{-1}
    const int64_t Q = bq * 64 + (tid & 63), P = bp * 64 + c;
    t[c][tid & 63] = (Q < n && P < n) ? Li[Q + n * P] : 0.0;
  }
  __syncthreads();
#pragma unroll 4
  for (int r = 0; r < 16; r++) {
    const int c = (tid >> 6) + 4 * r;  // column Q of L^{-T}
    const int64_t P = bp * 64 + (tid & 63), Q = bq * 64 + c;
    if (P < n && Q < n) Lt[P + n * Q] = t[tid & 63][c];
  }
}

__global__ __launch_bounds__(256) void transpose_kernel(const double* __restrict__ Li, int64_t n,
                                                        double* __restrict__ Lt) {
  __shared__ double t[64][65];
  transpose_tile(Li, n, Lt, blockIdx.x, t);
}

struct TransposeJob {
  const double* Li;
  double* Lt;
  int64_t n, first;
};

__global__ __launch_bounds__(256) void transpose_group_kernel(const TransposeJob* __restrict__ jobs, int njobs,
                                                              int64_t total) {
  __shared__ double t[64][65];
  for (int64_t u = blockIdx.x; u < total; u += gridDim.x) {
    const TransposeJob& J = jobs[find_job(jobs, njobs, u)];
    transpose_tile(J.Li, J.n, J.Lt, u - J.first, t);
    __syncthreads();  // t is reused by the next tile
  }
}

// Geometry of the six-slice operands.
struct I8Geom {
  int64_t nks, nqp, nvp;
};
I8Geom i8_geom(int64_t n, int64_t nvirt) {
  return I8Geom{round_up(n, (int64_t)128) / 32, round_up(n, (int64_t)64), round_up(nvirt, (int64_t)64)};
}
size_t x_slices_bytes_per_orb(const I8Geom& g) { return (size_t)g.nks * (size_t)(g.nvp / 32) * TI_SL * TI_FB; }

}  // namespace

size_t metric_i8_x_bytes(int64_t n, int64_t nvirt, int64_t nocc) {
  return x_slices_bytes_per_orb(i8_geom(n, nvirt)) * (size_t)nocc;
}

size_t metric_i8_linv_bytes(int64_t n) {
  const I8Geom g = i8_geom(n, 1);
  return (size_t)g.nks * (size_t)(g.nqp / 32) * TI_SL * TI_FB;
}

void metric_slice_linv(Ctx* c, hipStream_t s, int64_t n) {
  ScopedTimer t(s == c->stream ? c : nullptr, RIMP2_K_TRANSFORM);
  const I8Geom g = i8_geom(n, 1);
  if ((size_t)g.nks * (size_t)(g.nqp / 32) * TI_SL * TI_FB > 0xFFFFFFFFull)
    fail(RIMP2_ERR_ARG, "int8 transform: L^-1 slices exceed 4 GiB (naux=%lld)", (long long)n);
  double* Lt = static_cast<double*>(c->work.ensure((size_t)(n * n) * sizeof(double)));
  const unsigned tiles = (unsigned)ceil_div(n, (int64_t)64);
  hipLaunchKernelGGL(transpose_kernel, dim3(tiles * tiles), dim3(256), 0, s, c->Linv.as<double>(), n, Lt);
  RIMP2_HIP_CHECK(hipGetLastError());
  double* sA = static_cast<double*>(c->lt_sigma.ensure((size_t)g.nqp * sizeof(double)));
  void* Ap = c->lt_slices.ensure(metric_i8_linv_bytes(n));
  pack_slices6(s, Lt, n, n, 1, g.nks, g.nqp, sA, Ap);
}

void metric_reserve_x_i8(Ctx* c, int64_t n, int64_t nvirt, int64_t nocc) {
  const I8Geom g = i8_geom(n, nvirt);
  if (x_slices_bytes_per_orb(g) > 0xFFFFFFFFull)
    fail(RIMP2_ERR_ARG, "int8 transform: slices of one occupied orbital exceed 4 GiB (naux=%lld nvirt=%lld)",
         (long long)n, (long long)nvirt);
  c->x_packed.ensure(metric_i8_x_bytes(n, nvirt, nocc));
  c->x_sigma.ensure((size_t)(g.nvp * nocc) * sizeof(double));
}

void metric_slice_x_range(Ctx* c, hipStream_t s, int64_t n, int64_t nvirt, int64_t i0, int64_t ni,
                          const double* dXi) {
  const I8Geom g = i8_geom(n, nvirt);
  unsigned char* Xp = c->x_packed.as<unsigned char>() + x_slices_bytes_per_orb(g) * (size_t)i0;
  double* sX = c->x_sigma.as<double>() + g.nvp * i0;
  // on a side stream a capped grid, as the FP64 X packing: it runs beside the Cholesky
  pack_slices6(s, dXi, n, nvirt, ni, g.nks, g.nvp, sX, Xp, s == c->stream ? 0u : 512u);
}

void metric_transform_i8_range(Ctx* c, int64_t n, int64_t nvirt, int64_t i0, int64_t ni, double* dB) {
  if (ni <= 0) return;
  ScopedTimer t(c, RIMP2_K_TRANSFORM);
  const I8Geom g = i8_geom(n, nvirt);
  TransformI8Args a;
  a.Ap = c->lt_slices.as<unsigned char>();
  a.sA = c->lt_sigma.as<double>();
  a.B = dB;
  a.n = n;
  a.nvirt = nvirt;
  a.nvp = g.nvp;
  a.nqp = g.nqp;
  a.nks = g.nks;
  a.MT = g.nqp / 64;
  a.NT = g.nvp / 64;
  // one launch per run of orbitals whose tiles fit the grid cap (8 x the
  // tiles of each XCD's range, rounded up)
  const int64_t per = a.MT * a.NT, step = std::max<int64_t>(1, (kMaxGridBlocks - 8) / per);
  for (int64_t o = 0; o < ni; o += step) {
    a.i0 = i0 + o;
    a.norb = std::min(step, ni - o);
    a.Xp = c->x_packed.as<unsigned char>() + x_slices_bytes_per_orb(g) * (size_t)a.i0;
    a.sX = c->x_sigma.as<double>() + g.nvp * a.i0;
    const int64_t tiles = per * a.norb;
    hipLaunchKernelGGL(transform_i8_kernel, dim3((unsigned)(8 * ceil_div(tiles, (int64_t)8))), dim3(TI_THREADS), 0,
                       c->stream, a, tiles);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
}

void transform_i8_group(MetricGroup& G, hipStream_t s, const std::vector<I8GroupItem>& items) {
  if (items.empty()) return;
  // arena offsets (256-byte aligned): L^-T slices, X slices, and the scales
  std::vector<size_t> lo(items.size()), xo(items.size()), so(items.size());
  size_t ln = 0, xn = 0, sn = 0;
  auto take = [](size_t& at, size_t bytes) {
    const size_t o = at;
    at += (size_t)round_up((int64_t)bytes, (int64_t)256);
    return o;
  };
  for (size_t k = 0; k < items.size(); k++) {
    const I8GroupItem& it = items[k];
    const I8Geom g = i8_geom(it.n, it.nvirt);
    if ((size_t)g.nks * (size_t)(g.nqp / 32) * TI_SL * TI_FB > 0xFFFFFFFFull || x_slices_bytes_per_orb(g) > 0xFFFFFFFFull)
      fail(RIMP2_ERR_ARG, "int8 transform: slices exceed 4 GiB per slab (naux=%lld nvirt=%lld)", (long long)it.n,
           (long long)it.nvirt);
    lo[k] = take(ln, metric_i8_linv_bytes(it.n));
    xo[k] = take(xn, metric_i8_x_bytes(it.n, it.nvirt, it.nocc));
    so[k] = take(sn, (size_t)(g.nqp + g.nvp * it.nocc) * sizeof(double));
  }
  unsigned char* Ls = static_cast<unsigned char*>(G.Ls.ensure(ln));
  unsigned char* Xs = static_cast<unsigned char*>(G.Xs.ensure(xn));
  unsigned char* sg = static_cast<unsigned char*>(G.sig.ensure(sn));
  std::vector<TransposeJob> tj;
  std::vector<Slice6Item> sl;
  std::vector<TransformI8Job> tr;
  int64_t nt = 0, nu = 0;
  for (size_t k = 0; k < items.size(); k++) {
    const I8GroupItem& it = items[k];
    const I8Geom g = i8_geom(it.n, it.nvirt);
    double* sA = reinterpret_cast<double*>(sg + so[k]);
    double* sX = sA + g.nqp;
    tj.push_back(TransposeJob{it.Li, it.T, it.n, nt});
    nt += ceil_div(it.n, (int64_t)64) * ceil_div(it.n, (int64_t)64);
    sl.push_back(Slice6Item{it.T, it.n, it.n, 1, g.nks, g.nqp, sA, Ls + lo[k], nullptr});
    sl.push_back(Slice6Item{it.X, it.n, it.nvirt, it.nocc, g.nks, g.nvp, sX, Xs + xo[k], it.status});
    TransformI8Job j;
    j.a.Xp = Xs + xo[k];
    j.a.Ap = Ls + lo[k];
    j.a.sX = sX;
    j.a.sA = sA;
    j.a.B = it.B;
    j.a.n = it.n;
    j.a.nvirt = it.nvirt;
    j.a.nvp = g.nvp;
    j.a.nqp = g.nqp;
    j.a.nks = g.nks;
    j.a.i0 = 0;
    j.a.norb = it.nocc;
    j.a.MT = g.nqp / 64;
    j.a.NT = g.nvp / 64;
    j.first = nu;
    tr.push_back(j);
    nu += j.a.MT * j.a.NT * it.nocc;
  }
  G.iblob.clear();
  auto put = [&](const void* p, size_t bytes) {
    const size_t off = (size_t)round_up((int64_t)G.iblob.size(), (int64_t)256);
    G.iblob.resize(off + bytes);
    memcpy(G.iblob.data() + off, p, bytes);
    return off;
  };
  const size_t otj = put(tj.data(), tj.size() * sizeof(TransposeJob));
  const size_t otr = put(tr.data(), tr.size() * sizeof(TransformI8Job));
  unsigned char* d = static_cast<unsigned char*>(G.ijobs.ensure(G.iblob.size()));
  RIMP2_HIP_CHECK(hipMemcpyAsync(d, G.iblob.data(), G.iblob.size(), hipMemcpyHostToDevice, s));
  const int nj = (int)items.size();
  hipLaunchKernelGGL(transpose_group_kernel, dim3(grid_blocks(nt, 1)), dim3(256), 0, s,
                     (const TransposeJob*)(d + otj), nj, nt);
  RIMP2_HIP_CHECK(hipGetLastError());
  pack_slices6_group(sl, s, G.sjobs, G.sblob);
  hipLaunchKernelGGL(transform_i8_group_kernel, dim3(grid_blocks(nu, 1)), dim3(TI_THREADS), 0, s,
                     (const TransformI8Job*)(d + otr), nj, nu);
  RIMP2_HIP_CHECK(hipGetLastError());
}

}  // namespace rimp2
