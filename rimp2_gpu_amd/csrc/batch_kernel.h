// The batch's persistent pair kernel (shared by batch.hip, which launches the
// one-tile forms, and batch_mixed.hip, which launches the forms that carry
// both a 32- and a 64-wide tile path).  The two groups live in separate
// translation units: in one unit, a second kernel calling the same pair_tile
// specialization fails overload resolution in the host compilation pass
// (hipcc of ROCm 7.2), where the device-only body of pair_tile is invalid.
#pragma once
#include "pair_core.h"

namespace rimp2 {

struct BatchFrag {
  pairk::FragView f;
  int64_t tile;  // pair tile of this fragment (virtual block width)
  int64_t nocc, nb, ntp, npairs;
  int64_t wu_begin;  // first global work unit (launch order)
  int64_t pair_off;  // offset of this fragment's [nocc, nocc] pair matrix
  int64_t orig;      // index of the fragment in the caller's list
};

struct BatchParams {
  const BatchFrag* frags;
  int64_t nfrag, total;
  int64_t range[9];  // XCD x owns work units [range[x], range[x+1])
  double* partials;
};

// Launches the mixed-tile kernel of `mode` (fp64 or fp32); batch_mixed.hip.
void launch_batch_mixed(int mode, hipStream_t s, const BatchParams& p, int ncu);

namespace {

using namespace pairk;

template <int A, int B>
constexpr int cmax() { return A > B ? A : B; }
template <int A, int B>
constexpr int cmin() { return A < B ? A : B; }

// Tile paths TS (fragments with tile == TS) and TL (the others); TS == TL
// compiles one path.  The stage buffers are the larger of the two paths'.
template <class PolS, int TS, class PolL, int TL>
__global__ __launch_bounds__(PAIR_THREADS)
__attribute__((amdgpu_waves_per_eu(cmin<pair_waves_per_eu<PolS, TS>(), pair_waves_per_eu<PolL, TL>()>())))
void batch_pair_kernel(BatchParams p) {
  constexpr int SB = cmax<pair_stage_bytes<PolS, TS>(), pair_stage_bytes<PolL, TL>()>();
  __shared__ __align__(16) unsigned char rimp2_buf0[SB];
  __shared__ __align__(16) unsigned char rimp2_buf1[SB];
  __shared__ __align__(16) unsigned char rimp2_buf2[SB];
  __shared__ __align__(16) unsigned char rimp2_buf3[SB];
  const int64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
  const int64_t end = p.range[xcd + 1];
  int64_t wu = p.range[xcd] + slot;
  if (wu >= end) return;
  // fragment holding the first unit: last fragment with wu_begin <= wu
  int64_t lo = 0, hi = p.nfrag - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (uni(gload(&p.frags[mid].wu_begin)) <= wu) lo = mid;
    else hi = mid - 1;
  }
  // The units of this workgroup increase, so the fragment index only moves
  // forward: the fragment's fields are read again only when wu passes the
  // next fragment's first unit.
  FragView f;
  int64_t nb = 0, ntp = 1, begin = 0, next = 0, tile = TL;
  auto load_frag = [&](int64_t k) {
    const BatchFrag* q = p.frags + k;
    f.packed = uni(gload(&q->f.packed));
    f.sigma = uni(gload(&q->f.sigma));
    f.eps_o = uni(gload(&q->f.eps_o));
    f.eps_v = uni(gload(&q->f.eps_v));
    f.nvp = uni(gload(&q->f.nvp));
    f.nks = uni(gload(&q->f.nks));
    f.nrb = uni(gload(&q->f.nrb));
    f.nvirt = uni(gload(&q->f.nvirt));
    tile = uni(gload(&q->tile));
    nb = uni(gload(&q->nb));
    ntp = uni(gload(&q->ntp));
    begin = uni(gload(&q->wu_begin));
    next = k + 1 < p.nfrag ? uni(gload(&p.frags[k + 1].wu_begin)) : INT64_MAX;
  };
  load_frag(lo);
  for (; wu < end; wu += nslot) {
    if (wu >= next) {
      while (lo + 1 < p.nfrag && uni(gload(&p.frags[lo + 1].wu_begin)) <= wu) lo++;
      load_frag(lo);
    }
    const int64_t local = wu - begin;
    const int64_t pidx = local / ntp, tp = local % ntp;
    int64_t i, j, A, Bk;
    tri_decode(pidx, i, j);
    i = uni(i);
    j = uni(j);
    double e;
    if (TS != TL && tile == TS) {
      const int kind = uni_int(unit_decode(tp, nb, packs_diagonal<PolS, TS>(), A, Bk));
      e = pair_tile<PolS, TS, false, true>(f, i, j, uni(A), uni(Bk), kind, rimp2_buf0, rimp2_buf1, rimp2_buf2,
                                           rimp2_buf3);
    } else {
      const int kind = uni_int(unit_decode(tp, nb, packs_diagonal<PolL, TL>(), A, Bk));
      e = pair_tile<PolL, TL, false, true>(f, i, j, uni(A), uni(Bk), kind, rimp2_buf0, rimp2_buf1, rimp2_buf2,
                                           rimp2_buf3);
    }
    if (threadIdx.x == 0) p.partials[wu] = e;
  }
}

template <class PolS, int TS, class PolL, int TL>
void launch_batch_t(hipStream_t s, const BatchParams& p, int ncu) {
  int blocks = 0;
  RIMP2_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batch_pair_kernel<PolS, TS, PolL, TL>,
                                                               PAIR_THREADS, 0));
  const int grid = 8 * ((ncu * (blocks < 1 ? 1 : blocks) + 7) / 8);  // whole XCD slots
  hipLaunchKernelGGL((batch_pair_kernel<PolS, TS, PolL, TL>), dim3(grid), dim3(PAIR_THREADS), 0, s, p);
  RIMP2_HIP_CHECK(hipGetLastError());
}

}  // namespace
}  // namespace rimp2
