// Batched DEC fragments: many fragments of uneven (nocc, nvirt, naux) through
// ONE persistent pair launch captured in a hipGraph.
//
// The reference processes its fragments one after another (`do ifile=1,2`,
// RIMP2.F90:36) with two OpenACC queues per fragment (RIMP2.GPU.F90:64-78).
// Here every fragment is factored, transformed and packed once
// (rimp2_batch_prepare); the pair work of all fragments is then one global
// work list: work unit = (fragment, occupied pair i <= j, virtual tile pair
// A <= B).  Fragments are ordered by the cost of their work units (longest
// auxiliary loop first).  The list is cut into 8 contiguous ranges of equal
// cost, one per XCD (workgroup b runs on XCD b % 8, so neighbouring units,
// which share a fragment's operands, share that XCD's L2); inside its range a
// persistent workgroup takes every (G/8)-th unit, largest first.  The loop
// bounds are kernel arguments, so every wave leaves the loop together.  (A
// per-workgroup atomic queue was not used: its divergent "thread 0 takes the
// next unit" branch lets the compiler rotate the loop so that the other lanes
// re-enter the barrier without the atomic.)  A deterministic reduction
// then sums each fragment's partials per pair (Neumaier) and the pairs per
// fragment (fixed tree order).
//
// Tile choice: every fragment has its own pair tile (virtual block width).
// For the modes with two tiles (fp64, fp32: 32 and 64 wide) a candidate is
// all fragments at 32, all at 64, or per fragment the 32-wide tile where the
// 64-wide one pads the most (mixed, encoded -1; rule in frag_tile);
// the mixed launch carries both tile paths and branches per unit on the
// fragment's tile (uniform within a workgroup's unit).  The autotuner times
// the persistent launch for each candidate on the batch itself and caches
// the winner per (mode, size class of the batch); rimp2_batch_set_tile fixes
// a candidate instead.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "batch_kernel.h"

namespace rimp2 {

struct BatchState {
  int mode = -1;
  int tile = 0;         // tile candidate of the layout: 32, 64, or -1 (mixed per fragment)
  int forced_tile = 0;  // rimp2_batch_set_tile: 0 = autotune
  int64_t nfrag = 0, total_wu = 0, total_pairs_sq = 0;
  std::vector<int64_t> nocc, nvirt, naux;
  std::vector<size_t> b64_off;  // FP64 B of each fragment in the b64 arena (doubles)
  DevBuf b64, packed, scales, eps_o, eps_v, partials, pairs, e_frag, frags;
  int64_t range[9] = {0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ready = false;
  std::map<std::pair<int, int>, int> tuned;  // (mode, size class) -> tile
  double tune_ms[4] = {0, 0, 0, 0};
  int tune_tiles[4] = {0, 0, 0, 0};
  int ntune = 0;
  MetricGroup group;  // grouped metric pipeline of the prepare
  DevBuf pack_jobs[2];  // job tables of the grouped packing (one per tile class)
  std::vector<unsigned char> pack_blob[2];
  DevBuf stage;       // host inputs staged on the device (V, then X, of the fragments with a metric)
  DevBuf fstatus;     // two status ints per fragment (bits as the context's status word)
  ~BatchState() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (DevBuf* b : {&b64, &packed, &scales, &eps_o, &eps_v, &partials, &pairs, &e_frag, &frags, &stage, &fstatus,
                      &pack_jobs[0], &pack_jobs[1]})
      b->release();
    group.release();
  }
};

void batch_destroy(BatchState* b) { delete b; }

namespace {

using namespace pairk;

// One workgroup per fragment: pair sums (Neumaier over its tile partials, in
// order), the fragment's [nocc,nocc] pair matrix and its total energy.
__global__ __launch_bounds__(256) void batch_reduce_kernel(const BatchFrag* frags, const double* partials,
                                                           double* pairs, double* e_frag) {
  const BatchFrag& fr = frags[blockIdx.x];
  __shared__ double ss[256], cc[256];
  double s = 0.0, c = 0.0;
  for (int64_t p = threadIdx.x; p < fr.npairs; p += 256) {
    const double* q = partials + fr.wu_begin + p * fr.ntp;
    double ps = 0.0, pc = 0.0;
    for (int64_t t = 0; t < fr.ntp; t++) {
      double x = q[t], u = ps + x;
      pc += (fabs(ps) >= fabs(x)) ? (ps - u) + x : (x - u) + ps;
      ps = u;
    }
    const double e = ps + pc;
    int64_t i, j;
    tri_decode(p, i, j);
    pairs[fr.pair_off + i + fr.nocc * j] = e;
    pairs[fr.pair_off + j + fr.nocc * i] = e;
    const double x = (i == j) ? e : 2.0 * e;
    const double u = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - u) + x : (x - u) + s;
    s = u;
  }
  ss[threadIdx.x] = s;
  cc[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      double a = ss[threadIdx.x], b = ss[threadIdx.x + w], u = a + b;
      double e = (fabs(a) >= fabs(b)) ? (a - u) + b : (b - u) + a;
      ss[threadIdx.x] = u;
      cc[threadIdx.x] += cc[threadIdx.x + w] + e;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) e_frag[fr.orig] = ss[0] + cc[0];
}

// tile: the layout's candidate (32, 64, or -1 = mixed per fragment)
void launch_batch(int mode, int tile, hipStream_t s, const BatchParams& p, int ncu) {
  if (tile == -1) return launch_batch_mixed(mode, s, p, ncu);
  switch (mode) {
    case RIMP2_MODE_FP64:
      if (tile == 32) return launch_batch_t<PolF64, 32, PolF64, 32>(s, p, ncu);
      if (tile == 64) return launch_batch_t<PolF64, 64, PolF64, 64>(s, p, ncu);
      break;
    case RIMP2_MODE_FP32:
      if (tile == 32) return launch_batch_t<PolF32R16, 32, PolF32R16, 32>(s, p, ncu);
      if (tile == 64) return launch_batch_t<PolF32, 64, PolF32, 64>(s, p, ncu);
      break;
    case RIMP2_MODE_BF16X3:
      if (tile == 64) return launch_batch_t<PolBF16x3, 64, PolBF16x3, 64>(s, p, ncu);
      break;
    case RIMP2_MODE_I8S4:
      if (tile == 64) return launch_batch_t<PolI8<4, 4>, 64, PolI8<4, 4>, 64>(s, p, ncu);
      break;
    case RIMP2_MODE_I8S5:
      if (tile == 64) return launch_batch_t<PolI8<5, 4>, 64, PolI8<5, 4>, 64>(s, p, ncu);
      break;
  }
  fail(RIMP2_ERR_ARG, "no batch pair kernel for mode %d tile %d", mode, tile);
}

// Tile candidates of a mode.  bf16x3 and the exact-slice modes have the
// 64-wide tile only: a 128-wide tile's wave groups would spill at two waves
// per SIMD, and a 32-wide int8 tile needs the 16x16x64 shape, which pads the
// auxiliary dimension to 256.
std::vector<int> tile_candidates(int mode) {
  switch (mode) {
    case RIMP2_MODE_FP64:
    case RIMP2_MODE_FP32: return {32, 64, -1};
    default: return {64};
  }
}

bool valid_candidate(int mode, int tile) {
  for (int t : tile_candidates(mode))
    if (t == tile) return true;
  return false;
}

// Pair tile of a fragment of nvirt virtuals under candidate `tile`.  The
// mixed candidate takes the 32-wide tile where the 64-wide tile's padding
// executes at least 1.5x its MFMA work (nvirt <= 32 and 65..96 only): per
// useful flop the 32-wide path runs ~1.2x slower (a unit
// does a quarter of the work against the same fixed cost), so a smaller
// padding saving loses (measured at 256 DEC fragments: every fragment with
// less padding on the 32-wide tile was 18% slower than all at 64).
int frag_tile(int tile, int64_t nvirt) {
  if (tile != -1) return tile;
  const double w64 = (double)round_up(nvirt, 64), w32 = (double)round_up(nvirt, 32);
  return w64 * w64 >= 1.5 * w32 * w32 ? 32 : 64;
}

int num_cus(int device) {
  int n = 0;
  RIMP2_HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device));
  return n > 0 ? n : 1;
}

// Packs every fragment's FP64 B (from the b64 arena) for the pair tile that
// candidate `tile` gives it, builds the work-unit table in cost order and
// uploads it.
void layout_batch(Ctx* c, BatchState* b, int tile) {
  hipStream_t s = c->stream;
  const int64_t nf = b->nfrag;
  std::vector<PackGeom> g(nf);
  std::vector<int> tf(nf);
  std::vector<double> ucost(nf);  // cost of one work unit: auxiliary depth x tile area
  size_t packed_bytes = 0, scale_n = 0, evp_n = 0;
  std::vector<size_t> poff(nf), soff(nf), eoff(nf);
  for (int64_t f = 0; f < nf; f++) {
    tf[f] = frag_tile(tile, b->nvirt[f]);
    g[f] = make_geom(b->mode, b->naux[f], b->nvirt[f], tf[f]);
    ucost[f] = (double)g[f].nqp * (double)tf[f] * (double)tf[f];
    poff[f] = packed_bytes;
    packed_bytes += round_up((int64_t)(g[f].bytes_per_occ() * b->nocc[f]), 256);
    soff[f] = scale_n;
    scale_n += (size_t)(g[f].nvp * b->nocc[f]);
    eoff[f] = evp_n;
    evp_n += (size_t)g[f].nvp;
  }
  unsigned char* pk = static_cast<unsigned char*>(b->packed.ensure(packed_bytes));
  double* sc = mode_slices(b->mode) ? static_cast<double*>(b->scales.ensure(scale_n * sizeof(double))) : nullptr;
  double* evp = static_cast<double*>(b->eps_v.ensure(evp_n * sizeof(double)));
  // raw eps_v of fragment f sits in eps_o arena after all occupied energies
  size_t occ_total = 0;
  for (int64_t f = 0; f < nf; f++) occ_total += (size_t)b->nocc[f];
  const double* eo = b->eps_o.as<double>();
  const double* ev_raw = eo + occ_total;
  std::vector<size_t> ooff(nf), voff(nf);
  {
    size_t o = 0, v = 0;
    for (int64_t f = 0; f < nf; f++) {
      ooff[f] = o;
      voff[f] = v;
      o += (size_t)b->nocc[f];
      v += (size_t)b->nvirt[f];
    }
  }
  {
    // one grouped packing per tile class (the operand layouts of the two
    // tiles differ in fp32), each with its own job table
    std::vector<PackItem> items[2];
    for (int64_t f = 0; f < nf; f++)
      items[tf[f] == 64].push_back(PackItem{b->b64.as<double>() + b->b64_off[f], b->naux[f], b->nvirt[f],
                                            b->nocc[f], g[f], pk + poff[f], sc ? sc + soff[f] : nullptr,
                                            ev_raw + voff[f], evp + eoff[f]});
    ScopedTimer tm(c, RIMP2_K_PACK);
    for (int k = 0; k < 2; k++) pack_B_group(items[k], b->mode, s, b->pack_jobs[k], b->pack_blob[k]);
  }
  // launch order: most expensive work units first
  std::vector<int64_t> order(nf);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return ucost[x] > ucost[y]; });
  std::vector<BatchFrag> fr(nf);
  int64_t wu = 0, psq = 0;
  std::vector<int64_t> pair_off(nf);
  for (int64_t f = 0; f < nf; f++) {
    pair_off[f] = psq;
    psq += b->nocc[f] * b->nocc[f];
  }
  for (int64_t k = 0; k < nf; k++) {
    const int64_t f = order[k];
    BatchFrag& x = fr[k];
    x.f.packed = pk + poff[f];
    x.f.sigma = sc ? sc + soff[f] : nullptr;
    x.f.eps_o = eo + ooff[f];
    x.f.eps_v = evp + eoff[f];
    x.f.nvp = g[f].nvp;
    x.f.nks = g[f].nks;
    x.f.nrb = g[f].nrb;
    x.f.nvirt = b->nvirt[f];
    x.tile = tf[f];
    x.nocc = b->nocc[f];
    x.nb = g[f].nvp / tf[f];
    x.ntp = pair_units(x.nb, pair_packs_diagonal(b->mode, tf[f]));
    x.npairs = x.nocc * (x.nocc + 1) / 2;
    x.wu_begin = wu;
    x.pair_off = pair_off[f];
    x.orig = f;
    wu += x.npairs * x.ntp;
  }
  b->total_wu = wu;
  b->total_pairs_sq = psq;
  BatchFrag* dfr = static_cast<BatchFrag*>(b->frags.ensure(sizeof(BatchFrag) * nf));
  RIMP2_HIP_CHECK(hipMemcpyAsync(dfr, fr.data(), sizeof(BatchFrag) * nf, hipMemcpyHostToDevice, s));
  b->partials.ensure((size_t)wu * sizeof(double));
  b->pairs.ensure((size_t)psq * sizeof(double));
  b->e_frag.ensure((size_t)nf * sizeof(double));
  // 8 contiguous ranges of equal cost
  {
    std::vector<double> cstart(nf + 1, 0.0);
    for (int64_t k = 0; k < nf; k++)
      cstart[k + 1] = cstart[k] + (double)(fr[k].npairs * fr[k].ntp) * ucost[order[k]];
    b->range[0] = 0;
    int64_t k = 0;
    for (int x = 1; x < 8; x++) {
      const double target = cstart[nf] * x / 8.0;
      while (k + 1 < nf && cstart[k + 1] <= target) k++;
      const int64_t n = fr[k].npairs * fr[k].ntp;
      int64_t u = (int64_t)std::ceil((target - cstart[k]) / ucost[order[k]]);
      u = std::min<int64_t>(std::max<int64_t>(u, 0), n);
      b->range[x] = std::max<int64_t>(fr[k].wu_begin + u, b->range[x - 1]);
    }
    b->range[8] = wu;
  }
  b->tile = tile;
  RIMP2_HIP_CHECK(hipStreamSynchronize(s));
}

void enqueue_pairs(Ctx* c, BatchState* b) {
  hipStream_t s = c->stream;
  BatchParams p;
  p.frags = b->frags.as<BatchFrag>();
  p.nfrag = b->nfrag;
  p.total = b->total_wu;
  for (int x = 0; x < 9; x++) p.range[x] = b->range[x];
  p.partials = b->partials.as<double>();
  launch_batch(b->mode, b->tile, s, p, num_cus(c->device));
  hipLaunchKernelGGL(batch_reduce_kernel, dim3((unsigned)b->nfrag), dim3(256), 0, s, p.frags, p.partials,
                     b->pairs.as<double>(), b->e_frag.as<double>());
  RIMP2_HIP_CHECK(hipGetLastError());
}

void capture_graph(Ctx* c, BatchState* b) {
  if (b->exec) RIMP2_HIP_CHECK(hipGraphExecDestroy(b->exec));
  if (b->graph) RIMP2_HIP_CHECK(hipGraphDestroy(b->graph));
  b->exec = nullptr;
  b->graph = nullptr;
  RIMP2_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_pairs(c, b);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  RIMP2_HIP_CHECK(hipStreamEndCapture(c->stream, &b->graph));
  RIMP2_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

double time_once(Ctx* c, BatchState* b) {
  hipEvent_t e0, e1;
  RIMP2_HIP_CHECK(hipEventCreate(&e0));
  RIMP2_HIP_CHECK(hipEventCreate(&e1));
  RIMP2_HIP_CHECK(hipEventRecord(e0, c->stream));
  enqueue_pairs(c, b);
  RIMP2_HIP_CHECK(hipEventRecord(e1, c->stream));
  RIMP2_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  RIMP2_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

// size class of a batch: log2 of the mean virtual dimension
int size_class(const BatchState* b) {
  double m = 0;
  for (int64_t v : b->nvirt) m += (double)v;
  m /= (double)(b->nfrag > 0 ? b->nfrag : 1);
  int k = 0;
  while ((1 << (k + 1)) <= (int)m) k++;
  return k;
}

// Metric, transform and staging of B into the b64 arena for every fragment:
// one grouped pipeline (one launch per Cholesky step, inverse level, packing
// and transform for all fragments together); on_device: every pointer of
// the descriptors is device memory (inputs already in HBM), else host memory
// (staged on the device first).  Input checks set per-fragment status words,
// read once at the end.
void prepare(Ctx* c, int64_t nfrag, const rimp2_frag_desc* frags, int mode, bool on_device) {
  if (nfrag <= 0 || !frags) fail(RIMP2_ERR_ARG, "empty fragment list");
  if (mode < 0 || mode >= RIMP2_NUM_MODES) fail(RIMP2_ERR_ARG, "unknown precision mode %d", mode);
  if (!c->batch) c->batch = new BatchState();
  BatchState* b = c->batch;
  b->ready = false;
  b->mode = mode;
  b->nfrag = nfrag;
  b->nocc.resize(nfrag);
  b->nvirt.resize(nfrag);
  b->naux.resize(nfrag);
  b->b64_off.resize(nfrag);
  size_t b64n = 0, occ_total = 0, virt_total = 0, stage_n = 0;
  for (int64_t f = 0; f < nfrag; f++) {
    const rimp2_frag_desc& d = frags[f];
    if (d.nocc <= 0 || d.nvirt <= 0 || d.naux <= 0 || d.nocc > 65536 || d.nvirt > 65536 || d.naux > 65536)
      fail(RIMP2_ERR_ARG, "fragment %lld: bad dimensions", (long long)f);
    if (!d.Pia || !d.eps_o || !d.eps_v) fail(RIMP2_ERR_ARG, "fragment %lld: null array", (long long)f);
    if (mode_slices(mode) && d.naux > 16384)
      fail(RIMP2_ERR_ARG, "fragment %lld: exact-slice modes support naux <= 16384", (long long)f);
    if (!on_device) {
      // fmax / fmin skip a NaN, so every orbital energy is tested on its own
      double mo = -INFINITY, mv = INFINITY;
      bool finite = true;
      for (int64_t i = 0; i < d.nocc; i++) {
        finite = finite && std::isfinite(d.eps_o[i]);
        mo = fmax(mo, d.eps_o[i]);
      }
      for (int64_t a = 0; a < d.nvirt; a++) {
        finite = finite && std::isfinite(d.eps_v[a]);
        mv = fmin(mv, d.eps_v[a]);
      }
      if (!finite) fail(RIMP2_ERR_NUMERIC, "fragment %lld: non-finite orbital energy", (long long)f);
      if (!(2.0 * mo - 2.0 * mv < 0.0))
        fail(RIMP2_ERR_NUMERIC, "fragment %lld: orbital-energy denominator >= 0", (long long)f);
      if (d.V) stage_n += (size_t)(d.naux * d.naux + d.naux * d.nvirt * d.nocc);
    }
    b->nocc[f] = d.nocc;
    b->nvirt[f] = d.nvirt;
    b->naux[f] = d.naux;
    b->b64_off[f] = b64n;
    b64n += (size_t)(d.naux * d.nvirt * d.nocc);
    occ_total += (size_t)d.nocc;
    virt_total += (size_t)d.nvirt;
  }
  double* b64 = static_cast<double*>(b->b64.ensure(b64n * sizeof(double)));
  double* eo = static_cast<double*>(b->eps_o.ensure((occ_total + virt_total) * sizeof(double)));
  double* ev = eo + occ_total;
  int* fst = static_cast<int*>(b->fstatus.ensure((size_t)nfrag * 2 * sizeof(int)));
  double* stage = stage_n ? static_cast<double*>(b->stage.ensure(stage_n * sizeof(double))) : nullptr;
  hipStream_t s = c->stream;
  RIMP2_HIP_CHECK(hipMemsetAsync(fst, 0, (size_t)nfrag * 2 * sizeof(int), s));
  std::vector<mg::Item> items;
  std::vector<mg::Copy> copies;
  std::vector<mg::Denom> denoms;
  // host inputs: the orbital energies of all fragments go over in one copy
  std::vector<double> heps(on_device ? 0 : occ_total + virt_total);
  size_t o = 0, v = 0, so = 0;
  for (int64_t f = 0; f < nfrag; f++) {
    const rimp2_frag_desc& d = frags[f];
    int* st = fst + 2 * f;
    const int64_t nx = d.naux * d.nvirt * d.nocc;
    double* dst = b64 + b->b64_off[f];
    if (on_device) {
      copies.push_back(mg::Copy{d.eps_o, eo + o, d.nocc, st});
      copies.push_back(mg::Copy{d.eps_v, ev + v, d.nvirt, st});
      if (d.V) items.push_back(mg::Item{d.naux, d.nvirt, d.nocc, d.V, d.Pia, dst, st});
      else copies.push_back(mg::Copy{d.Pia, dst, nx, st});  // already fitted: Pia is B
    } else {
      memcpy(heps.data() + o, d.eps_o, (size_t)d.nocc * sizeof(double));
      memcpy(heps.data() + occ_total + v, d.eps_v, (size_t)d.nvirt * sizeof(double));
      if (d.V) {
        double* sV = stage + so;
        double* sX = sV + d.naux * d.naux;
        so += (size_t)(d.naux * d.naux + nx);
        h2d_copy(c, c->stream, sV, d.V, (size_t)(d.naux * d.naux) * sizeof(double));
        h2d_copy(c, c->stream, sX, d.Pia, (size_t)nx * sizeof(double));
        items.push_back(mg::Item{d.naux, d.nvirt, d.nocc, sV, sX, dst, st});
      } else {
        h2d_copy(c, c->stream, dst, d.Pia, (size_t)nx * sizeof(double));
        copies.push_back(mg::Copy{dst, dst, nx, st});
      }
    }
    denoms.push_back(mg::Denom{eo + o, ev + v, d.nocc, d.nvirt, st});
    o += (size_t)d.nocc;
    v += (size_t)d.nvirt;
  }
  if (!on_device) h2d_copy(c, s, eo, heps.data(), heps.size() * sizeof(double));
  // the exact-slice modes transform on int8 MFMA (six slices per operand)
  metric_group_run(b->group, s, items, copies, denoms, mode_slices(mode) != 0);
  std::vector<int> hst((size_t)nfrag * 2);
  RIMP2_HIP_CHECK(hipMemcpyAsync(hst.data(), fst, hst.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  RIMP2_HIP_CHECK(hipStreamSynchronize(s));
  for (int64_t f = 0; f < nfrag; f++) {
    const int bits = hst[2 * f];
    if (bits & 4) fail(RIMP2_ERR_NUMERIC, "fragment %lld: non-finite value in an input", (long long)f);
    if (bits & 8) fail(RIMP2_ERR_NUMERIC, "fragment %lld: orbital-energy denominator >= 0", (long long)f);
    if (bits & 2)
      fail(RIMP2_ERR_NOT_SPD, "fragment %lld: metric not positive definite (pivot %d)", (long long)f, hst[2 * f + 1]);
  }
  // the single-fragment state no longer matches the context buffers
  c->have_metric = false;
  c->have_B = false;
  // autotune the pair tile for (mode, size class), unless the caller fixed it
  const std::pair<int, int> key(mode, size_class(b));
  auto it = b->tuned.find(key);
  b->ntune = 0;
  int best = 0;
  if (b->forced_tile) {
    if (!valid_candidate(mode, b->forced_tile))
      fail(RIMP2_ERR_ARG, "tile candidate %d is not available in precision mode %d", b->forced_tile, mode);
    best = b->forced_tile;
  } else if (it != b->tuned.end()) {
    best = it->second;
  } else {
    double best_ms = 1e300;
    for (int t : tile_candidates(mode)) {
      layout_batch(c, b, t);
      time_once(c, b);  // warm-up (first launch of the instantiation)
      const double ms = time_once(c, b);
      if (b->ntune < 4) {
        b->tune_tiles[b->ntune] = t;
        b->tune_ms[b->ntune] = ms;
        b->ntune++;
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    b->tuned[key] = best;
  }
  layout_batch(c, b, best);
  capture_graph(c, b);
  b->ready = true;
}

}  // namespace
}  // namespace rimp2

using namespace rimp2;

extern "C" {

int rimp2_batch_prepare(void* ctx, int64_t nfrag, const rimp2_frag_desc* frags, int mode) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  prepare(c, nfrag, frags, mode, false);
  RIMP2_API_END(ctx)
}

int rimp2_batch_prepare_device(void* ctx, int64_t nfrag, const rimp2_frag_desc* frags, int mode) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  prepare(c, nfrag, frags, mode, true);
  RIMP2_API_END(ctx)
}

int rimp2_batch_run(void* ctx, double* d_e_frag) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  BatchState* b = c->batch;
  if (!b || !b->ready) fail(RIMP2_ERR_STATE, "rimp2_batch_prepare must run first");
  ScopedTimer t(c, RIMP2_K_BATCH);
  RIMP2_HIP_CHECK(hipGraphLaunch(b->exec, c->stream));
  if (d_e_frag)
    RIMP2_HIP_CHECK(hipMemcpyAsync(d_e_frag, b->e_frag.ptr, (size_t)b->nfrag * sizeof(double),
                                   hipMemcpyDeviceToDevice, c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_batch_result(void* ctx, double* e_frag, double* eps_pair_packed) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  BatchState* b = c->batch;
  if (!b || !b->ready) fail(RIMP2_ERR_STATE, "rimp2_batch_prepare must run first");
  if (!e_frag) fail(RIMP2_ERR_ARG, "null e_frag");
  RIMP2_HIP_CHECK(hipMemcpyAsync(e_frag, b->e_frag.ptr, (size_t)b->nfrag * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
  if (eps_pair_packed)
    RIMP2_HIP_CHECK(hipMemcpyAsync(eps_pair_packed, b->pairs.ptr, (size_t)b->total_pairs_sq * sizeof(double),
                                   hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_batch(void* ctx, int64_t nfrag, const rimp2_frag_desc* frags, int mode, double* e_frag,
                double* eps_pair_packed) {
  int st = rimp2_batch_prepare(ctx, nfrag, frags, mode);
  if (st) return st;
  st = rimp2_batch_run(ctx, nullptr);
  if (st) return st;
  return rimp2_batch_result(ctx, e_frag, eps_pair_packed);
}

int rimp2_batch_tuning(void* ctx, int* n, int64_t* tiles, double* ms) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (!n || !tiles || !ms) fail(RIMP2_ERR_ARG, "null output pointer");
  if (!c->batch || !c->batch->ready) fail(RIMP2_ERR_STATE, "rimp2_batch_prepare must run first");
  const BatchState* b = c->batch;
  *n = b->ntune;
  for (int k = 0; k < b->ntune; k++) {
    tiles[k] = b->tune_tiles[k];
    ms[k] = b->tune_ms[k];
  }
  RIMP2_API_END(ctx)
}

int rimp2_batch_set_tile(void* ctx, int64_t tile) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (tile != 0 && tile != 32 && tile != 64 && tile != -1)
    fail(RIMP2_ERR_ARG, "tile candidate %lld: expected 0 (autotune), 32, 64 or -1 (mixed)", (long long)tile);
  if (!c->batch) c->batch = new BatchState();
  c->batch->forced_tile = (int)tile;
  RIMP2_API_END(ctx)
}

int rimp2_batch_tile(void* ctx, int64_t* tile) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (!tile) fail(RIMP2_ERR_ARG, "null tile");
  if (!c->batch || !c->batch->ready) fail(RIMP2_ERR_STATE, "rimp2_batch_prepare must run first");
  *tile = c->batch->tile;
  RIMP2_API_END(ctx)
}

}  // extern "C"
