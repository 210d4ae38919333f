// Fused RI-MP2 pair kernel.
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
#include <algorithm>

#include "pair_core.h"

namespace rimp2 {
namespace {

using namespace pairk;

struct PairParams {
  FragView f;
  double* partials;             // [npairs * ntp]
  const int64_t* pair_list;     // optional list of unique pair indices (j*(j+1)/2+i)
  int64_t nocc;
  int64_t nb, ntp;              // virtual blocks, work units per occupied pair
  int64_t npairs;               // pairs processed (list length or all)
  int64_t nwu;                  // npairs * ntp
  int64_t chunk;                // work units per XCD slot
};

template <class Pol, int TM>
__global__ __launch_bounds__(PAIR_THREADS) __attribute__((amdgpu_waves_per_eu(pair_waves_per_eu<Pol, TM>())))
void pair_kernel(PairParams p) {
  RIMP2_PAIR_STAGE_BUFFERS(Pol, TM);
  // XCD-aware mapping: consecutive workgroups land on different XCDs
  // (round-robin dispatch), so each XCD owns a contiguous run of work units
  // (neighbouring units share an occupied pair and hit the same L2); slot s
  // of XCD x takes units x * chunk + s, + nslot, ...  (one unit per
  // workgroup unless the grid had to be capped; the loop bounds are uniform
  // over the workgroup).
  const int64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
  for (int64_t k = slot; k < p.chunk; k += nslot) {
    const int64_t wu = xcd * p.chunk + k;
    if (wu >= p.nwu) break;
    const int64_t pidx = wu / p.ntp, tp = wu % p.ntp;
    int64_t i, j, A, Bk;
    tri_decode(p.pair_list ? uni(gload(p.pair_list + pidx)) : pidx, i, j);
    const int kind = uni_int(unit_decode(tp, p.nb, packs_diagonal<Pol, TM>(), A, Bk));
    const double e = pair_tile<Pol, TM>(p.f, uni(i), uni(j), uni(A), uni(Bk), kind, rimp2_buf0, rimp2_buf1,
                                        rimp2_buf2, rimp2_buf3);
    if (threadIdx.x == 0) p.partials[wu] = e;
  }
}

// Amplitude form over one chunk of occupied orbitals: the pairs i <= j with
// i in [i0, i0 + nchunk), ordered by i, then j; units of a pair are
// consecutive (same XCD mapping as pair_kernel).
template <class Pol, int TM>
__global__ __launch_bounds__(PAIR_THREADS) void pair_ampl_kernel(PairParams p, AmplOut o, int64_t nchunk) {
  RIMP2_PAIR_STAGE_BUFFERS(Pol, TM);
  const int64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
  for (int64_t k = slot; k < p.chunk; k += nslot) {
    const int64_t wu = xcd * p.chunk + k;
    if (wu >= p.nwu) break;
    const int64_t pidx = wu / p.ntp, tp = wu % p.ntp;
    // pairs of orbital i0 + z: nj - z  (nj = nocc - i0)
    int64_t z = 0, rem = pidx;
    while (z + 1 < nchunk && rem >= o.nj - z) {
      rem -= o.nj - z;
      z++;
    }
    const int64_t i = o.i0 + z, j = i + rem;
    int64_t A, Bk;
    const int kind = unit_decode(tp, p.nb, packs_diagonal<Pol, TM>(), A, Bk);
    pair_tile<Pol, TM, true>(p.f, i, j, A, Bk, kind, rimp2_buf0, rimp2_buf1, rimp2_buf2, rimp2_buf3, &o);
  }
}

// Per pair: Neumaier sum of its ntp partials in order -> eps_pair (both
// triangles) and the pair's contribution to E (x2 off the diagonal).
__global__ void pair_reduce_kernel(const double* partials, int64_t ntp, int64_t npairs, const int64_t* pair_list,
                                   int64_t nocc, double* eps_pair, double* contrib) {
  const int64_t pidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= npairs) return;
  double s = 0.0, c = 0.0;
  const double* q = partials + pidx * ntp;
  for (int64_t t = 0; t < ntp; t++) {
    double x = q[t], u = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - u) + x : (x - u) + s;
    s = u;
  }
  const double e = s + c;
  int64_t i, j;
  tri_decode(pair_list ? pair_list[pidx] : pidx, i, j);
  if (eps_pair) {
    eps_pair[i + nocc * j] = e;
    eps_pair[j + nocc * i] = e;
  }
  contrib[pidx] = (i == j) ? e : 2.0 * e;
}

// E = sum of contributions; 256 strided Neumaier partial sums combined in a
// fixed tree order (deterministic).
__global__ __launch_bounds__(256) void total_kernel(const double* contrib, int64_t n, double* etot) {
  __shared__ double ss[256], cc[256];
  double s = 0.0, c = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += 256) {
    double x = contrib[k], u = s + x;
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
  if (threadIdx.x == 0) *etot = ss[0] + cc[0];
}

__global__ void pad_eps_v_kernel(const double* ev, int64_t nvirt, int64_t nvp, double* out) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a < nvp) out[a] = (a < nvirt) ? ev[a] : 1.0;
}

// status bit 4: non-finite value;  bit 8: some denominator e_i+e_j-e_a-e_b >= 0
// The check runs on a capped grid beside the Cholesky, so each thread keeps
// four 16-byte loads in flight per iteration (one 8-byte load each reached
// 1.6 TB/s); each wave instruction reads 1 KB contiguous.
__global__ void finite_check_kernel(const double* x, int64_t n, int* status) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, nth = (int64_t)gridDim.x * blockDim.x;
  bool bad = false;
  if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
    const double2* x2 = reinterpret_cast<const double2*>(x);
    const int64_t n2 = n / 2;
    for (int64_t b = 0; b < n2 && !bad; b += 4 * nth) {
      double2 v[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int64_t i = b + e * nth + tid;
        v[e] = i < n2 ? x2[i] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int e = 0; e < 4; e++) bad |= !isfinite(v[e].x) || !isfinite(v[e].y);
    }
    if (tid == 0 && (n & 1)) bad |= !isfinite(x[n - 1]);
  } else {
    for (int64_t k = tid; k < n && !bad; k += nth) bad = !isfinite(x[k]);
  }
  if (bad) atomicOr(status, 4);
}
__global__ void denom_check_kernel(const double* eo, int64_t nocc, const double* ev, int64_t nvirt, int* status) {
  __shared__ double mo, mv;
  if (threadIdx.x == 0) {
    mo = -INFINITY;
    mv = INFINITY;
    for (int64_t i = 0; i < nocc; i++) mo = fmax(mo, eo[i]);
    for (int64_t a = 0; a < nvirt; a++) mv = fmin(mv, ev[a]);
    if (!(2.0 * mo - 2.0 * mv < 0.0)) atomicOr(status, 8);
  }
}

template <class Pol, int TM>
void launch_pair(hipStream_t s, const PairParams& p) {
  const int64_t nslot = p.chunk < kMaxGridBlocks / 8 ? p.chunk : kMaxGridBlocks / 8;
  hipLaunchKernelGGL((pair_kernel<Pol, TM>), dim3((unsigned)(8 * nslot)), dim3(PAIR_THREADS), 0, s, p);
  RIMP2_HIP_CHECK(hipGetLastError());
}

}  // namespace

// Tile (virtual block width) per mode and shape.  Larger tiles raise MFMA
// operand reuse; smaller ones keep small fragments from idling CUs.
int choose_tile(int mode, int64_t nvirt, int64_t nocc) {
  (void)nocc;
  switch (mode) {
    case RIMP2_MODE_FP64:
    case RIMP2_MODE_FP32: return nvirt <= 48 ? 32 : 64;
    case RIMP2_MODE_BF16X3: return 64;  // a 128 tile's split wave groups would spill at two waves per SIMD
    default: return 64;
  }
}

// Whether the pair kernel of (mode, tile) packs two diagonal blocks per unit.
bool pair_packs_diagonal(int mode, int tile) {
  switch (mode) {
    case RIMP2_MODE_FP64: return tile == 32 ? packs_diagonal<PolF64, 32>() : packs_diagonal<PolF64, 64>();
    case RIMP2_MODE_FP32: return tile == 32 ? packs_diagonal<PolF32R16, 32>() : packs_diagonal<PolF32, 64>();
    case RIMP2_MODE_BF16X3: return packs_diagonal<PolBF16x3, 64>();
    case RIMP2_MODE_I8S4: return packs_diagonal<PolI8<4, 4>, 64>();
    default: return packs_diagonal<PolI8<5, 4>, 64>();
  }
}

static void dispatch_pair(int mode, int tile, hipStream_t s, const PairParams& p) {
  switch (mode) {
    case RIMP2_MODE_FP64:
      if (tile == 32) return launch_pair<PolF64, 32>(s, p);
      if (tile == 64) return launch_pair<PolF64, 64>(s, p);
      break;
    case RIMP2_MODE_FP32:
      if (tile == 32) return launch_pair<PolF32R16, 32>(s, p);
      if (tile == 64) return launch_pair<PolF32, 64>(s, p);
      break;
    case RIMP2_MODE_BF16X3:
      if (tile == 64) return launch_pair<PolBF16x3, 64>(s, p);
      break;
    case RIMP2_MODE_I8S4:
      if (tile == 64) return launch_pair<PolI8<4, 4>, 64>(s, p);
      break;
    case RIMP2_MODE_I8S5:
      if (tile == 64) return launch_pair<PolI8<5, 4>, 64>(s, p);
      break;
  }
  fail(RIMP2_ERR_ARG, "no pair kernel for mode %d tile %d", mode, tile);
}

// Device bytes pair_energies_core needs for its partials and padded virtual
// energies (both 256-byte aligned, partials first).
size_t pair_scratch_bytes(const PackGeom& g, int tile, int64_t npairs, size_t* evp_offset) {
  const int64_t ntp = pair_units(g.nvp / tile, pair_packs_diagonal(g.mode, tile));
  const size_t part = (size_t)round_up((int64_t)((size_t)(npairs * ntp + npairs + 1) * sizeof(double)), 256);
  if (evp_offset) *evp_offset = part;
  return part + (size_t)round_up(g.nvp * (int64_t)sizeof(double), 256);
}

// The pair kernel + reductions on stream s with caller-provided scratch (evp:
// nvp doubles; part: npairs * ntp + npairs + 1 doubles); allocates nothing.
// Every eps_pair entry is written when all pairs run; with a pair list the
// matrix is cleared first (skipped pairs stay zero).
void pair_energies_core(Ctx* timer, hipStream_t s, int tile, const PackGeom& g, const void* packed,
                        const double* scales, int64_t nocc, int64_t nvirt, const double* d_eps_o,
                        const double* d_eps_v, double* d_eps_pair, double* d_etot, const int64_t* d_pair_list,
                        int64_t npairs_listed, double* evp, double* part) {
  const int64_t nb = g.nvp / tile;
  const int64_t ntp = pair_units(nb, pair_packs_diagonal(g.mode, tile));
  const int64_t npairs = d_pair_list ? npairs_listed : nocc * (nocc + 1) / 2;
  hipLaunchKernelGGL(pad_eps_v_kernel, dim3((unsigned)ceil_div(g.nvp, 256)), dim3(256), 0, s, d_eps_v, nvirt, g.nvp,
                     evp);
  RIMP2_HIP_CHECK(hipGetLastError());
  const size_t npart = (size_t)(npairs * ntp);
  double* contrib = part + npart;
  if (d_eps_pair && d_pair_list)
    RIMP2_HIP_CHECK(hipMemsetAsync(d_eps_pair, 0, (size_t)(nocc * nocc) * sizeof(double), s));
  if (npairs > 0) {
    PairParams p;
    p.f.packed = static_cast<const unsigned char*>(packed);
    p.f.sigma = scales;
    p.f.eps_o = d_eps_o;
    p.f.eps_v = evp;
    p.f.nvp = g.nvp;
    p.f.nks = g.nks;
    p.f.nrb = g.nrb;
    p.f.nvirt = nvirt;
    p.partials = part;
    p.pair_list = d_pair_list;
    p.nocc = nocc;
    p.nb = nb;
    p.ntp = ntp;
    p.npairs = npairs;
    p.nwu = npairs * ntp;
    p.chunk = ceil_div(p.nwu, 8);
    {
      ScopedTimer t(timer, RIMP2_K_PAIR);
      dispatch_pair(g.mode, tile, s, p);
    }
    ScopedTimer t(timer, RIMP2_K_REDUCE);
    hipLaunchKernelGGL(pair_reduce_kernel, dim3((unsigned)ceil_div(npairs, 256)), dim3(256), 0, s, part, ntp, npairs,
                       d_pair_list, nocc, d_eps_pair, contrib);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(total_kernel, dim3(1), dim3(256), 0, s, contrib, npairs, d_etot);
  RIMP2_HIP_CHECK(hipGetLastError());
}

// Runs the pair kernel + reductions on ctx->stream, with the context's
// partials and padded-energy buffers.  d_pair_list (optional) restricts the
// work to the listed unique pairs (screening).
void pair_energies_device(Ctx* c, const PackGeom& g, const void* packed, const double* scales, int64_t nocc,
                          int64_t nvirt, const double* d_eps_o, const double* d_eps_v, double* d_eps_pair,
                          double* d_etot, const int64_t* d_pair_list, int64_t npairs_listed) {
  const int64_t npairs = d_pair_list ? npairs_listed : nocc * (nocc + 1) / 2;
  const int64_t ntp = pair_units(g.nvp / c->tile, pair_packs_diagonal(g.mode, c->tile));
  double* evp = static_cast<double*>(c->eps_v.ensure((size_t)g.nvp * sizeof(double)));
  double* part = static_cast<double*>(c->partials.ensure((size_t)(npairs * ntp + npairs + 1) * sizeof(double)));
  pair_energies_core(c, c->stream, c->tile, g, packed, scales, nocc, nvirt, d_eps_o, d_eps_v, d_eps_pair, d_etot,
                     d_pair_list, npairs_listed, evp, part);
}

// Amplitude slab of one chunk (pair kernel, 64-wide virtual blocks) on
// stream s: T[a + nv b + nv^2 (j - i0) + nv^2 (nocc - i0) (i - i0)] =
// w_ij V^{ij}_ab / Delta for i in [i0, i0 + nchunk), j >= i (entries j < i
// are not written).  packed: the FP64 or exact-slice i8s5 packing of the
// 3-index tensor, geometry g (tile 64); sigma: the i8s5 row scales.
void pair_amplitudes_device(hipStream_t s, const PackGeom& g, const void* packed, const double* sigma, int64_t nocc,
                            int64_t nvirt, const double* d_eo, const double* d_ev, int64_t i0, int64_t nchunk,
                            double* T) {
  constexpr int TM = 64;
  if ((g.mode != RIMP2_MODE_FP64 && g.mode != RIMP2_MODE_I8S5) || g.nvp % TM)
    fail(RIMP2_ERR_ARG, "amplitude slab needs the FP64 or i8s5 64-wide packing");
  if (g.mode == RIMP2_MODE_I8S5 && !sigma) fail(RIMP2_ERR_ARG, "i8s5 amplitude slab without row scales");
  if (i0 < 0 || nchunk < 1 || i0 + nchunk > nocc) fail(RIMP2_ERR_ARG, "bad occupied chunk");
  const int64_t nb = g.nvp / TM;
  const bool i8 = g.mode == RIMP2_MODE_I8S5;
  const int64_t ntp = pair_units(nb, i8 ? packs_diagonal<PolI8<5, 4>, TM>() : packs_diagonal<PolF64, TM>());
  const int64_t nj = nocc - i0;
  const int64_t npairs = nchunk * nj - nchunk * (nchunk - 1) / 2;
  PairParams p;
  p.f.packed = static_cast<const unsigned char*>(packed);
  p.f.sigma = sigma;
  p.f.eps_o = d_eo;
  p.f.eps_v = d_ev;
  p.f.nvp = g.nvp;
  p.f.nks = g.nks;
  p.f.nrb = g.nrb;
  p.f.nvirt = nvirt;
  p.partials = nullptr;
  p.pair_list = nullptr;
  p.nocc = nocc;
  p.nb = nb;
  p.ntp = ntp;
  p.npairs = npairs;
  p.nwu = npairs * ntp;
  p.chunk = ceil_div(p.nwu, 8);
  AmplOut o{T, d_eo, d_ev, nvirt, nj, i0};
  const int64_t nslot = p.chunk < kMaxGridBlocks / 8 ? p.chunk : kMaxGridBlocks / 8;
  if (i8)
    hipLaunchKernelGGL((pair_ampl_kernel<PolI8<5, 4>, TM>), dim3((unsigned)(8 * nslot)), dim3(PAIR_THREADS), 0, s, p,
                       o, nchunk);
  else
    hipLaunchKernelGGL((pair_ampl_kernel<PolF64, TM>), dim3((unsigned)(8 * nslot)), dim3(PAIR_THREADS), 0, s, p, o,
                       nchunk);
  RIMP2_HIP_CHECK(hipGetLastError());
}

void check_finite_on(hipStream_t s, int* status, const double* p, int64_t n, unsigned max_blocks) {
  const unsigned blocks = (unsigned)std::min<int64_t>(max_blocks, std::max<int64_t>(1, ceil_div(n, 256 * 8)));
  hipLaunchKernelGGL(finite_check_kernel, dim3(blocks), dim3(256), 0, s, p, n, status);
  RIMP2_HIP_CHECK(hipGetLastError());
}

void check_denom_on(hipStream_t s, int* status, const double* eo, int64_t nocc, const double* ev, int64_t nvirt) {
  hipLaunchKernelGGL(denom_check_kernel, dim3(1), dim3(64), 0, s, eo, nocc, ev, nvirt, status);
  RIMP2_HIP_CHECK(hipGetLastError());
}

void check_inputs_device(Ctx* c, const double* p, int64_t n, const char* what) {
  (void)what;
  check_finite_on(c->stream, static_cast<int*>(c->status.ensure(64)), p, n);
}

void denom_check_device(Ctx* c, const double* eo, int64_t nocc, const double* ev, int64_t nvirt) {
  check_denom_on(c->stream, static_cast<int*>(c->status.ensure(64)), eo, nocc, ev, nvirt);
}

}  // namespace rimp2
