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
// Per-lane Kahan sums are reduced by wave shuffles and LDS to one FP64 partial
// per workgroup; the partials are summed per pair and then over pairs in a
// fixed order with Neumaier compensation (deterministic).
#include "rimp2_common.h"

namespace rimp2 {
namespace {

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

// ---------------------------------------------------------------------------
struct PairParams {
  const unsigned char* packed;  // packed B (layout in pack.hip)
  const double* sigma;          // int8 row scales [nocc][nvp] (null otherwise)
  const double* eps_o;          // [nocc]
  const double* eps_v;          // [nvp], padding entries = 1.0
  double* partials;             // [npairs * ntp]
  const int64_t* pair_list;     // optional list of unique pair indices (j*(j+1)/2+i)
  int64_t nocc, nvp, nks, nrb;
  int64_t nb, ntp;              // virtual blocks, tile pairs per occupied pair
  int64_t npairs;               // pairs processed (list length or all)
  int64_t nwu;                  // npairs * ntp
  int64_t chunk;                // work units per XCD slot
};

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

template <class Pol, int TM>
__global__ __launch_bounds__(256) void pair_kernel(PairParams p) {
  constexpr int FPB = TM / Pol::RB;           // fragments per block side
  constexpr int WT = TM / 2;                  // wave tile side
  constexpr int FW = WT / Pol::RB;            // fragments per wave side
  constexpr int FB = 64 * Pol::LB;            // bytes per fragment
  constexpr int OPB = Pol::KSTEPS * FPB * Pol::P * FB;  // bytes per operand per stage
  static_assert(FW >= 1, "wave tile smaller than one MFMA fragment");
  extern __shared__ __align__(16) unsigned char lds[];

  // XCD-aware mapping: consecutive workgroups land on different XCDs
  // (round-robin dispatch), so give each XCD slot a contiguous run of work
  // units: neighbouring units share an occupied pair and hit the same L2.
  const int64_t bid = blockIdx.x;
  const int64_t wu = (bid & 7) * p.chunk + (bid >> 3);
  if (wu >= p.nwu) return;
  const int64_t pidx = wu / p.ntp, tp = wu % p.ntp;
  int64_t i, j, A, Bk;
  tri_decode(p.pair_list ? p.pair_list[pidx] : pidx, i, j);
  tri_decode(tp, A, Bk);  // A <= Bk
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
  const size_t per_ks = (size_t)p.nrb * Pol::P * FB;
  const size_t per_occ = (size_t)p.nks * per_ks;
  const unsigned char* src[4];
  src[0] = p.packed + i * per_occ + (size_t)A * FPB * Pol::P * FB;
  src[1] = p.packed + j * per_occ + (size_t)Bk * FPB * Pol::P * FB;
  src[2] = p.packed + j * per_occ + (size_t)A * FPB * Pol::P * FB;
  src[3] = p.packed + i * per_occ + (size_t)Bk * FPB * Pol::P * FB;
  const int nops = same_occ ? 2 : 4;
  constexpr int CHUNK = FPB * Pol::P * FB;  // contiguous bytes per operand per k-step

  for (int64_t ks0 = 0; ks0 < p.nks; ks0 += Pol::KSTEPS) {
    // global -> LDS, 16 bytes per thread per iteration
    for (int op = 0; op < nops; op++) {
      const unsigned char* s = src[op] + (size_t)ks0 * per_ks;
      unsigned char* d = lds + op * OPB;
      for (int off = tid * 16; off < OPB; off += 256 * 16) {
        const int ks = off / CHUNK, r = off % CHUNK;
        *reinterpret_cast<uint4*>(d + off) = *reinterpret_cast<const uint4*>(s + (size_t)ks * per_ks + r);
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < Pol::KSTEPS; ks++) {
      typename Pol::frag_t ra[FW][Pol::P], cb[FW][Pol::P];
      const unsigned char* base = lds + ks * CHUNK;
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
    __syncthreads();
  }

  // ---- epilogue: energy terms in FP64 ----
  const double ei = p.eps_o[i], ej = p.eps_o[j];
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
          v1 *= p.sigma[i * p.nvp + a] * p.sigma[j * p.nvp + b];
          if (!same_occ) v2 *= p.sigma[j * p.nvp + a] * p.sigma[i * p.nvp + b];
          else v2 = v1;
        }
        const double inv = 1.0 / (ei + ej - p.eps_v[a] - p.eps_v[b]);
        double e;
        if (same_blk) e = v1 * (2.0 * v1 - v2);
        else e = 2.0 * (v1 * v1 + v2 * v2 - v1 * v2);
        kahan_add(s, c, e * inv);
      }
  double tot = s - c;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
  __shared__ double wsum[4];
  if (lane == 0) wsum[wave] = tot;
  __syncthreads();
  if (tid == 0) p.partials[wu] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
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
    if (threadIdx.x < w) {
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
__global__ void finite_check_kernel(const double* x, int64_t n, int* status) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; k < n; k += (int64_t)gridDim.x * blockDim.x)
    if (!isfinite(x[k])) { atomicOr(status, 4); return; }
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
  constexpr int FPB = TM / Pol::RB;
  constexpr int OPB = Pol::KSTEPS * FPB * Pol::P * 64 * Pol::LB;
  const size_t lds = 4 * (size_t)OPB;
  static bool attr_set = false;
  if (!attr_set) {
    RIMP2_HIP_CHECK(hipFuncSetAttribute((const void*)pair_kernel<Pol, TM>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  const unsigned grid = (unsigned)(8 * p.chunk);
  hipLaunchKernelGGL((pair_kernel<Pol, TM>), dim3(grid), dim3(256), lds, s, p);
  RIMP2_HIP_CHECK(hipGetLastError());
}

}  // namespace

// Tile (virtual block width) per mode and shape.  Larger tiles raise MFMA
// operand reuse; smaller ones keep small fragments from idling CUs.
int choose_tile(int mode, int64_t nvirt, int64_t nocc) {
  (void)nocc;
  switch (mode) {
    case RIMP2_MODE_FP64: return nvirt <= 48 ? 32 : 64;
    case RIMP2_MODE_FP32:
    case RIMP2_MODE_BF16X3: return nvirt <= 96 ? 64 : 128;
    default: return 64;
  }
}

static void dispatch_pair(int mode, int tile, hipStream_t s, const PairParams& p) {
  switch (mode) {
    case RIMP2_MODE_FP64:
      if (tile == 32) return launch_pair<PolF64, 32>(s, p);
      if (tile == 64) return launch_pair<PolF64, 64>(s, p);
      break;
    case RIMP2_MODE_FP32:
      if (tile == 64) return launch_pair<PolF32, 64>(s, p);
      if (tile == 128) return launch_pair<PolF32, 128>(s, p);
      break;
    case RIMP2_MODE_BF16X3:
      if (tile == 64) return launch_pair<PolBF16x3, 64>(s, p);
      if (tile == 128) return launch_pair<PolBF16x3, 128>(s, p);
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

// Runs the pair kernel + reductions on ctx->stream.  d_pair_list (optional)
// restricts the work to the listed unique pairs (screening).
void pair_energies_device(Ctx* c, const PackGeom& g, const void* packed, const double* scales, int64_t nocc,
                          int64_t nvirt, const double* d_eps_o, const double* d_eps_v, double* d_eps_pair,
                          double* d_etot, const int64_t* d_pair_list, int64_t npairs_listed) {
  hipStream_t s = c->stream;
  const int tile = c->tile;
  const int64_t nb = g.nvp / tile;
  const int64_t ntp = nb * (nb + 1) / 2;
  const int64_t npairs = d_pair_list ? npairs_listed : nocc * (nocc + 1) / 2;
  double* evp = static_cast<double*>(c->eps_v.ensure((size_t)g.nvp * sizeof(double)));
  hipLaunchKernelGGL(pad_eps_v_kernel, dim3((unsigned)ceil_div(g.nvp, 256)), dim3(256), 0, s, d_eps_v, nvirt, g.nvp,
                     evp);
  RIMP2_HIP_CHECK(hipGetLastError());
  const size_t npart = (size_t)(npairs * ntp);
  double* part = static_cast<double*>(c->partials.ensure((npart + (size_t)npairs + 1) * sizeof(double)));
  double* contrib = part + npart;
  if (d_eps_pair) RIMP2_HIP_CHECK(hipMemsetAsync(d_eps_pair, 0, (size_t)(nocc * nocc) * sizeof(double), s));
  if (npairs > 0) {
    PairParams p;
    p.packed = static_cast<const unsigned char*>(packed);
    p.sigma = scales;
    p.eps_o = d_eps_o;
    p.eps_v = evp;
    p.partials = part;
    p.pair_list = d_pair_list;
    p.nocc = nocc;
    p.nvp = g.nvp;
    p.nks = g.nks;
    p.nrb = g.nrb;
    p.nb = nb;
    p.ntp = ntp;
    p.npairs = npairs;
    p.nwu = npairs * ntp;
    p.chunk = ceil_div(p.nwu, 8);
    {
      ScopedTimer t(c, RIMP2_K_PAIR);
      dispatch_pair(g.mode, tile, s, p);
    }
    ScopedTimer t(c, RIMP2_K_REDUCE);
    hipLaunchKernelGGL(pair_reduce_kernel, dim3((unsigned)ceil_div(npairs, 256)), dim3(256), 0, s, part, ntp, npairs,
                       d_pair_list, nocc, d_eps_pair, contrib);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(total_kernel, dim3(1), dim3(256), 0, s, contrib, npairs, d_etot);
  RIMP2_HIP_CHECK(hipGetLastError());
}

void check_inputs_device(Ctx* c, const double* p, int64_t n, const char* what) {
  (void)what;
  int* status = static_cast<int*>(c->status.ensure(64));
  hipLaunchKernelGGL(finite_check_kernel, dim3(256), dim3(256), 0, c->stream, p, n, status);
  RIMP2_HIP_CHECK(hipGetLastError());
}

void denom_check_device(Ctx* c, const double* eo, int64_t nocc, const double* ev, int64_t nvirt) {
  int* status = static_cast<int*>(c->status.ensure(64));
  hipLaunchKernelGGL(denom_check_kernel, dim3(1), dim3(64), 0, c->stream, eo, nocc, ev, nvirt, status);
  RIMP2_HIP_CHECK(hipGetLastError());
}

}  // namespace rimp2
