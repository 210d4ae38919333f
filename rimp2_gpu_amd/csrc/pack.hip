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

// Auxiliary padding quantum per mode: four LDS stages of the pair kernel (its
// break-free loop body runs four stage buffers against two register sets).
static int64_t q_quantum(int mode) {
  switch (mode) {
    case RIMP2_MODE_FP64: return 4 * 16;
    case RIMP2_MODE_FP32: return 4 * 16;
    case RIMP2_MODE_BF16X3: return 4 * 16;
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

// sigma[i*nvp + a] = max_Q |B[Q,a,i]| (0 for padding rows); one wave per row.
__global__ __launch_bounds__(256) void rowmax_kernel(const double* __restrict__ B, int64_t naux, int64_t nvirt,
                                                     int64_t nocc, int64_t nvp, double* __restrict__ sigma) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nvp * nocc) return;
  const int64_t i = row / nvp, a = row % nvp;
  double m = 0.0;
  if (a < nvirt) {
    const double* col = B + naux * (a + nvirt * i);
    for (int64_t q = lane; q < naux; q += 64) m = fmax(m, fabs(col[q]));
  }
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

// One packed element t (fragment f = t >> 6 = (i * nks + ks) * nrb + rb, lane
// t & 63) per mode; the split-BF16 and exact-slice modes write every plane of
// the fragment.
template <int MODE>
__device__ inline void pack_one(const PackArgs& p, int64_t t, int64_t nvp);

template <>
__device__ inline void pack_one<RIMP2_MODE_FP64>(const PackArgs& p, int64_t t, int64_t) {
  const int lane = (int)(t & 63);
  const int64_t f = t >> 6;
  const int64_t rb = f % p.nrb, ks = (f / p.nrb) % p.nks, i = f / (p.nrb * p.nks);
  const int64_t a = rb * 16 + (lane & 15), q = ks * 4 + (lane >> 4);
  static_cast<double*>(p.dst)[t] = bget(p, q, a, i);
}

template <>
__device__ inline void pack_one<RIMP2_MODE_FP32>(const PackArgs& p, int64_t t, int64_t) {
  const int lane = (int)(t & 63);
  const int64_t f = t >> 6;
  const int64_t rb = f % p.nrb, ks = (f / p.nrb) % p.nks, i = f / (p.nrb * p.nks);
  const int64_t a = rb * 32 + (lane & 31), q = ks * 2 + (lane >> 5);
  static_cast<float*>(p.dst)[t] = (float)bget(p, q, a, i);
}

// FP32 in 16x16x4 operand order (the fp64 fragment shape with 4-byte lanes).
constexpr int PACK_F32R16 = 16;  // packer id beside the mode numbers
template <>
__device__ inline void pack_one<PACK_F32R16>(const PackArgs& p, int64_t t, int64_t) {
  const int lane = (int)(t & 63);
  const int64_t f = t >> 6;
  const int64_t rb = f % p.nrb, ks = (f / p.nrb) % p.nks, i = f / (p.nrb * p.nks);
  const int64_t a = rb * 16 + (lane & 15), q = ks * 4 + (lane >> 4);
  static_cast<float*>(p.dst)[t] = (float)bget(p, q, a, i);
}

// bf16 round-to-nearest-even of a float (no NaN inputs reach here).
__device__ inline uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// planes: 0 = hi = bf16(b), 1 = lo = bf16(b - hi)
template <>
__device__ inline void pack_one<RIMP2_MODE_BF16X3>(const PackArgs& p, int64_t t, int64_t) {
  const int lane = (int)(t & 63);
  const int64_t f = t >> 6;
  const int64_t rb = f % p.nrb, ks = (f / p.nrb) % p.nks, i = f / (p.nrb * p.nks);
  const int64_t a = rb * 32 + (lane & 31), q0 = ks * 16 + 8 * (lane >> 5);
  uint16_t hi[8], lo[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    float b = (float)bget(p, q0 + e, a, i);
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
}

template <int S>
__device__ inline void pack_i8_one(const PackArgs& p, int64_t t, int64_t nvp) {
  const int lane = (int)(t & 63);
  const int64_t f = t >> 6;
  const int64_t rb = f % p.nrb, ks = (f / p.nrb) % p.nks, i = f / (p.nrb * p.nks);
  const int64_t a = rb * 32 + (lane & 31), q0 = ks * 32 + 16 * (lane >> 5);
  const double sg = p.sigma[i * nvp + a];
  const double inv = (sg > 0.0) ? 127.0 / sg : 0.0;
  int8_t sl[S][16];
#pragma unroll
  for (int e = 0; e < 16; e++) {
    double r = bget(p, q0 + e, a, i) * inv;
#pragma unroll
    for (int k = 0; k < S; k++) {
      double s = rint(r);
      s = fmin(127.0, fmax(-127.0, s));
      sl[k][e] = (int8_t)s;
      r = (r - s) * 254.0;
    }
  }
  uint4* dst = static_cast<uint4*>(p.dst);
#pragma unroll
  for (int k = 0; k < S; k++) {
    uint4 v;
    memcpy(&v, sl[k], 16);
    dst[(f * S + k) * 64 + lane] = v;
  }
}
template <>
__device__ inline void pack_one<RIMP2_MODE_I8S4>(const PackArgs& p, int64_t t, int64_t nvp) {
  pack_i8_one<4>(p, t, nvp);
}
template <>
__device__ inline void pack_one<RIMP2_MODE_I8S5>(const PackArgs& p, int64_t t, int64_t nvp) {
  pack_i8_one<5>(p, t, nvp);
}

// Exact-slice packing by tiles: one workgroup per (i, rb, 4 k-steps) = 32 rows
// x 128 auxiliary values.  The tile is read with coalesced loads (one row of
// 1 KB per wave instruction) into LDS, then thread (k-step f, lane) quantises the
// 16 values of its fragment element from LDS and writes the S planes (each
// wave: 1 KB per plane).  The per-element path reads 16 strided bytes from 64
// rows per load instead.
template <int S>
__global__ __launch_bounds__(256) void pack_i8_tile_kernel(PackArgs p, int64_t nvp) {
  __shared__ double tile[32][129];
  const int64_t kt = p.nks / 4, ntiles = p.nocc * kt * p.nrb;
  const int tid = threadIdx.x;
  for (int64_t u = blockIdx.x; u < ntiles; u += gridDim.x) {
    const int64_t rb = u % p.nrb, kq = (u / p.nrb) % kt, i = u / (p.nrb * kt);
    const int64_t q0 = kq * 128;
    {  // wave w reads rows 8w .. 8w+7, one row (128 q = 1 KB) per instruction, lane l q = 2l, 2l+1
      const int w = tid >> 6, l = tid & 63;
#pragma unroll
      for (int rr = 0; rr < 8; rr++) {
        const int row = 8 * w + rr;
        tile[row][2 * l] = bget(p, q0 + 2 * l, rb * 32 + row, i);
        tile[row][2 * l + 1] = bget(p, q0 + 2 * l + 1, rb * 32 + row, i);
      }
    }
    __syncthreads();
    const int f = tid >> 6, lane = tid & 63, row = lane & 31;
    const int64_t a = rb * 32 + row;
    const double sg = p.sigma[i * nvp + a];
    const double inv = (sg > 0.0) ? 127.0 / sg : 0.0;
    int8_t sl[S][16];
#pragma unroll
    for (int e = 0; e < 16; e++) {
      double x = tile[row][f * 32 + 16 * (lane >> 5) + e] * inv;
#pragma unroll
      for (int k = 0; k < S; k++) {
        double sv = rint(x);
        sv = fmin(127.0, fmax(-127.0, sv));
        sl[k][e] = (int8_t)sv;
        x = (x - sv) * 254.0;
      }
    }
    const int64_t frag = (i * p.nks + kq * 4 + f) * p.nrb + rb;
    uint4* dst = static_cast<uint4*>(p.dst);
#pragma unroll
    for (int k = 0; k < S; k++) {
      uint4 v;
      memcpy(&v, sl[k], 16);
      dst[(frag * S + k) * 64 + lane] = v;
    }
    __syncthreads();  // the tile is reused by the next iteration
  }
}

// one thread per (i, ks, rb, lane)
template <int MODE>
__global__ __launch_bounds__(256) void pack_kernel(PackArgs p, int64_t nvp) {
  RIMP2_GRID_STRIDE(t, p.nocc * p.nks * p.nrb * 64) pack_one<MODE>(p, t, nvp);
}

// Grouped packing of many fragments (one launch per step for a batch).  Job j
// owns the global items [first, next job's first); a chunk of 256 consecutive
// items finds its first job once (uniform), each thread steps on from there.
struct PackJob {
  PackArgs p;
  double* sigma;  // row scales written by rowmax_group_kernel (p.sigma reads them)
  int64_t nvp;
  int64_t first;
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

template <int MODE>
__global__ __launch_bounds__(256) void pack_group_kernel(const PackJob* __restrict__ jobs, int njobs, int64_t total) {
  for (int64_t base = (int64_t)blockIdx.x * 256; base < total; base += (int64_t)gridDim.x * 256) {
    int j = pack_job_find(jobs, njobs, base);
    const int64_t t = base + threadIdx.x;
    if (t >= total) continue;
    while (j + 1 < njobs && jobs[j + 1].first <= t) j++;
    pack_one<MODE>(jobs[j].p, t - jobs[j].first, jobs[j].nvp);
  }
}

// Row scales of the exact-slice modes for many fragments: one wave per row;
// jobs own rows [first, next first) (row = i * nvp + a).
__global__ __launch_bounds__(256) void rowmax_group_kernel(const PackJob* __restrict__ jobs, int njobs, int64_t total) {
  const int lane = threadIdx.x & 63;
  for (int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < total; row += (int64_t)gridDim.x * 4) {
    const PackJob& J = jobs[pack_job_find(jobs, njobs, row)];
    const int64_t r = row - J.first, i = r / J.nvp, a = r % J.nvp;
    double m = 0.0;
    if (a < J.p.nvirt) {
      const double* col = J.p.B + J.p.naux * (a + J.p.nvirt * i);
      for (int64_t q = lane; q < J.p.naux; q += 64) m = fmax(m, fabs(col[q]));
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

void pack_B(Ctx* c, const double* dB, int64_t naux, int64_t nvirt, int64_t nocc, const PackGeom& g, void* dst,
            double* dscales, hipStream_t s, unsigned max_blocks) {
  ScopedTimer tm(c, RIMP2_K_PACK);
  PackArgs p{dB, naux, nvirt, nocc, g.nks, g.nrb, dscales, dst};
  const int64_t threads = nocc * g.nks * g.nrb * 64;
  unsigned blocks = grid_blocks(threads, 256);
  if (max_blocks && blocks > max_blocks) blocks = max_blocks;  // grid-stride kernels
  switch (g.mode) {
    case RIMP2_MODE_FP64:
      hipLaunchKernelGGL(pack_kernel<RIMP2_MODE_FP64>, dim3(blocks), dim3(256), 0, s, p, g.nvp);
      break;
    case RIMP2_MODE_FP32:
      if (g.rb == 16) hipLaunchKernelGGL(pack_kernel<PACK_F32R16>, dim3(blocks), dim3(256), 0, s, p, g.nvp);
      else hipLaunchKernelGGL(pack_kernel<RIMP2_MODE_FP32>, dim3(blocks), dim3(256), 0, s, p, g.nvp);
      break;
    case RIMP2_MODE_BF16X3:
      hipLaunchKernelGGL(pack_kernel<RIMP2_MODE_BF16X3>, dim3(blocks), dim3(256), 0, s, p, g.nvp);
      break;
    case RIMP2_MODE_I8S4:
    case RIMP2_MODE_I8S5: {
      hipLaunchKernelGGL(rowmax_kernel, dim3((unsigned)ceil_div(g.nvp * nocc, 4)), dim3(256), 0, s, dB, naux, nvirt,
                         nocc, g.nvp, dscales);
      RIMP2_HIP_CHECK(hipGetLastError());
      // tiles of 4 k-steps (the auxiliary padding is a multiple of four stages of 32)
      if (g.nks % 4) fail(RIMP2_ERR_ARG, "exact-slice packing: %lld k-steps not a multiple of 4", (long long)g.nks);
      const unsigned tb = grid_blocks(nocc * (g.nks / 4) * g.nrb * 64, 64);  // tiles, capped
      if (g.mode == RIMP2_MODE_I8S4)
        hipLaunchKernelGGL(pack_i8_tile_kernel<4>, dim3(tb), dim3(256), 0, s, p, g.nvp);
      else
        hipLaunchKernelGGL(pack_i8_tile_kernel<5>, dim3(tb), dim3(256), 0, s, p, g.nvp);
      break;
    }
    default: fail(RIMP2_ERR_ARG, "unknown precision mode %d", g.mode);
  }
  RIMP2_HIP_CHECK(hipGetLastError());
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
    pj.push_back(j);
    ne += it.nocc * it.g.nks * it.g.nrb * 64;
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
    hipLaunchKernelGGL(rowmax_group_kernel, dim3(grid_blocks(nr * 64, 256)), dim3(256), 0, s,
                       (const PackJob*)(d + orj), nj, nr);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  const dim3 gb(grid_blocks(ne, 256));
  const PackJob* jp = (const PackJob*)(d + opj);
  switch (mode) {
    case RIMP2_MODE_FP64: hipLaunchKernelGGL(pack_group_kernel<RIMP2_MODE_FP64>, gb, dim3(256), 0, s, jp, nj, ne); break;
    case RIMP2_MODE_FP32:
      if (items[0].g.rb == 16) hipLaunchKernelGGL(pack_group_kernel<PACK_F32R16>, gb, dim3(256), 0, s, jp, nj, ne);
      else hipLaunchKernelGGL(pack_group_kernel<RIMP2_MODE_FP32>, gb, dim3(256), 0, s, jp, nj, ne);
      break;
    case RIMP2_MODE_BF16X3:
      hipLaunchKernelGGL(pack_group_kernel<RIMP2_MODE_BF16X3>, gb, dim3(256), 0, s, jp, nj, ne);
      break;
    case RIMP2_MODE_I8S4: hipLaunchKernelGGL(pack_group_kernel<RIMP2_MODE_I8S4>, gb, dim3(256), 0, s, jp, nj, ne); break;
    case RIMP2_MODE_I8S5: hipLaunchKernelGGL(pack_group_kernel<RIMP2_MODE_I8S5>, gb, dim3(256), 0, s, jp, nj, ne); break;
    default: fail(RIMP2_ERR_ARG, "unknown precision mode %d", mode);
  }
  RIMP2_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(pad_group_kernel, dim3(grid_blocks(np, 256)), dim3(256), 0, s, (const PadJob*)(d + odj), nj, np);
  RIMP2_HIP_CHECK(hipGetLastError());
}

}  // namespace rimp2
