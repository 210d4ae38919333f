// DEC local-basis 4-index arrays (toccEOS, tvirtEOS, goccEOS, gvirtEOS) and
// the general FP64 GEMM behind them.
//
// Reference: RIMP2.F90:67-170 (Major Steps 5-8, CPU) and RIMP2.GPU.F90:103-260
// (the same steps as OpenACC kernels + cublasDgemm_v2).  Every output is a
// 4-index transform of the canonical amplitude or integral:
//   t(I,J,A,B) = g(A,I,B,J) / (e_I + e_J - e_B - e_A),  g = sum_alpha C(alpha,A,I) C(alpha,B,J)
//   toccEOS(Al,Il,Bl,Jl) = sum UvirtT(A,Al) UoccEOST(I,Il) UvirtT(B,Bl) UoccEOST(J,Jl) t(I,J,A,B)
//   tvirtEOS(Al,Il,Bl,Jl) = sum UvirtEOST(A,Al) UoccT(I,Il) UvirtEOST(B,Bl) UoccT(J,Jl) t(I,J,A,B)
//   goccEOS / gvirtEOS: the same transforms of g, formed as C3^T C3 of the
//   transformed 3-index tensor (RIMP2.F90:136-170).
//
// Order of work here (not the reference's): both amplitude arrays share one
// t, built chunk by chunk over I (t never exists whole) for each unordered
// occupied pair once (t(A,B,J,I) = t(B,A,I,J)) by the FP64 pair kernel in its
// amplitude form (pair.hip: the energy kernel's operand pipeline, with an
// epilogue that writes w V / Delta); from each chunk the occupied pair of
// toccEOS and the virtual pair of tvirtEOS are contracted at once (the index
// of each array that shrinks most goes first), the remaining two indices of
// each are transformed after the last chunk, and a symmetrization restores
// the pairs J < I.  Every other step is one launch of the FP64 MFMA GEMM
// below with two-level batch strides, so the reorders of toccB / tvirtB
// (RIMP2.F90:456-508) are just output strides.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pair_core.h"

namespace rimp2 {

struct EosState {
  DevBuf in, work, out, fo, pk, sg;
};
void eos_destroy(EosState* s) {
  if (!s) return;
  s->in.release();
  s->work.release();
  s->out.release();
  s->fo.release();
  s->pk.release();
  s->sg.release();
  delete s;
}

namespace {

using v4d = double __attribute__((ext_vector_type(4)));

// C[z] = alpha op(A[z]) op(B[z]) + beta C[z] for batch members z = z1 + Z1 z2,
// member offsets z1 * s?1 + z2 * s?2 (elements).
struct DgemmArgs {
  int64_t M, N, K;
  double alpha, beta;
  const double* A;
  int64_t lda, sA1, sA2;
  const double* B;
  int64_t ldb, sB1, sB2;
  double* C;
  int64_t ldc, sC1, sC2;
  int64_t Z1, zbase;
  int64_t tm, tiles;  // row tiles, tiles per member
  int a_vec, b_vec;   // 16-byte loads allowed (even strides, aligned bases)
  int lower;          // symmetric product: tiles strictly above the diagonal are skipped
};

// 256 threads = 4 waves in 2x2, wave tile (BM/2)x(BN/2) of 16x16 MFMA blocks
// (v_mfma_f64_16x16x4f64; lane l holds A[l%16][l/16], B[l/16][l%16], result
// d[r] = D[4r + l/16][l%16]).  K tiles go through LDS with the next
// tile fetched into registers while the MFMAs run.  LDS layouts keep every
// fragment read free of bank conflicts: a stored-untransposed A as As[k][m]
// (pitch BM+16), a transposed A as As[m][k] (pitch KP); B likewise.
// At least two waves per SIMD: the compiler then keeps the accumulators of
// the larger tiles in VGPRs with no spills (measured 70 ms against 99 ms for
// the one-wave build, 76 ms for 64x64, on a batch of eight 1041 x 103059 x
// 2173 products).
// Tiles are numbered row tile fastest and laid out so that each XCD (blocks
// b with the same b % 8) runs one contiguous range of them: the row tiles
// that share a column panel of op(B) share that XCD's L2.
template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void dgemm_kernel(const DgemmArgs p) {
  // K tile: 32 on the 96x128 tile (twice the loads in flight per barrier:
  // -2.9 % against 16 on the products above), 16 elsewhere (the 128x128
  // tile spills at 32)
  constexpr int BK = (BM == 96) ? 32 : 16;
  // pitch of the k-contiguous LDS layouts: a half-wave reads 16 rows at k and
  // 16 at k + 1; 17 (BK 16) and 34 (BK 32) doubles put them on disjoint banks
  constexpr int KP = BK == 32 ? BK + 2 : BK + 1;
  constexpr int PA = TA ? KP : BM + 16;
  constexpr int PB = TB ? BN + 16 : KP;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int NA = BM * BK / 512, NB = BN * BK / 512;
  __shared__ double As[TA ? BM * PA : BK * PA];
  __shared__ double Bs[TB ? BK * PB : BN * PB];

  const int64_t G = gridDim.x;  // a multiple of 8
  const int64_t b = blockIdx.x;
  const int64_t t = (b & 7) * (G >> 3) + (b >> 3);
  if (t >= p.tiles) return;
  const int64_t z = p.zbase + blockIdx.y;
  const int64_t z1 = z % p.Z1, z2 = z / p.Z1;
  const double* __restrict__ A = p.A + z1 * p.sA1 + z2 * p.sA2;
  const double* __restrict__ B = p.B + z1 * p.sB1 + z2 * p.sB2;
  double* __restrict__ C = p.C + z1 * p.sC1 + z2 * p.sC2;
  const int64_t M = p.M, N = p.N, K = p.K, lda = p.lda, ldb = p.ldb;
  const int64_t m0 = (t % p.tm) * BM, n0 = (t / p.tm) * BN;
  if (p.lower && n0 > m0 + BM - 1) return;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;

  v4d acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  double2 ra[NA], rb[NB];
  // element (m, k) of op(A): A[m + lda k] (TA: A[k + lda m]); pairs run along
  // the contiguous index
  auto load_a = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < NA; s++) {
      const int e = tid + 256 * s;
      double2 v = {0.0, 0.0};
      if (!TA) {
        const int k = e / (BM / 2), m = (e % (BM / 2)) * 2;
        const int64_t gm = m0 + m, gk = k0 + k;
        if (gk < K) {
          const double* q = A + gm + lda * gk;
          if (p.a_vec && gm + 1 < M) v = *reinterpret_cast<const double2*>(q);
          else {
            if (gm < M) v.x = q[0];
            if (gm + 1 < M) v.y = q[1];
          }
        }
      } else {
        const int m = e / (BK / 2), k = (e % (BK / 2)) * 2;
        const int64_t gm = m0 + m, gk = k0 + k;
        if (gm < M) {
          const double* q = A + gk + lda * gm;
          if (p.a_vec && gk + 1 < K) v = *reinterpret_cast<const double2*>(q);
          else {
            if (gk < K) v.x = q[0];
            if (gk + 1 < K) v.y = q[1];
          }
        }
      }
      ra[s] = v;
    }
  };
  // element (k, n) of op(B): B[k + ldb n] (TB: B[n + ldb k])
  auto load_b = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < NB; s++) {
      const int e = tid + 256 * s;
      double2 v = {0.0, 0.0};
      if (!TB) {
        const int n = e / (BK / 2), k = (e % (BK / 2)) * 2;
        const int64_t gn = n0 + n, gk = k0 + k;
        if (gn < N) {
          const double* q = B + gk + ldb * gn;
          if (p.b_vec && gk + 1 < K) v = *reinterpret_cast<const double2*>(q);
          else {
            if (gk < K) v.x = q[0];
            if (gk + 1 < K) v.y = q[1];
          }
        }
      } else {
        const int k = e / (BN / 2), n = (e % (BN / 2)) * 2;
        const int64_t gn = n0 + n, gk = k0 + k;
        if (gk < K) {
          const double* q = B + gn + ldb * gk;
          if (p.b_vec && gn + 1 < N) v = *reinterpret_cast<const double2*>(q);
          else {
            if (gn < N) v.x = q[0];
            if (gn + 1 < N) v.y = q[1];
          }
        }
      }
      rb[s] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int s = 0; s < NA; s++) {
      const int e = tid + 256 * s;
      if (!TA) {
        const int k = e / (BM / 2), m = (e % (BM / 2)) * 2;
        As[k * PA + m] = ra[s].x;
        As[k * PA + m + 1] = ra[s].y;
      } else {
        const int m = e / (BK / 2), k = (e % (BK / 2)) * 2;
        As[m * PA + k] = ra[s].x;
        As[m * PA + k + 1] = ra[s].y;
      }
    }
#pragma unroll
    for (int s = 0; s < NB; s++) {
      const int e = tid + 256 * s;
      if (!TB) {
        const int n = e / (BK / 2), k = (e % (BK / 2)) * 2;
        Bs[n * PB + k] = rb[s].x;
        Bs[n * PB + k + 1] = rb[s].y;
      } else {
        const int k = e / (BN / 2), n = (e % (BN / 2)) * 2;
        Bs[k * PB + n] = rb[s].x;
        Bs[k * PB + n + 1] = rb[s].y;
      }
    }
  };

  if (K > 0) {
    load_a(0);
    load_b(0);
  }
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    store();
    __syncthreads();
    if (k0 + BK < K) {
      load_a(k0 + BK);
      load_b(k0 + BK);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kq = ks + (lane >> 4);
      double af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) {
        const int m = wm + 16 * i + (lane & 15);
        af[i] = TA ? As[m * PA + kq] : As[kq * PA + m];
      }
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int n = wn + 16 * j + (lane & 15);
        bf[j] = TB ? Bs[kq * PB + n] : Bs[n * PB + kq];
      }
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++) {
      const int64_t gn = n0 + wn + 16 * j + (lane & 15);
      if (gn >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t gm = m0 + wm + 16 * i + 4 * r + (lane >> 4);
        if (gm >= M) continue;
        double* q = C + gm + p.ldc * gn;
        double v = p.alpha * acc[i][j][r];
        if (p.beta != 0.0) v += p.beta * *q;
        *q = v;
      }
    }
}

// Host description of one batched GEMM (element strides).
struct Gemm {
  bool ta = false, tb = false;
  int64_t M = 0, N = 0, K = 0;
  double alpha = 1.0, beta = 0.0;
  const double* A = nullptr;
  int64_t lda = 1, sA1 = 0, sA2 = 0;
  const double* B = nullptr;
  int64_t ldb = 1, sB1 = 0, sB2 = 0;
  double* C = nullptr;
  int64_t ldc = 1, sC1 = 0, sC2 = 0;
  int64_t Z1 = 1, Z2 = 1;
  bool lower = false;  // C symmetric: only tiles touching the lower triangle
};

bool even_aligned(const double* base, int64_t ld, int64_t s1, int64_t s2) {
  return (reinterpret_cast<uintptr_t>(base) & 15) == 0 && (ld & 1) == 0 && (s1 & 1) == 0 && (s2 & 1) == 0;
}

template <int BM, int BN>
void launch_tiles(hipStream_t s, const Gemm& g, DgemmArgs a) {
  a.tm = ceil_div(g.M, BM);
  a.tiles = a.tm * ceil_div(g.N, BN);
  const int64_t gx = round_up(a.tiles, 8);
  // one dispatch counts blocks x 256 work-items in 32 bits; grid.y <= 65535
  if (gx > kMaxGridBlocks) fail(RIMP2_ERR_ARG, "GEMM of %lld x %lld exceeds the launch limit", (long long)g.M,
                                (long long)g.N);
  const int64_t zmax = std::max<int64_t>(1, std::min<int64_t>(65535, kMaxGridBlocks / gx));
  const int64_t Z = g.Z1 * g.Z2;
  for (int64_t z0 = 0; z0 < Z; z0 += zmax) {
    a.zbase = z0;
    const dim3 grid((unsigned)gx, (unsigned)std::min<int64_t>(zmax, Z - z0));
    if (!g.ta && !g.tb) hipLaunchKernelGGL((dgemm_kernel<BM, BN, false, false>), grid, dim3(256), 0, s, a);
    else if (g.ta && !g.tb) hipLaunchKernelGGL((dgemm_kernel<BM, BN, true, false>), grid, dim3(256), 0, s, a);
    else if (!g.ta && g.tb) hipLaunchKernelGGL((dgemm_kernel<BM, BN, false, true>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((dgemm_kernel<BM, BN, true, true>), grid, dim3(256), 0, s, a);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
}

// Block tile of one GEMM.  An FP64 64x64 tile moves 8 flops per operand byte
// through L2; 128x128 doubles that, and 96x128 (13.7) pads an awkward M less
// (M = 1041: 1.4 % against 10.7 % at 128).  Small or skinny problems keep
// 64x64 so that their grids stay large.
enum TileShape { TILE_64x64, TILE_96x128, TILE_128x128 };
TileShape pick_tile(const Gemm& g) {
  if (g.N < 128 || g.M < 96 || g.K < 64) return TILE_64x64;
  const double w128 = (double)round_up(g.M, 128) / (double)g.M, w96 = (double)round_up(g.M, 96) / (double)g.M;
  return (w96 < w128 - 0.03) ? TILE_96x128 : TILE_128x128;
}

void run_gemm(hipStream_t s, const Gemm& g) {
  if (g.M <= 0 || g.N <= 0 || g.Z1 <= 0 || g.Z2 <= 0) return;
  DgemmArgs a;
  a.M = g.M;
  a.N = g.N;
  a.K = g.K;
  a.alpha = g.alpha;
  a.beta = g.beta;
  a.A = g.A;
  a.lda = g.lda;
  a.sA1 = g.sA1;
  a.sA2 = g.sA2;
  a.B = g.B;
  a.ldb = g.ldb;
  a.sB1 = g.sB1;
  a.sB2 = g.sB2;
  a.C = g.C;
  a.ldc = g.ldc;
  a.sC1 = g.sC1;
  a.sC2 = g.sC2;
  a.Z1 = g.Z1;
  a.zbase = 0;
  a.lower = g.lower ? 1 : 0;
  a.a_vec = even_aligned(g.A, g.lda, g.sA1, g.sA2);
  a.b_vec = even_aligned(g.B, g.ldb, g.sB1, g.sB2);
  switch (pick_tile(g)) {
    case TILE_64x64: launch_tiles<64, 64>(s, g, a); break;
    case TILE_96x128: launch_tiles<96, 128>(s, g, a); break;
    case TILE_128x128: launch_tiles<128, 128>(s, g, a); break;
  }
}

// C[m + ldc n] = C[n + ldc m] for n > m (n x n), through 32 x 32 LDS tiles:
// block (bi, bj), bi < bj, reads tile (bj, bi) of the lower triangle row-
// contiguously and writes its transpose to tile (bi, bj).
__global__ __launch_bounds__(256) void mirror_lower_kernel(double* C, int64_t n, int64_t ldc) {
  __shared__ double tile[32][33];
  const int64_t nt = (n + 31) / 32;
  int64_t bi, bj;
  pairk::tri_decode((int64_t)blockIdx.x, bi, bj);  // bi <= bj
  if (bi == bj || bj >= nt) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {  // source: rows bj*32 + tx, cols bi*32 + r  (lower: row > col)
    const int64_t row = bj * 32 + tx, col = bi * 32 + r;
    if (row < n && col < n) tile[r][tx] = C[row + ldc * col];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {  // destination: rows bi*32 + tx, cols bj*32 + r
    const int64_t row = bi * 32 + tx, col = bj * 32 + r;
    if (row < n && col < n) C[row + ldc * col] = tile[tx][r];
  }
}

__global__ void zero_kernel(double* x, int64_t n) {
  RIMP2_GRID_STRIDE(t, n) x[t] = 0.0;
}

// In place: x(Al,Il,Bl,Jl) <- x(Al,Il,Bl,Jl) + x(Bl,Jl,Al,Il) over an
// [nv, no, nv, no] array (each pair of elements read and written by one thread).
__global__ void sym_pairs_kernel(double* x, int64_t nv, int64_t no) {
  const int64_t n1 = nv * no, n = n1 * n1;
  RIMP2_GRID_STRIDE(t, n) {
    const int64_t u = t % n1, v = t / n1;  // u = (Al, Il), v = (Bl, Jl)
    const int64_t q = v + n1 * u;
    if (t < q) {
      const double a = x[t], b = x[q];
      x[t] = a + b;
      x[q] = a + b;
    } else if (t == q) {
      x[t] = 2.0 * x[t];
    }
  }
}

// y(Al,Bl,J,I) = x(Al,Bl,J,I) + x(Bl,Al,I,J) over [vE, vE, no, no] arrays.
__global__ void sym_swap_kernel(const double* __restrict__ x, double* __restrict__ y, int64_t vE, int64_t no) {
  const int64_t n = vE * vE * no * no;
  RIMP2_GRID_STRIDE(t, n) {
    const int64_t al = t % vE, bl = (t / vE) % vE, j = (t / (vE * vE)) % no, i = t / (vE * vE * no);
    y[t] = x[t] + x[bl + vE * (al + vE * (i + no * j))];
  }
}

// Sizes (elements) of the work arena of the EOS pipeline with ci occupied
// orbitals per amplitude chunk.
struct EosPlan {
  int64_t nv, no, nE, vE, nba, ci;
  int64_t T, P, S, Q, R, c2, c3;  // element counts
  size_t off[8];                  // byte offsets: T P S S' Q R R' c2 (c3 after)
  size_t bytes;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// p_all: the toccEOS partial P of every occupied orbital is kept (one I
// contraction at the end); else P holds one chunk and S accumulates per chunk.
EosPlan eos_plan(int64_t nv, int64_t no, int64_t nE, int64_t vE, int64_t nba, int64_t ci, bool p_all) {
  EosPlan p;
  p.nv = nv;
  p.no = no;
  p.nE = nE;
  p.vE = vE;
  p.nba = nba;
  p.ci = ci;
  p.T = nv * nv * no * ci;
  p.P = nv * nv * nE * (p_all ? no : ci);
  p.S = nv * nv * nE * nE;
  p.Q = vE * nv * no * ci;
  p.R = vE * vE * no * no;
  p.c2 = nba * nv * std::max(nE, no);
  p.c3 = nba * std::max(nv * nE, vE * no);
  const int64_t sizes[9] = {p.T, p.P, p.S, p.S, p.Q, p.R, p.R, p.c2, p.c3};
  size_t o = 0;
  for (int i = 0; i < 9; i++) {
    if (i < 8) p.off[i] = o;
    o += align256((size_t)sizes[i] * sizeof(double));
  }
  p.bytes = o;
  return p;
}

// Chunk of occupied orbitals per amplitude slab: the t slab (nv^2 no doubles
// per orbital) and its tvirtEOS partial transform stay within 8 GiB.
int64_t eos_chunk(int64_t nv, int64_t no, int64_t nE, int64_t vE, bool p_all) {
  const double per = (double)(nv * nv * no + (p_all ? 0 : nv * nv * nE) + vE * nv * no) * sizeof(double);
  const double budget = 8.0 * (1ull << 30);
  int64_t ci = (int64_t)(budget / per);
  return std::max<int64_t>(1, std::min<int64_t>(no, ci));
}

}  // namespace

// Queues Steps 5-8 on c->stream (device arrays; outputs in the reference's
// layouts, RIMP2.F90:92,123,146,169).
void eos_arrays_device(Ctx* c, int64_t nv, int64_t no, int64_t nE, int64_t vE, int64_t nba, const double* Cal,
                       const double* eo, const double* ev, const double* Uo5, const double* Uv5, const double* Uo6,
                       const double* Uv6, double* toccEOS, double* tvirtEOS, double* goccEOS, double* gvirtEOS,
                       int64_t ci) {
  if (!c->eos) c->eos = new EosState();
  // Automatic chunking keeps P whole when it fits 32 GiB (nv^2 nE no doubles);
  // an explicit chunk bounds the work memory, so P then holds one chunk.
  const bool p_all = ci <= 0 && (double)(nv * nv * nE * no) * sizeof(double) <= 32.0 * (1ull << 30);
  if (ci <= 0) ci = eos_chunk(nv, no, nE, vE, p_all);
  ci = std::min(ci, no);
  const EosPlan pl = eos_plan(nv, no, nE, vE, nba, ci, p_all);
  unsigned char* w = static_cast<unsigned char*>(c->eos->work.ensure(pl.bytes));
  double* T = reinterpret_cast<double*>(w + pl.off[0]);
  double* P = reinterpret_cast<double*>(w + pl.off[1]);
  double* S = reinterpret_cast<double*>(w + pl.off[2]);
  double* S2 = reinterpret_cast<double*>(w + pl.off[3]);
  double* Q = reinterpret_cast<double*>(w + pl.off[4]);
  double* R = reinterpret_cast<double*>(w + pl.off[5]);
  double* R2 = reinterpret_cast<double*>(w + pl.off[6]);
  double* c2 = reinterpret_cast<double*>(w + pl.off[7]);
  double* c3 = reinterpret_cast<double*>(w + pl.off[7] + align256((size_t)pl.c2 * sizeof(double)));
  hipStream_t s = c->stream;
  ScopedTimer tm(c, RIMP2_K_EOS);
  const int64_t nv2 = nv * nv;

  // The 3-index tensor in the pair kernel's operand order (64-wide virtual
  // blocks), once per call: FP64, or exact-slice int8 with row scales (int32
  // slice sums exact for NBA <= 16384: 5 * 127^2 * 16511 < 2^31).
  if (c->eos_mode == RIMP2_MODE_I8S5 && nba > 16384)
    fail(RIMP2_ERR_ARG, "exact-slice amplitude slab supports NBA <= 16384 (int32 accumulation bound)");
  const PackGeom pg = make_geom(c->eos_mode, nba, nv, 64);
  void* pk = c->eos->pk.ensure(pg.bytes_per_occ() * (size_t)no);
  double* sigma = c->eos_mode == RIMP2_MODE_I8S5
                      ? static_cast<double*>(c->eos->sg.ensure((size_t)(no * pg.nvp) * sizeof(double)))
                      : nullptr;
  pack_B(c, Cal, nba, nv, no, pg, pk, sigma, s);
  // Each unordered occupied pair once: t(A,B,J,I) = t(B,A,I,J), so a chunk
  // of I builds J >= i0 only, weighted by w_IJ (1 above the diagonal, 1/2 on
  // it, 0 below).  Any 4-index transform X of the weighted half satisfies
  // full(Al,Bl,J,I) = X(Al,Bl,J,I) + X(Bl,Al,I,J), applied at the end.
  {
    const int64_t nR = vE * vE * no * no;
    hipLaunchKernelGGL(zero_kernel, dim3(grid_blocks(nR, 256)), dim3(256), 0, s, R, nR);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  for (int64_t i0 = 0; i0 < no; i0 += ci) {
    const int64_t cz = std::min(ci, no - i0), nj = no - i0;
    {  // t(A, B, J - i0, I - i0), J >= I:  w V^{IJ}_AB / Delta on the FP64 pair kernel   (RIMP2.F90:399-404)
      ScopedTimer ta(c, RIMP2_K_EOS_AMPL);
      for (int64_t z = 1; z < cz; z++) {  // J in [i0, I): weight 0
        const int64_t n = nv2 * z;
        hipLaunchKernelGGL(zero_kernel, dim3(grid_blocks(n, 256)), dim3(256), 0, s, T + nv2 * nj * z, n);
        RIMP2_HIP_CHECK(hipGetLastError());
      }
      pair_amplitudes_device(s, pg, pk, sigma, no, nv, eo, ev, i0, cz, T);
    }
    {  // toccEOS, occupied J: P(A,B,Jl,I) = sum_J t(A,B,J,I) UoccEOST(J,Jl)
      Gemm g;
      g.M = nv2;
      g.N = nE;
      g.K = nj;
      g.A = T;
      g.lda = nv2;
      g.sA1 = nv2 * nj;
      g.B = Uo5 + i0;
      g.ldb = no;
      g.C = p_all ? P + nv2 * nE * i0 : P;
      g.ldc = nv2;
      g.sC1 = nv2 * nE;
      g.Z1 = cz;
      run_gemm(s, g);
    }
    if (!p_all) {  // toccEOS, occupied I: S(A,B,Jl,Il) += sum_I P(A,B,Jl,I) UoccEOST(I,Il)
      Gemm g;
      g.M = nv2 * nE;
      g.N = nE;
      g.K = cz;
      g.A = P;
      g.lda = nv2 * nE;
      g.B = Uo5 + i0;
      g.ldb = no;
      g.beta = i0 > 0 ? 1.0 : 0.0;
      g.C = S;
      g.ldc = nv2 * nE;
      run_gemm(s, g);
    }
    {  // tvirtEOS, virtual A: Q(Al, B, J, I) = sum_A UvirtEOST(A,Al) t(A,B,J,I)
      Gemm g;
      g.ta = true;
      g.M = vE;
      g.N = nv * nj * cz;
      g.K = nv;
      g.A = Uv6;
      g.lda = nv;
      g.B = T;
      g.ldb = nv;
      g.C = Q;
      g.ldc = vE;
      run_gemm(s, g);
    }
    {  // tvirtEOS, virtual B: R(Al,Bl,J,I) = sum_B Q(Al,B,J,I) UvirtEOST(B,Bl)
      Gemm g;
      g.M = vE;
      g.N = vE;
      g.K = nv;
      g.A = Q;
      g.lda = vE;
      g.sA1 = vE * nv;
      g.sA2 = vE * nv * nj;
      g.B = Uv6;
      g.ldb = nv;
      g.C = R + vE * vE * (i0 + no * i0);
      g.ldc = vE;
      g.sC1 = vE * vE;
      g.sC2 = vE * vE * no;
      g.Z1 = nj;
      g.Z2 = cz;
      run_gemm(s, g);
    }
  }
  if (p_all) {  // toccEOS, occupied I: S(A,B,Jl,Il) = sum_I P(A,B,Jl,I) UoccEOST(I,Il)
    Gemm g;
    g.M = nv2 * nE;
    g.N = nE;
    g.K = no;
    g.A = P;
    g.lda = nv2 * nE;
    g.B = Uo5;
    g.ldb = no;
    g.C = S;
    g.ldc = nv2 * nE;
    run_gemm(s, g);
  }
  {  // toccEOS, virtual A: S2(Al,B,Jl,Il) = sum_A UvirtT(A,Al) S(A,B,Jl,Il)
    Gemm g;
    g.ta = true;
    g.M = nv;
    g.N = nv * nE * nE;
    g.K = nv;
    g.A = Uv5;
    g.lda = nv;
    g.B = S;
    g.ldb = nv;
    g.C = S2;
    g.ldc = nv;
    run_gemm(s, g);
  }
  {  // X(Al,Bl,Jl,Il) = sum_B S2(Al,B,Jl,Il) UvirtT(B,Bl), stored at toccEOS(Al,Il,Bl,Jl)
    Gemm g;
    g.M = nv;
    g.N = nv;
    g.K = nv;
    g.A = S2;
    g.lda = nv;
    g.sA1 = nv2;
    g.sA2 = nv2 * nE;
    g.B = Uv5;
    g.ldb = nv;
    g.C = toccEOS;
    g.ldc = nv * nE;
    g.sC1 = nv2 * nE;
    g.sC2 = nv;
    g.Z1 = nE;
    g.Z2 = nE;
    run_gemm(s, g);
  }
  {  // toccEOS(Al,Il,Bl,Jl) = X(Al,Bl,Jl,Il) + X(Bl,Al,Il,Jl)   (toccB reorder, RIMP2.F90:456-481)
    const int64_t n = nv2 * nE * nE;
    hipLaunchKernelGGL(sym_pairs_kernel, dim3(grid_blocks(n, 256)), dim3(256), 0, s, toccEOS, nv, nE);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  {  // R2(Al,Bl,J,I) = R(Al,Bl,J,I) + R(Bl,Al,I,J): the full virtual transform of every ordered pair
    const int64_t n = vE * vE * no * no;
    hipLaunchKernelGGL(sym_swap_kernel, dim3(grid_blocks(n, 256)), dim3(256), 0, s, R, R2, vE, no);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  {  // tvirtEOS, occupied I: R(Al,Bl,J,Il) = sum_I R2(Al,Bl,J,I) UoccT(I,Il)
    Gemm g;
    g.M = vE * vE * no;
    g.N = no;
    g.K = no;
    g.A = R2;
    g.lda = vE * vE * no;
    g.B = Uo6;
    g.ldb = no;
    g.C = R;
    g.ldc = vE * vE * no;
    run_gemm(s, g);
  }
  {  // tvirtEOS(Al,Il,Bl,Jl) = sum_J R(Al,Bl,J,Il) UoccT(J,Jl)   (tvirtB reorder, RIMP2.F90:483-508)
    Gemm g;
    g.M = vE;
    g.N = no;
    g.K = no;
    g.A = R;
    g.lda = vE * vE;
    g.sA1 = vE;
    g.sA2 = vE * vE * no;
    g.B = Uo6;
    g.ldb = no;
    g.C = tvirtEOS;
    g.ldc = vE * no * vE;
    g.sC1 = vE * no;
    g.sC2 = vE;
    g.Z1 = vE;
    g.Z2 = no;
    run_gemm(s, g);
  }
  // Step 7 (RIMP2.F90:136-147): goccEOS = C3^T C3 with
  // C3(alpha,Bl,Jl) = sum_B UvirtT(B,Bl) sum_J C(alpha,B,J) UoccEOST(J,Jl)
  // Step 8 (RIMP2.F90:158-170): gvirtEOS likewise with UoccT and UvirtEOST.
  for (int step = 7; step <= 8; step++) {
    const int64_t nO = step == 7 ? nE : no, nV = step == 7 ? nv : vE;
    const double* Uo = step == 7 ? Uo5 : Uo6;
    const double* Uv = step == 7 ? Uv5 : Uv6;
    double* out = step == 7 ? goccEOS : gvirtEOS;
    {
      Gemm g;
      g.M = nba * nv;
      g.N = nO;
      g.K = no;
      g.A = Cal;
      g.lda = nba * nv;
      g.B = Uo;
      g.ldb = no;
      g.C = c2;
      g.ldc = nba * nv;
      run_gemm(s, g);
    }
    {
      Gemm g;
      g.M = nba;
      g.N = nV;
      g.K = nv;
      g.A = c2;
      g.lda = nba;
      g.sA1 = nba * nv;
      g.B = Uv;
      g.ldb = nv;
      g.C = c3;
      g.ldc = nba;
      g.sC1 = nba * nV;
      g.Z1 = nO;
      run_gemm(s, g);
    }
    {  // C3^T C3 is symmetric: the lower tiles, then the mirror
      Gemm g;
      g.ta = true;
      g.lower = true;
      g.M = nV * nO;
      g.N = nV * nO;
      g.K = nba;
      g.A = c3;
      g.lda = nba;
      g.B = c3;
      g.ldb = nba;
      g.C = out;
      g.ldc = nV * nO;
      run_gemm(s, g);
      const int64_t n = nV * nO, nt = ceil_div(n, 32);
      hipLaunchKernelGGL(mirror_lower_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, s, out, n, n);
      RIMP2_HIP_CHECK(hipGetLastError());
    }
  }
}

// First-order intermediates (RIMP2.F90:172-251), with the reference's own
// call arguments: UoccT multiplies CalphaOcc as an nocctot x nocctot matrix
// (nocctot = nocc, checked by the caller), and CalphaVV is multiplied by
// UvirtT with K = ldb = nocc (RIMP2.F90:231-235): the first nocc virtual
// columns of CalphaVV against UvirtT's storage read as an nocc x nvirt matrix.
//   djik(a,Il,Jl,K) = sum_alpha C3(alpha,a,Il) C4(alpha,Jl,K)   (RIMP2_calc_gen4DimFO, RIMP2.F90:616-641)
//   blad(Al,J,Bl,b) = sum_alpha B3(alpha,Al,J) B4(alpha,Bl,b)
void first_order_device(Ctx* c, int64_t nv, int64_t no, int64_t nE, int64_t vE, int64_t nba, const double* Cal,
                        const double* Cocc, const double* Cvv, const double* Uo5, const double* Uv5,
                        const double* Uo6, const double* Uv6, double* djik, double* blad) {
  if (!c->eos) c->eos = new EosState();
  const int64_t nt = no;
  const int64_t n_c2 = nba * nv * std::max(nE, no), n_c3 = nba * std::max(nv * nE, vE * no);
  const int64_t n_o2 = nba * no * nt, n_c4 = nba * nE * nt, n_v2 = nba * vE * no, n_b4 = nba * vE * nv;
  const size_t o_c3 = align256((size_t)n_c2 * 8), o_o2 = o_c3 + align256((size_t)n_c3 * 8),
               o_c4 = o_o2 + align256((size_t)n_o2 * 8), o_v2 = o_c4 + align256((size_t)n_c4 * 8),
               o_b4 = o_v2 + align256((size_t)n_v2 * 8), total = o_b4 + align256((size_t)n_b4 * 8);
  unsigned char* w = static_cast<unsigned char*>(c->eos->fo.ensure(total));
  double* c2 = reinterpret_cast<double*>(w);
  double* c3 = reinterpret_cast<double*>(w + o_c3);
  double* o2 = reinterpret_cast<double*>(w + o_o2);
  double* c4 = reinterpret_cast<double*>(w + o_c4);
  double* v2 = reinterpret_cast<double*>(w + o_v2);
  double* b4 = reinterpret_cast<double*>(w + o_b4);
  hipStream_t s = c->stream;
  ScopedTimer tm(c, RIMP2_K_EOS);
  auto gemm = [&](bool ta, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, int64_t sA1,
                  const double* B, int64_t ldb, double* C, int64_t ldc, int64_t sC1, int64_t Z1) {
    Gemm g;
    g.ta = ta;
    g.M = M;
    g.N = N;
    g.K = K;
    g.A = A;
    g.lda = lda;
    g.sA1 = sA1;
    g.B = B;
    g.ldb = ldb;
    g.C = C;
    g.ldc = ldc;
    g.sC1 = sC1;
    g.Z1 = Z1;
    run_gemm(s, g);
  };
  // djik: C3(alpha,Bl,Jl) = sum_B UvirtT(B,Bl) sum_J Calpha(alpha,B,J) UoccEOST(J,Jl)   (RIMP2.F90:181-187)
  gemm(false, nba * nv, nE, no, Cal, nba * nv, 0, Uo5, no, c2, nba * nv, 0, 1);
  gemm(false, nba, nv, nv, c2, nba, nba * nv, Uv5, nv, c3, nba, nba * nv, nE);
  // C4(alpha,Il,K) = sum_i UoccEOST(i,Il) sum_k CalphaOcc(alpha,i,k) UoccT(k,K)   (RIMP2.F90:193-205)
  gemm(false, nba * no, nt, nt, Cocc, nba * no, 0, Uo6, nt, o2, nba * no, 0, 1);
  gemm(false, nba, nE, no, o2, nba, nba * no, Uo5, no, c4, nba, nba * nE, nt);
  gemm(true, nv * nE, nE * nt, nba, c3, nba, 0, c4, nba, djik, nv * nE, 0, 1);
  // blad: B3(alpha,Al,J) = sum_B UvirtEOST(B,Al) sum_j Calpha(alpha,B,j) UoccT(j,J)   (RIMP2.F90:220-228)
  gemm(false, nba * nv, no, no, Cal, nba * nv, 0, Uo6, no, c2, nba * nv, 0, 1);
  gemm(false, nba, vE, nv, c2, nba, nba * nv, Uv6, nv, c3, nba, nba * vE, no);
  // B4(alpha,Al,b) = sum_a UvirtEOST(a,Al) sum_{k < nocc} CalphaVV(alpha,a,k) UvirtT[k + nocc b]   (RIMP2.F90:231-244),
  // summed as sum_{k < nocc} [sum_a UvirtEOST(a,Al) CalphaVV(alpha,a,k)] UvirtT[k + nocc b]: the virtual
  // contraction on the nocc columns the product uses instead of on all nvirt of them
  gemm(false, nba, vE, nv, Cvv, nba, nba * nv, Uv6, nv, v2, nba, nba * vE, no);
  gemm(false, nba * vE, nv, no, v2, nba * vE, 0, Uv5, no, b4, nba * vE, 0, 1);
  gemm(true, vE * no, vE * nv, nba, c3, nba, 0, b4, nba, blad, vE * no, 0, 1);
}

void dgemm_device(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                  int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  Gemm g;
  g.ta = ta;
  g.tb = tb;
  g.M = m;
  g.N = n;
  g.K = k;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  run_gemm(s, g);
}

}  // namespace rimp2

using namespace rimp2;

namespace {

int parse_op(int op, const char* which) {
  switch (op) {
    case 0: case 'N': case 'n': return 0;
    case 1: case 2: case 'T': case 't': case 'C': case 'c': return 1;
    default: fail(RIMP2_ERR_ARG, "%s = %d is not N (0) or T (1, 2)", which, op);
  }
}

void eos_validate(int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba) {
  if (nvirt < 1 || nocc < 1 || noccEOS < 1 || nvirtEOS < 1 || nba < 1)
    fail(RIMP2_ERR_ARG, "dimensions must be >= 1 (nvirt=%lld nocc=%lld noccEOS=%lld nvirtEOS=%lld nba=%lld)",
         (long long)nvirt, (long long)nocc, (long long)noccEOS, (long long)nvirtEOS, (long long)nba);
  if (noccEOS > nocc || nvirtEOS > nvirt)
    fail(RIMP2_ERR_ARG, "noccEOS=%lld > nocc=%lld or nvirtEOS=%lld > nvirt=%lld", (long long)noccEOS,
         (long long)nocc, (long long)nvirtEOS, (long long)nvirt);
}

void first_order_validate(int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba,
                          int64_t nocctot) {
  eos_validate(nvirt, nocc, noccEOS, nvirtEOS, nba);
  if (nocctot != nocc)
    fail(RIMP2_ERR_ARG, "nocctot=%lld != nocc=%lld: the reference multiplies CalphaOcc by UoccT(nocc,nocc) as an "
         "nocctot x nocctot matrix (RIMP2.F90:195-198), in bounds only for nocctot = nocc",
         (long long)nocctot, (long long)nocc);
  if (nocc > nvirt)
    fail(RIMP2_ERR_ARG, "nocc=%lld > nvirt=%lld: the reference's CalphaVV product (K = ldb = nocc, "
         "RIMP2.F90:231-235) reads past UvirtT", (long long)nocc, (long long)nvirt);
}

}  // namespace

extern "C" {

int rimp2_dgemm(void* ctx, int transa, int transb, int64_t m, int64_t n, int64_t k, const double* alpha,
                const double* d_A, int64_t lda, const double* d_B, int64_t ldb, const double* beta, double* d_C,
                int64_t ldc) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  const int ta = parse_op(transa, "transa"), tb = parse_op(transb, "transb");
  if (m < 0 || n < 0 || k < 0) fail(RIMP2_ERR_ARG, "negative dimension (m=%lld n=%lld k=%lld)", (long long)m,
                                    (long long)n, (long long)k);
  if (!alpha || !beta) fail(RIMP2_ERR_ARG, "null alpha or beta");
  const int64_t ra = ta ? k : m, rb = tb ? n : k;
  if (lda < std::max<int64_t>(1, ra) || ldb < std::max<int64_t>(1, rb) || ldc < std::max<int64_t>(1, m))
    fail(RIMP2_ERR_ARG, "leading dimension too small (lda=%lld ldb=%lld ldc=%lld)", (long long)lda, (long long)ldb,
         (long long)ldc);
  if (m == 0 || n == 0) return RIMP2_OK;
  if (!d_C || (k > 0 && (!d_A || !d_B))) fail(RIMP2_ERR_ARG, "null device pointer");
  // alpha = 0: A and B are not read (BLAS), so a NaN there cannot reach C
  dgemm_device(c->stream, ta, tb, m, n, *alpha == 0.0 ? 0 : k, *alpha, d_A, lda, d_B, ldb, *beta, d_C, ldc);
  RIMP2_API_END(ctx)
}

int rimp2_eos_arrays_device(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba,
                            const double* d_Calpha, const double* d_EVocc, const double* d_EVvirt,
                            const double* d_UoccEOST, const double* d_UvirtT, const double* d_UoccT,
                            const double* d_UvirtEOST, double* d_toccEOS, double* d_tvirtEOS, double* d_goccEOS,
                            double* d_gvirtEOS) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  eos_validate(nvirt, nocc, noccEOS, nvirtEOS, nba);
  if (!d_Calpha || !d_EVocc || !d_EVvirt || !d_UoccEOST || !d_UvirtT || !d_UoccT || !d_UvirtEOST || !d_toccEOS ||
      !d_tvirtEOS || !d_goccEOS || !d_gvirtEOS)
    fail(RIMP2_ERR_ARG, "null device pointer");
  reset_status_async(c);
  {
    const int64_t nv = nvirt, no = nocc;
    const int64_t nin[7] = {nba * nv * no, no, nv, no * noccEOS, nv * nv, no * no, nv * nvirtEOS};
    const double* din[7] = {d_Calpha, d_EVocc, d_EVvirt, d_UoccEOST, d_UvirtT, d_UoccT, d_UvirtEOST};
    for (int i = 0; i < 7; i++) check_inputs_device(c, din[i], nin[i], "eos input");
  }
  eos_arrays_device(c, nvirt, nocc, noccEOS, nvirtEOS, nba, d_Calpha, d_EVocc, d_EVvirt, d_UoccEOST, d_UvirtT,
                    d_UoccT, d_UvirtEOST, d_toccEOS, d_tvirtEOS, d_goccEOS, d_gvirtEOS, c->eos_chunk);
  c->status_unread = true;
  RIMP2_API_END(ctx)
}

int rimp2_eos_arrays(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba,
                     const double* Calpha, const double* EVocc, const double* EVvirt, const double* UoccEOST,
                     const double* UvirtT, const double* UoccT, const double* UvirtEOST, double* toccEOS,
                     double* tvirtEOS, double* goccEOS, double* gvirtEOS) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  eos_validate(nvirt, nocc, noccEOS, nvirtEOS, nba);
  if (!Calpha || !EVocc || !EVvirt || !UoccEOST || !UvirtT || !UoccT || !UvirtEOST || !toccEOS || !tvirtEOS ||
      !goccEOS || !gvirtEOS)
    fail(RIMP2_ERR_ARG, "null pointer");
  if (!c->eos) c->eos = new EosState();
  const int64_t nv = nvirt, no = nocc, nE = noccEOS, vE = nvirtEOS;
  // inputs, in order, and outputs
  const int64_t nin[7] = {nba * nv * no, no, nv, no * nE, nv * nv, no * no, nv * vE};
  const double* hin[7] = {Calpha, EVocc, EVvirt, UoccEOST, UvirtT, UoccT, UvirtEOST};
  const int64_t nout[4] = {nv * nE * nv * nE, vE * no * vE * no, nv * nE * nv * nE, vE * no * vE * no};
  double* hout[4] = {toccEOS, tvirtEOS, goccEOS, gvirtEOS};
  size_t offi[7], offo[4], bi = 0, bo = 0;
  for (int i = 0; i < 7; i++) {
    offi[i] = bi;
    bi += align256((size_t)nin[i] * sizeof(double));
  }
  for (int i = 0; i < 4; i++) {
    offo[i] = bo;
    bo += align256((size_t)nout[i] * sizeof(double));
  }
  unsigned char* din = static_cast<unsigned char*>(c->eos->in.ensure(bi));
  unsigned char* dout = static_cast<unsigned char*>(c->eos->out.ensure(bo));
  reset_status(c);
  for (int i = 0; i < 7; i++)
    h2d_copy(c, c->stream, din + offi[i], hin[i], (size_t)nin[i] * sizeof(double));
  const double* d[7];
  for (int i = 0; i < 7; i++) d[i] = reinterpret_cast<const double*>(din + offi[i]);
  for (int i = 0; i < 7; i++) check_inputs_device(c, d[i], nin[i], "eos input");
  check_status(c);
  double* o[4];
  for (int i = 0; i < 4; i++) o[i] = reinterpret_cast<double*>(dout + offo[i]);
  eos_arrays_device(c, nv, no, nE, vE, nba, d[0], d[1], d[2], d[3], d[4], d[5], d[6], o[0], o[1], o[2], o[3], c->eos_chunk);
  for (int i = 0; i < 4; i++)
    RIMP2_HIP_CHECK(hipMemcpyAsync(hout[i], o[i], (size_t)nout[i] * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_first_order_device(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS,
                              int64_t nba, int64_t nocctot, const double* d_Calpha, const double* d_CalphaOcc,
                              const double* d_CalphaVV, const double* d_UoccEOST, const double* d_UvirtT,
                              const double* d_UoccT, const double* d_UvirtEOST, double* d_djik, double* d_blad) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  first_order_validate(nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot);
  if (!d_Calpha || !d_CalphaOcc || !d_CalphaVV || !d_UoccEOST || !d_UvirtT || !d_UoccT || !d_UvirtEOST || !d_djik ||
      !d_blad)
    fail(RIMP2_ERR_ARG, "null device pointer");
  reset_status_async(c);
  {
    const int64_t nv = nvirt, no = nocc;
    const int64_t nin[7] = {nba * nv * no, nba * no * nocctot, nba * nv * nv, no * noccEOS, nv * nv, no * no,
                            nv * nvirtEOS};
    const double* din[7] = {d_Calpha, d_CalphaOcc, d_CalphaVV, d_UoccEOST, d_UvirtT, d_UoccT, d_UvirtEOST};
    for (int i = 0; i < 7; i++) check_inputs_device(c, din[i], nin[i], "first-order input");
  }
  first_order_device(c, nvirt, nocc, noccEOS, nvirtEOS, nba, d_Calpha, d_CalphaOcc, d_CalphaVV, d_UoccEOST, d_UvirtT,
                     d_UoccT, d_UvirtEOST, d_djik, d_blad);
  c->status_unread = true;
  RIMP2_API_END(ctx)
}

int rimp2_first_order(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba,
                      int64_t nocctot, const double* Calpha, const double* CalphaOcc, const double* CalphaVV,
                      const double* UoccEOST, const double* UvirtT, const double* UoccT, const double* UvirtEOST,
                      double* djik, double* blad) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  first_order_validate(nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot);
  if (!Calpha || !CalphaOcc || !CalphaVV || !UoccEOST || !UvirtT || !UoccT || !UvirtEOST || !djik || !blad)
    fail(RIMP2_ERR_ARG, "null pointer");
  if (!c->eos) c->eos = new EosState();
  const int64_t nv = nvirt, no = nocc, nE = noccEOS, vE = nvirtEOS;
  const int64_t nin[7] = {nba * nv * no, nba * no * nocctot, nba * nv * nv, no * nE, nv * nv, no * no, nv * vE};
  const double* hin[7] = {Calpha, CalphaOcc, CalphaVV, UoccEOST, UvirtT, UoccT, UvirtEOST};
  const int64_t nout[2] = {nv * nE * nE * nocctot, vE * no * vE * nv};
  double* hout[2] = {djik, blad};
  size_t offi[7], offo[2], bi = 0, bo = 0;
  for (int i = 0; i < 7; i++) {
    offi[i] = bi;
    bi += align256((size_t)nin[i] * sizeof(double));
  }
  for (int i = 0; i < 2; i++) {
    offo[i] = bo;
    bo += align256((size_t)nout[i] * sizeof(double));
  }
  unsigned char* din = static_cast<unsigned char*>(c->eos->in.ensure(bi));
  unsigned char* dout = static_cast<unsigned char*>(c->eos->out.ensure(bo));
  reset_status(c);
  for (int i = 0; i < 7; i++)
    h2d_copy(c, c->stream, din + offi[i], hin[i], (size_t)nin[i] * sizeof(double));
  const double* d[7];
  for (int i = 0; i < 7; i++) d[i] = reinterpret_cast<const double*>(din + offi[i]);
  for (int i = 0; i < 7; i++) check_inputs_device(c, d[i], nin[i], "first-order input");
  check_status(c);
  double* o[2];
  for (int i = 0; i < 2; i++) o[i] = reinterpret_cast<double*>(dout + offo[i]);
  first_order_device(c, nv, no, nE, vE, nba, d[0], d[1], d[2], d[3], d[4], d[5], d[6], o[0], o[1]);
  for (int i = 0; i < 2; i++)
    RIMP2_HIP_CHECK(hipMemcpyAsync(hout[i], o[i], (size_t)nout[i] * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_eos_set_mode(void* ctx, int mode) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (mode != RIMP2_MODE_FP64 && mode != RIMP2_MODE_I8S5)
    fail(RIMP2_ERR_ARG, "4-index amplitude mode %d: only RIMP2_MODE_FP64 and RIMP2_MODE_I8S5", mode);
  c->eos_mode = mode;
  RIMP2_API_END(ctx)
}

int rimp2_eos_set_chunk(void* ctx, int64_t nocc_per_chunk) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (nocc_per_chunk < 0) fail(RIMP2_ERR_ARG, "nocc_per_chunk = %lld < 0", (long long)nocc_per_chunk);
  c->eos_chunk = nocc_per_chunk;
  RIMP2_API_END(ctx)
}

}  // extern "C"
