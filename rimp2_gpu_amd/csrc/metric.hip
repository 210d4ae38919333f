// FP64 metric step on the GPU:  V = L L^T (blocked Cholesky),  L^{-1}
// (blocked triangular inverse) and the metric transform B = L^{-1} X, all on
// one FP64-MFMA GEMM kernel (v_mfma_f64_16x16x4f64) plus two small LDS
// kernels for the 64x64 diagonal blocks.
//
// The reference receives the fitted tensor Calpha(NBA,nvirt,nocc) from its
// parent code (RIMP2.F90:48-49, "alphaCD ... in the diagonal basis"); this file
// builds that tensor from the raw (P|ia) and the metric V_PQ.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "pair_core.h"

namespace rimp2 {
namespace {

using pairk::lds_ptr_t;
using pairk::pipeline_barrier;

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NB = (int)mg::kPanel;  // Cholesky / inverse panel width

// ---------------------------------------------------------------------------
// Generic column-major FP64 GEMM on MFMA:
//   C = alpha * A * op(B) + beta * C,   A: M x K (lda),  op(B): K x N.
// TRANSB: op(B)[k,n] = B[n + ldb*k], else B[k + ldb*n].
// TRI_A : A is lower triangular (A[m,k] = 0 for k > m): the K loop of a row
//         tile stops at its last row.
// LOWER_C: only tiles touching the lower triangle (n <= m) are computed;
//         tiles strictly above the diagonal return at once.
// 256 threads = 4 waves in 2x2, wave tile (BM/2) x (BN/2) of 16x16 MFMA blocks.
// MFMA f64 16x16x4 operand layout (lane l): A[l%16][l/16], B[l/16][l%16];
// result d[r] = D[4r + l/16][l%16].
// Pipelining: the next K tile is fetched into registers (16-byte loads where
// the leading dimensions allow) while the MFMAs consume the current one.
// LDS: As[k][m] with pitch BM+16 doubles and Bs[n][k] with pitch 17 doubles,
// so both fragment reads are free of bank conflicts.
// ---------------------------------------------------------------------------
// One BM x BN tile (row tile mt, column tile nt) of C; all 256 threads.
// LDS of one tile: As[BK * (BM + 16)] then Bs[BN * (BK + 1)] doubles.
template <int BM, int BN>
constexpr int gemm_tile_lds_doubles() {
  return 16 * (BM + 16) + BN * 17;
}

template <int BM, int BN, bool TRANSB, bool TRI_A, bool LOWER_C>
__device__ __attribute__((always_inline)) inline void gemm_f64_tile(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                                     const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int64_t mt,
                                     int64_t nt, double* __restrict__ lds) {
  constexpr int BK = 16;
  constexpr int PA = BM + 16;  // As[k][m]
  constexpr int PB = BK + 1;   // Bs[n][k]
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int NA = BM * BK / 512;  // double2 loads of A per thread per K tile
  constexpr int NBL = BN * BK / 512; // double2 loads of B per thread per K tile
  static_assert(gemm_tile_lds_doubles<BM, BN>() == BK * PA + BN * PB, "tile LDS layout");
  double* __restrict__ As = lds;
  double* __restrict__ Bs = lds + BK * PA;

  const int64_t m0 = mt * BM;
  const int64_t n0 = nt * BN;
  if (LOWER_C && n0 > m0 + BM - 1) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const bool a_vec = ((lda & 1) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
  const bool b_vec = ((ldb & 1) == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);

  v4d acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  int64_t kend = K;
  if (TRI_A) kend = (m0 + BM < K) ? (m0 + BM) : K;

  double2 ra[NA], rb[NBL];
  // A tile BM x BK: element pair (m, m+1) of column k; BM/2 pairs per column
  auto load_a = [&](int64_t k0) {
#pragma unroll
    for (int t = 0; t < NA; t++) {
      const int e = tid + 256 * t;
      const int k = e / (BM / 2), m = (e % (BM / 2)) * 2;
      const int64_t gm = m0 + m, gk = k0 + k;
      double2 v = {0.0, 0.0};
      if (gk < K) {
        if (a_vec && gm + 1 < M) v = *reinterpret_cast<const double2*>(A + gm + lda * gk);
        else {
          if (gm < M) v.x = A[gm + lda * gk];
          if (gm + 1 < M) v.y = A[gm + 1 + lda * gk];
        }
      }
      ra[t] = v;
    }
  };
  // op(B) tile BK x BN
  auto load_b = [&](int64_t k0) {
#pragma unroll
    for (int t = 0; t < NBL; t++) {
      const int e = tid + 256 * t;
      double2 v = {0.0, 0.0};
      if (TRANSB) {  // pairs along n: (n, n+1) of row k
        const int k = e / (BN / 2), n = (e % (BN / 2)) * 2;
        const int64_t gn = n0 + n, gk = k0 + k;
        if (gk < K) {
          if (b_vec && gn + 1 < N) v = *reinterpret_cast<const double2*>(B + gn + ldb * gk);
          else {
            if (gn < N) v.x = B[gn + ldb * gk];
            if (gn + 1 < N) v.y = B[gn + 1 + ldb * gk];
          }
        }
      } else {  // pairs along k: (k, k+1) of column n
        const int n = e / (BK / 2), k = (e % (BK / 2)) * 2;
        const int64_t gn = n0 + n, gk = k0 + k;
        if (gn < N) {
          if (b_vec && gk + 1 < K) v = *reinterpret_cast<const double2*>(B + gk + ldb * gn);
          else {
            if (gk < K) v.x = B[gk + ldb * gn];
            if (gk + 1 < K) v.y = B[gk + 1 + ldb * gn];
          }
        }
      }
      rb[t] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int t = 0; t < NA; t++) {
      const int e = tid + 256 * t;
      const int k = e / (BM / 2), m = (e % (BM / 2)) * 2;
      As[k * PA + m] = ra[t].x;
      As[k * PA + m + 1] = ra[t].y;
    }
#pragma unroll
    for (int t = 0; t < NBL; t++) {
      const int e = tid + 256 * t;
      if (TRANSB) {
        const int k = e / (BN / 2), n = (e % (BN / 2)) * 2;
        Bs[n * PB + k] = rb[t].x;
        Bs[(n + 1) * PB + k] = rb[t].y;
      } else {
        const int n = e / (BK / 2), k = (e % (BK / 2)) * 2;
        Bs[n * PB + k] = rb[t].x;
        Bs[n * PB + k + 1] = rb[t].y;
      }
    }
  };

  if (kend > 0) {
    load_a(0);
    load_b(0);
  }
  for (int64_t k0 = 0; k0 < kend; k0 += BK) {
    store();
    __syncthreads();
    if (k0 + BK < kend) {
      load_a(k0 + BK);
      load_b(k0 + BK);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) af[i] = As[(ks + lane / 16) * PA + wm + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; j++) bf[j] = Bs[(wn + 16 * j + (lane & 15)) * PB + ks + lane / 16];
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int64_t gm = m0 + wm + 16 * i + 4 * r + lane / 16;
        int64_t gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < M && gn < N) {
          if (LOWER_C && gn > gm) continue;
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * C[gm + ldc * gn];
          C[gm + ldc * gn] = v;
        }
      }
}

template <int BM, int BN, bool TRANSB, bool TRI_A, bool LOWER_C>
__device__ __attribute__((always_inline)) inline void gemm_f64_tile(int64_t M, int64_t N, int64_t K, double alpha,
                                                                    const double* A, int64_t lda, const double* B,
                                                                    int64_t ldb, double beta, double* C, int64_t ldc,
                                                                    int64_t mt, int64_t nt) {
  __shared__ double lds[gemm_tile_lds_doubles<BM, BN>()];
  gemm_f64_tile<BM, BN, TRANSB, TRI_A, LOWER_C>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mt, nt, lds);
}

template <int BM, int BN, bool TRANSB, bool TRI_A, bool LOWER_C>
__global__ __launch_bounds__(256) void gemm_f64_kernel(int64_t M, int64_t N, int64_t K, double alpha,
                                                       const double* A, int64_t lda, const double* B, int64_t ldb,
                                                       double beta, double* C, int64_t ldc, int64_t sA,
                                                       int64_t sB, int64_t sC) {
  // batch member blockIdx.z (element strides sA, sB, sC between members)
  const int64_t z = blockIdx.z;
  gemm_f64_tile<BM, BN, TRANSB, TRI_A, LOWER_C>(M, N, K, alpha, A + z * sA, lda, B + z * sB, ldb, beta, C + z * sC,
                                                ldc, blockIdx.x, blockIdx.y);
}

// Job lookup of the grouped kernels: the last job j with jobs[j].first <= t
// (jobs in ascending order of first, jobs[0].first = 0).
template <class Job>
__device__ inline int job_find(const Job* __restrict__ jobs, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Grouped GEMM: independent problems of any shapes in one launch (job
// tables of metric_group.h); LOWER_C jobs list only the lower tiles,
// t = mt (mt + 1) / 2 + nt.
using mg::GemmJob;

template <int BM, int BN, bool TRANSB, bool TRI_A, bool LOWER_C>
__global__ __launch_bounds__(256) void gemm_f64_group_kernel(const GemmJob* __restrict__ jobs, int njobs) {
  const int64_t t = blockIdx.x;
  const GemmJob& J = jobs[job_find(jobs, njobs, t)];
  const int64_t u = t - J.first;
  int64_t mt, nt;
  if (LOWER_C) pairk::tri_decode(u, nt, mt);
  else {
    mt = u % J.tm;
    nt = u / J.tm;
  }
  gemm_f64_tile<BM, BN, TRANSB, TRI_A, LOWER_C>(J.M, J.N, J.K, J.alpha, J.A, J.lda, J.B, J.ldb, J.beta, J.C, J.ldc,
                                                mt, nt);
}

template <int BM, int BN, bool TRANSB, bool TRI_A, bool LOWER_C>
void gemm_f64(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int64_t batch = 1, int64_t sA = 0,
              int64_t sB = 0, int64_t sC = 0) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  dim3 grid((unsigned)ceil_div(M, BM), (unsigned)ceil_div(N, BN), (unsigned)batch);
  hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, TRANSB, TRI_A, LOWER_C>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda,
                     B, ldb, beta, C, ldc, sA, sB, sC);
  RIMP2_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Orders this wave's LDS stores before its following LDS loads as seen by
// the other lanes (no instruction at wavefront scope; constrains the compiler).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Diagonal block: factor A[k0:k0+kb, k0:k0+kb] = L11 L11^T, write L11 back
// (zeros above the diagonal) and L11^{-1} into Linv at the same block.
// status bit 1 is set on a non-positive pivot; status[1] receives the global
// index of the first failing pivot (1-based).
//
// The block is padded to 64 x 64 with the identity, so there is no ragged
// case.  Factorisation: one wave, lane r holds row r in registers, the column
// loop is unrolled (static register indices) and runs the unscaled update
//     a_r[q] -= (a_r[c] / d_c) * a_q[c],   q > c,   d_c = a_c[c];
// column c is published through LDS (lane r writes a_r[c], every lane then
// reads the entries it needs with broadcast loads, two per instruction).  A
// wave's LDS accesses complete in order, so a wavefront-scope fence (no
// instruction) is all the ordering needed: the whole dependency chain lives
// in one wave.  The scaling by 1/sqrt(d_c) is applied once at the end.
// Inverse: all four waves, in 16 x 16 blocks (diagonal blocks by forward
// substitution, the blocks below by distance on f64 MFMA; see below); X goes
// through LDS to be stored with coalesced writes by all four waves.
// ---------------------------------------------------------------------------
struct PotrfLds {
  double Ls[NB][NB + 1];
  double Xs[NB][NB + 1];
  __align__(16) double colb[2][NB];
  __align__(16) double invd[NB];
  double Ts[3][16][17];  // per-wave 16 x 16 products of the inverse
};

__device__ __attribute__((always_inline)) inline void potrf_diag_block(double* A, int64_t lda, double* Linv, int64_t ldi, int64_t k0, int kb,
                                        int* status, PotrfLds& sh) {
  auto& Ls = sh.Ls;
  auto& Xs = sh.Xs;
  auto& colb = sh.colb;
  auto& invd = sh.invd;
  auto& Ts = sh.Ts;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (wave == 0) {
    const int r = lane;
    double a[NB];
#pragma unroll
    for (int q = 0; q < NB; q++)
      a[q] = (r < kb && q < kb) ? (q <= r ? A[(k0 + r) + lda * (k0 + q)] : 0.0) : (r == q ? 1.0 : 0.0);
    double piv = 1.0;
    int bad_at = 0;
#pragma unroll
    for (int c = 0; c < NB; c++) {
      double* col = colb[c & 1];
      col[r] = a[c];
      wave_lds_sync();
      double d = col[c];
      if (!(d > 0.0)) {
        if (bad_at == 0) bad_at = c + 1;
        d = 1.0;
      }
      if (r == c) piv = d;
      const double lrc = a[c] / d;
#pragma unroll
      for (int q = c + 1; q < NB; q++) a[q] = fma(-lrc, col[q], a[q]);
    }
    const double sp = sqrt(piv);
    invd[r] = 1.0 / sp;
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < NB; q++) Ls[r][q] = q < r ? a[q] * invd[q] : (q == r ? sp : 0.0);
    wave_lds_sync();
    if (bad_at && r == 0) {
      atomicCAS(status + 1, 0, (int)(k0 + bad_at));
      atomicOr(status, 2);
    }
  }
  __syncthreads();
  // L11^{-1} on all four waves, in 16 x 16 blocks: wave w inverts the diagonal
  // block D_w by forward substitution (lane j < 16: column j), then the
  // blocks below the diagonal by distance d = i - j:
  //     X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj
  // (from (L X)_ij = 0), the 16 x 16 products on f64 MFMA 16x16x4.
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e & 63, c = e >> 6;
    if ((c >> 4) > (r >> 4)) Xs[r][c] = 0.0;
  }
  if (lane < 16) {
    const int j = lane, o = 16 * wave;
    double x[16];
#pragma unroll
    for (int c = 0; c < 16; c++) {
      double s = (c == j) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < c; k++) s = fma(-Ls[o + c][o + k], x[k], s);
      x[c] = s * invd[o + c];
    }
#pragma unroll
    for (int c = 0; c < 16; c++) Xs[o + c][o + j] = x[c];
  }
  __syncthreads();
  for (int d = 1; d < 4; d++) {
    if (wave < 4 - d) {
      const int j = wave, i = wave + d;
      // MFMA operands: A[l % 16][l / 16], B[l / 16][l % 16]; result D[4 r + l / 16][l % 16]
      v4d t = {0.0, 0.0, 0.0, 0.0};
      for (int k = j; k < i; k++)
#pragma unroll
        for (int s4 = 0; s4 < 4; s4++)
          t = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[16 * i + (lane & 15)][16 * k + 4 * s4 + (lane >> 4)],
                                                   Xs[16 * k + 4 * s4 + (lane >> 4)][16 * j + (lane & 15)], t, 0, 0,
                                                   0);
#pragma unroll
      for (int r = 0; r < 4; r++) Ts[wave][4 * r + (lane >> 4)][lane & 15] = t[r];
      wave_lds_sync();
      v4d x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s4 = 0; s4 < 4; s4++)
        x = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[16 * i + (lane & 15)][16 * i + 4 * s4 + (lane >> 4)],
                                                 Ts[wave][4 * s4 + (lane >> 4)][lane & 15], x, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; r++) Xs[16 * i + 4 * r + (lane >> 4)][16 * j + (lane & 15)] = -x[r];
    }
    __syncthreads();
  }
  {
    const int r = lane;
#pragma unroll
    for (int t = 0; t < NB / 4; t++) {
      const int q = wave + 4 * t;
      if (r < kb && q < kb) {
        Linv[(k0 + r) + ldi * (k0 + q)] = q <= r ? Xs[r][q] : 0.0;
        A[(k0 + r) + lda * (k0 + q)] = Ls[r][q];
      }
    }
  }
}

__global__ __launch_bounds__(256) void potrf_diag_kernel(double* A, int64_t lda, double* Linv, int64_t ldi,
                                                         int64_t k0, int kb, int* status) {
  __shared__ PotrfLds sh;
  potrf_diag_block(A, lda, Linv, ldi, k0, kb, status, sh);
}

// Trailing update of a Cholesky step, A22 -= L21 L21^T (lower 64 x 64 tiles,
// one per workgroup), with the factor of the next diagonal block in the same
// launch: that block is tile (0, 0), so its workgroup factors it as soon as it
// has updated it, while the other workgroups are still updating theirs.  (The
// diagonal factor is a ~30 us latency-bound chain; here it runs under the
// update instead of after it.)
// The two phases share one LDS buffer, so the kernel holds the factor's 74 KB
// (two workgroups per CU), not that plus the tile's 19 KB.
__global__ __launch_bounds__(256) void syrk_potrf_kernel(int64_t mr, int64_t kb, const double* L21, double* L,
                                                         int64_t n, double* Linv, int64_t r0, int kbn, int* status) {
  static_assert(sizeof(PotrfLds) >= gemm_tile_lds_doubles<64, 64>() * sizeof(double), "shared LDS");
  __shared__ PotrfLds sh;
  gemm_f64_tile<64, 64, true, false, true>(mr, mr, kb, -1.0, L21, n, L21, n, 1.0, L + r0 + n * r0, n, blockIdx.x,
                                           blockIdx.y, &sh.Ls[0][0]);
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    __syncthreads();  // the block's update is written, and the tile's LDS is free
    potrf_diag_block(L, n, Linv, n, r0, kbn, status, sh);
  }
}

// Grouped: workgroup g factors the diagonal block of job g.
using mg::PotrfJob;

__global__ __launch_bounds__(256) void potrf_diag_group_kernel(const PotrfJob* __restrict__ jobs) {
  __shared__ PotrfLds sh;
  const PotrfJob& J = jobs[blockIdx.x];
  potrf_diag_block(J.A, J.n, J.Linv, J.n, J.k0, J.kb, J.status, sh);
}

// ---------------------------------------------------------------------------
// Metric transform on packed operands.
//   Lp[ks][mb][lane]: A fragments of L^{-1}, lane l <-> (m = 16 mb + l%16,
//                     k = 4 ks + l/16), zero above the diagonal and in padding;
//   Xp[i][ks][rb][lane]: B fragments of (P|ia), lane l <-> (q = 4 ks + l/16,
//                     a = 16 rb + l%16)  (the FP64 operand layout of pack.hip);
// so every 512-byte fragment is contiguous in HBM and in LDS.  One workgroup
// computes a 128 (Q rows) x 128 (virtuals of one occupied) tile of B = L^{-1} X
// with a four-stage LDS ring fed by buffer_load_dwordx4 ... lds (32 KB per
// 16-wide K stage); the K loop of a row tile stops at its last row (L^{-1} is
// lower triangular).  Waves 4 x 2, wave tile 32 x 64 = 2 x 4 MFMA blocks.
// ---------------------------------------------------------------------------
constexpr int TR_TM = (int)mg::kTrTM, TR_TN = (int)mg::kTrTN, TR_KQ = (int)mg::kTrKQ;  // K padding: four 16-wide stages

__global__ void pack_linv_kernel(const double* __restrict__ Li, int64_t n, int64_t MB, int64_t KS,
                                 double* __restrict__ Lp) {
  RIMP2_GRID_STRIDE(t, KS * MB * 64) {
    const int lane = (int)(t & 63);
    const int64_t f = t >> 6, mb = f % MB, ks = f / MB;
    const int64_t m = mb * 16 + (lane & 15), k = ks * 4 + (lane >> 4);
    Lp[t] = (m < n && k < n && k <= m) ? Li[m + n * k] : 0.0;
  }
}

using mg::TransformArgs;

// One 128 (Q rows) x 128 (virtuals of occupied i) tile of B = L^{-1} X.
// Eight waves on a 4 x 2 grid, wave tile 32 x 64 (2 x 4 MFMA blocks of 16x16),
// two waves per SIMD.  Same operand pipeline as the pair kernel: four LDS
// stage buffers (32 KB per 16-deep K stage: 16 KB of L^{-1}, 16 KB of X) fed
// by buffer_load_dwordx4 ... lds, next-stage operands read under the current
// stage's MFMAs, one DMA + six LDS reads per eight MFMAs.
// A wave stops its LDS reads and MFMAs after the last stage where its 32-row
// strip of L^{-1} has nonzeros (row q ends at column q), and does none in a
// strip that lies wholly in the padding beyond naux; it still issues its share
// of the DMA and meets every barrier.  The products skipped are exact zeros,
// so B is unchanged.  Strips pair up on the SIMDs as (0, 3) and (1, 2)
// (waves w and w + 4 share a SIMD), so in the diagonal block each SIMD keeps
// 10 of the 16 strip-stages of a full one.
__device__ __attribute__((always_inline)) inline void transform_unit(const TransformArgs& p, int64_t wu, unsigned char* __restrict__ buf0,
                                      unsigned char* __restrict__ buf1, unsigned char* __restrict__ buf2,
                                      unsigned char* __restrict__ buf3) {
  constexpr int FB = 512;                     // bytes per fragment
  constexpr int NW = 8;                       // waves
  constexpr int A_STAGE = 4 * 8 * FB;         // 4 k-steps x 8 row blocks of L^{-1}
  constexpr int B_STAGE = 4 * 8 * FB;         // 4 k-steps x 8 column blocks of X
  constexpr int SB = A_STAGE + B_STAGE;       // 32 KB
  constexpr int NI = SB / (NW * 1024);        // 4 DMA transfers per wave per stage (0-1: L^{-1}, 2-3: X)
  static_assert(A_STAGE == 2 * NW * 1024, "transfers 0 and 1 of every wave carry the L^{-1} slab");
  const int64_t mt = wu % p.MT, col = wu / p.MT, i = col / p.NT, nt = col % p.NT;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave < 4 ? wave >> 1 : 5 - (wave >> 1), wc = wave & 1;  // strips 0 0 1 1 3 3 2 2
  const int64_t m0 = mt * TR_TM;
  // L^{-1} is lower triangular: the K loop of row tile mt stops at its last
  // row (m0 + 128 is a whole number of four-stage groups; so is KS * 4)
  int64_t kend = m0 + TR_TM;
  if (kend > p.KS * 4) kend = p.KS * 4;
  const int64_t nst = kend / 16;
  // stages [0, wst) carry this wave's nonzeros; wst is even (strips are two stages deep)
  int64_t wst = (m0 + 32 * (wr + 1)) / 16;
  if (wst > nst) wst = nst;
  if (m0 + 32 * wr >= p.naux) wst = 0;
  const size_t a_ks = (size_t)p.MB * FB, b_ks = (size_t)p.NRB * FB;
  const size_t a_off = (size_t)(mt * 8) * FB, b_off = (size_t)nt * 8 * FB;
  const __amdgpu_buffer_rsrc_t ra = pairk::slab_rsrc(p.Lp + a_off, (size_t)p.KS * a_ks - a_off);
  const __amdgpu_buffer_rsrc_t rb =
      pairk::slab_rsrc(p.Xp + (size_t)i * p.KS * b_ks + b_off, (size_t)p.KS * b_ks - b_off);
  uint32_t voff[NI];
#pragma unroll
  for (int n = 0; n < NI; n++) {
    const int wb = (n * NW + wave) * 1024;
    if (n < 2) voff[n] = (uint32_t)((wb / (8 * FB)) * a_ks + wb % (8 * FB) + lane * 16);
    else voff[n] = (uint32_t)(((wb - A_STAGE) / (8 * FB)) * b_ks + (wb - A_STAGE) % (8 * FB) + lane * 16);
  }

  v4d acc[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = v4d{0.0, 0.0, 0.0, 0.0};

  auto issue = [&](int64_t st, unsigned char* dst) {
    const uint32_t sa = (uint32_t)((size_t)st * 4 * a_ks), sb = (uint32_t)((size_t)st * 4 * b_ks);
#pragma unroll
    for (int n = 0; n < NI; n++)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(n < 2 ? ra : rb, (lds_ptr_t)(dst + (n * NW + wave) * 1024), 16,
                                               voff[n], n < 2 ? sa : sb, 0, 0);
  };
  struct Ops {
    double af[4][2], bf[4][4];
  };
  auto load = [&](const unsigned char* stg, Ops& o) {
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
#pragma unroll
      for (int x = 0; x < 2; x++)
        o.af[ks][x] = *reinterpret_cast<const double*>(stg + (ks * 8 + wr * 2 + x) * FB + lane * 8);
#pragma unroll
      for (int y = 0; y < 4; y++)
        o.bf[ks][y] = *reinterpret_cast<const double*>(stg + A_STAGE + (ks * 8 + wc * 4 + y) * FB + lane * 8);
    }
  };
  auto mma = [&](const Ops& o) {
#pragma unroll
    for (int ks = 0; ks < 4; ks++)
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.af[ks][x], o.bf[ks][y], acc[x][y], 0, 0, 0);
  };
  auto step = [&](int64_t st, Ops& cur, Ops& nxt, unsigned char* freed, const unsigned char* next) {
    RIMP2_WAIT_VMCNT_LGKM0(2 * NI);
    pipeline_barrier();
    issue(st + 4 < nst ? st + 4 : nst - 1, freed);
    load(next, nxt);
    mma(cur);
#pragma unroll
    for (int g = 0; g < NI; g++) {
      __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x0100, 6, 0);
      __builtin_amdgcn_sched_group_barrier(0x0008, 8, 0);
    }
  };
  // a stage past wst: the DMA and the barrier only
  auto idle = [&](int64_t st, unsigned char* freed) {
    RIMP2_WAIT_VMCNT_LGKM0(2 * NI);
    pipeline_barrier();
    issue(st + 4 < nst ? st + 4 : nst - 1, freed);
  };
  Ops oa, ob;
  issue(0, buf0);
  issue(1, buf1);
  issue(2, buf2);
  issue(3, buf3);
  RIMP2_WAIT_VMCNT(3 * NI);
  pipeline_barrier();
  if (wst > 0) load(buf0, oa);
  // every loop bound below is uniform over the wave; all waves run nst stages
  const int64_t wst4 = wst & ~(int64_t)3;
  int64_t st = 0;
  for (; st < wst4; st += 4) {
    step(st, oa, ob, buf0, buf1);
    step(st + 1, ob, oa, buf1, buf2);
    step(st + 2, oa, ob, buf2, buf3);
    step(st + 3, ob, oa, buf3, buf0);
  }
  if (st < wst) {  // wst - st == 2
    step(st, oa, ob, buf0, buf1);
    step(st + 1, ob, oa, buf1, buf2);
    idle(st + 2, buf2);
    idle(st + 3, buf3);
    st += 4;
  }
  for (; st < nst; st += 4) {
    idle(st, buf0);
    idle(st + 1, buf1);
    idle(st + 2, buf2);
    idle(st + 3, buf3);
  }
  RIMP2_WAIT_VMCNT_LGKM0(0);
  // epilogue: D layout d[r] = D[4r + l/16][l%16]
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t q = m0 + wr * 32 + x * 16 + 4 * r + (lane >> 4);
        const int64_t a = nt * TR_TN + wc * 64 + y * 16 + (lane & 15);
        if (q < p.naux && a < p.nvirt) p.B[q + p.naux * (a + p.nvirt * i)] = acc[x][y][r];
      }
  __syncthreads();  // every wave is done with the stage buffers before a next unit refills them
}

// Work unit of the transform kernels: row tiles t and MT-1-t of one column
// tile (col = wu / MTH, t = wu % MTH, MTH = ceil(MT / 2)).  Row tile t runs
// t + 1 K tiles (L^{-1} is lower triangular), so every unit spans MT + 1 K
// tiles and the workgroups of a launch finish together instead of waiting on
// the longest row tiles at the end.
__device__ __attribute__((always_inline)) inline void transform_pair(const TransformArgs& p, int64_t wu, unsigned char* __restrict__ buf0,
                                      unsigned char* __restrict__ buf1, unsigned char* __restrict__ buf2,
                                      unsigned char* __restrict__ buf3) {
  const int64_t mth = (p.MT + 1) >> 1, col = wu / mth, t = wu - col * mth;
  transform_unit(p, col * p.MT + t, buf0, buf1, buf2, buf3);
  if (p.MT - 1 - t != t) transform_unit(p, col * p.MT + (p.MT - 1 - t), buf0, buf1, buf2, buf3);
}

__global__ __launch_bounds__(512) void transform_kernel(TransformArgs p) {
  constexpr int SB = 32768;
  __shared__ __align__(16) unsigned char buf0[SB];
  __shared__ __align__(16) unsigned char buf1[SB];
  __shared__ __align__(16) unsigned char buf2[SB];
  __shared__ __align__(16) unsigned char buf3[SB];
  // XCD-aware: slot s of XCD x takes units x * chunk + s, + nslot, ... (the
  // loop bounds are uniform over the workgroup)
  const int64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
  for (int64_t k = slot; k < p.chunk; k += nslot) {
    const int64_t wu = xcd * p.chunk + k;
    if (wu >= p.nwu) break;
    transform_pair(p, wu, buf0, buf1, buf2, buf3);
  }
}

__global__ void zero_upper_kernel(double* A, int64_t n, int64_t lda) {
  RIMP2_GRID_STRIDE(e, n * n) {
    int64_t r = e % n, c = e / n;
    if (c > r) A[r + lda * c] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Grouped metric pipeline: the Cholesky steps, inverse levels, packing and
// transform of many fragments, one launch per step for the whole group, the
// per-fragment arguments in job tables in device memory.
// ---------------------------------------------------------------------------

// Copy src -> dst (dst == src: check only); bit 4 of status on a non-finite value.
using mg::CopyJob;

// Grid-stride over [0, total) in chunks of 64 x 256 consecutive elements.
// A chunk inside one job (the common case: a metric holds n^2 >= 16384
// elements from n = 128) runs with the job's fields in scalar registers and 64
// independent loads per thread; a chunk across a job boundary steps each
// element to its job.
__global__ __launch_bounds__(256) void copy_check_group_kernel(const CopyJob* __restrict__ jobs, int njobs,
                                                               int64_t total) {
  constexpr int64_t CH = 64 * 256;
  const int tid = threadIdx.x;
  for (int64_t base = (int64_t)blockIdx.x * CH; base < total; base += (int64_t)gridDim.x * CH) {
    const int64_t end = base + CH < total ? base + CH : total;
    int j = job_find(jobs, njobs, base);
    const int64_t jend = j + 1 < njobs ? jobs[j + 1].first : total;
    if (end <= jend) {
      const CopyJob J = jobs[j];
      const double* src = J.src + (base - J.first);
      double* dst = J.dst + (base - J.first);
      const bool copy = J.dst != J.src;
      const int64_t cnt = end - base;
      bool bad = false;
#pragma unroll 16
      for (int r = 0; r < 64; r++) {
        const int64_t e = (int64_t)r * 256 + tid;
        if (e < cnt) {
          const double v = src[e];
          if (copy) dst[e] = v;
          bad |= !isfinite(v);
        }
      }
      if (bad) atomicOr(J.status, 4);
      continue;
    }
    for (int64_t t = base + tid; t < end; t += 256) {
      while (j + 1 < njobs && jobs[j + 1].first <= t) j++;
      const CopyJob& J = jobs[j];
      const int64_t e = t - J.first;
      const double v = J.src[e];
      if (J.dst != J.src) J.dst[e] = v;
      if (!isfinite(v)) atomicOr(J.status, 4);
    }
  }
}

// One wave per fragment: bit 8 of status when max(eo) >= min(ev).
using mg::DenomJob;

__global__ __launch_bounds__(64) void denom_group_kernel(const DenomJob* __restrict__ jobs) {
  const DenomJob& J = jobs[blockIdx.x];
  const int lane = threadIdx.x;
  double mo = -INFINITY, mv = INFINITY;
  for (int64_t i = lane; i < J.nocc; i += 64) mo = fmax(mo, J.eo[i]);
  for (int64_t a = lane; a < J.nvirt; a += 64) mv = fmin(mv, J.ev[a]);
  for (int off = 32; off > 0; off >>= 1) {
    mo = fmax(mo, __shfl_xor(mo, off));
    mv = fmin(mv, __shfl_xor(mv, off));
  }
  if (lane == 0 && !(2.0 * mo - 2.0 * mv < 0.0)) atomicOr(J.status, 8);
}

// L^{-1} of one fragment into A-fragment order (as pack_linv_kernel).
using mg::LinvPackJob;

// One 64-row x 64-column tile of L^{-1} per workgroup (tiles the transform
// never reads are skipped: linv_tile_read): coalesced column reads (512
// bytes per wave) into LDS, zeros above the diagonal and beyond n, then the
// tile's 16 k-steps x 4 row blocks written as 512-byte fragments.
__global__ __launch_bounds__(256) void pack_linv_group_kernel(const LinvPackJob* __restrict__ jobs, int njobs,
                                                              int64_t total) {
  __shared__ double tile[64][65];  // [column][row]
  const int tid = threadIdx.x;
  for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
    const LinvPackJob& J = jobs[job_find(jobs, njobs, t)];
    const int64_t u = t - J.first, kt = J.KS / 16;
    const int64_t mbt = u / kt, kst = u - mbt * kt;
    if (kst >= 2 * (mbt / 2 + 1)) continue;  // not read (mg::linv_tile_read); uniform over the workgroup
    const int64_t m0 = mbt * 64, k0 = kst * 64;
    double v[16];  // all 16 loads in flight before the LDS stores
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int64_t m = m0 + (tid & 63), k = k0 + (tid >> 6) + 4 * r;
      v[r] = (m < J.n && k < J.n && k <= m) ? J.Li[m + J.n * k] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 16; r++) tile[(tid >> 6) + 4 * r][tid & 63] = v[r];
    __syncthreads();
    double* out = J.Lp + ((kst * 16) * J.MB + mbt * 4) * 64;
#pragma unroll 4
    for (int r = 0; r < 16; r++) {
      const int e = 256 * r + tid, fr = e >> 6, lane = e & 63;
      const int ksl = fr >> 2, mbl = fr & 3;
      out[((int64_t)ksl * J.MB + mbl) * 64 + lane] = tile[4 * ksl + (lane >> 4)][16 * mbl + (lane & 15)];
    }
    __syncthreads();  // the tile is reused by the next one
  }
}

// One 64 (q) x 64 (column) tile of X per workgroup: coalesced column reads
// (512 bytes per wave) into LDS, then the tile's 16 k-steps x 4 row blocks
// written as 512-byte fragments of Xp[g][ks][rb][lane].
using mg::XPackJob;

__global__ __launch_bounds__(256) void pack_x_group_kernel(const XPackJob* __restrict__ jobs, int njobs) {
  __shared__ double tile[64][65];  // [column][q]
  const int64_t t = blockIdx.x;
  const XPackJob& J = jobs[job_find(jobs, njobs, t)];
  const int64_t u = t - J.first, kt = J.nks / 16, rt = J.nrb / 4;
  const int64_t rbt = u % rt, kst = (u / rt) % kt, g = u / (rt * kt);
  const int64_t q0 = kst * 64, c0 = rbt * 64;
  const double* Xg = J.X + J.naux * J.cols * g;
  const int tid = threadIdx.x;
  bool bad = false;
  double v[16];  // all 16 loads in flight before the LDS stores
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int cl = (tid >> 6) + 4 * r;
    const int64_t q = q0 + (tid & 63), c = c0 + cl;
    v[r] = (q < J.naux && c < J.cols) ? Xg[q + J.naux * c] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < 16; r++) {
    bad |= !isfinite(v[r]);
    tile[(tid >> 6) + 4 * r][tid & 63] = v[r];
  }
  __syncthreads();
  double* out = J.Xp + ((g * J.nks + kst * 16) * J.nrb + rbt * 4) * 64;
#pragma unroll 4
  for (int r = 0; r < 16; r++) {
    const int e = 256 * r + tid, fr = e >> 6, lane = e & 63;
    const int ksl = fr >> 2, rbl = fr & 3;
    out[((int64_t)ksl * J.nrb + rbl) * 64 + lane] = tile[16 * rbl + (lane & 15)][4 * ksl + (lane >> 4)];
  }
  if (bad) atomicOr(J.status, 4);
}

using mg::TransformJob;

// XCD x takes the units [r.b[x], r.b[x + 1]) (ranges of equal cost, from the
// host); its workgroups stride through them.
__global__ __launch_bounds__(512) void transform_group_kernel(const TransformJob* __restrict__ jobs, int njobs,
                                                              mg::XcdRanges r) {
  constexpr int SB = 32768;
  __shared__ __align__(16) unsigned char buf0[SB];
  __shared__ __align__(16) unsigned char buf1[SB];
  __shared__ __align__(16) unsigned char buf2[SB];
  __shared__ __align__(16) unsigned char buf3[SB];
  const int64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
  for (int64_t wu = r.b[xcd] + slot; wu < r.b[xcd + 1]; wu += nslot) {
    const int j = job_find(jobs, njobs, wu);
    const TransformArgs a = jobs[j].a;
    transform_pair(a, wu - jobs[j].first, buf0, buf1, buf2, buf3);
  }
}

}  // namespace

using mg::linv_packed_bytes;
using mg::x_packed_bytes;

MetricView ctx_metric_view(Ctx* c) {
  return MetricView{c->stream, c->V, c->Linv, c->work, c->linv_packed, c->x_packed,
                    static_cast<int*>(c->status.ensure(64)), c};
}

// dV_in (device, n x n) -> m.L holds L (lower, zeros above), m.Linv holds
// L^{-1}, m.linv_packed L^{-1} in transform-operand order.  A non-positive
// pivot sets bit 2 of m.status (cleared by the caller) and m.status[1] to the
// first failing pivot.  Buffers grow on demand: a caller that queues work for
// several inputs on one view sizes them for the largest first.
void metric_factor(const MetricView& m, int64_t n, const double* dV_in) {
  hipStream_t s = m.stream;
  double* L = static_cast<double*>(m.L.ensure((size_t)(n * n) * sizeof(double)));
  double* Li = static_cast<double*>(m.Linv.ensure((size_t)(n * n) * sizeof(double)));
  int* status = m.status;
  if (dV_in != L) RIMP2_HIP_CHECK(hipMemcpyAsync(L, dV_in, (size_t)(n * n) * sizeof(double), hipMemcpyDeviceToDevice, s));
  RIMP2_HIP_CHECK(hipMemsetAsync(Li, 0, (size_t)(n * n) * sizeof(double), s));
  {
    ScopedTimer t(m.timer, RIMP2_K_CHOLESKY);
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, s, L, n, Li, n, (int64_t)0, (int)std::min<int64_t>(n, NB),
                       status);
    RIMP2_HIP_CHECK(hipGetLastError());
    for (int64_t k0 = 0; k0 < n; k0 += NB) {
      const int kb = (int)((n - k0) < NB ? (n - k0) : NB);
      const int64_t r0 = k0 + kb, mr = n - r0;
      if (mr <= 0) break;
      // panel: L21 = A21 * L11^{-T}   (in place; one N tile covers kb <= 64)
      gemm_f64<64, 64, true, false, false>(s, mr, kb, kb, 1.0, L + r0 + n * k0, n, Li + k0 + n * k0, n, 0.0,
                                           L + r0 + n * k0, n);
      // trailing: A22 -= L21 L21^T  (lower tiles only; K = 64, so small
      // tiles: the update is latency-bound and wants many workgroups), and
      // the next diagonal block's factor in the same launch.  Splitting the
      // update over a second stream instead measured slower: the extra launch
      // and stream wait per step cost more than the 15 us hidden.
      const unsigned tiles = (unsigned)ceil_div(mr, (int64_t)64);
      hipLaunchKernelGGL(syrk_potrf_kernel, dim3(tiles, tiles), dim3(256), 0, s, mr, (int64_t)kb, L + r0 + n * k0, L,
                         n, Li, r0, (int)std::min<int64_t>(mr, NB), status);
      RIMP2_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(zero_upper_kernel, dim3(grid_blocks(n * n, 256)), dim3(256), 0, s, L, n, n);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  {
    // L^{-1} by recursive doubling over the 64-wide diagonal blocks whose
    // inverses potrf_diag already wrote:  for [[P,0],[Q,R]],
    // inv = [[P^-1, 0], [-R^-1 Q P^-1, R^-1]].
    // The groups of one level are independent and (but for a ragged last
    // one) equally shaped with constant pointer strides, so each level is one
    // batched launch per product.
    ScopedTimer t(m.timer, RIMP2_K_TRTRI);
    const int64_t nblk = ceil_div(n, NB);
    double* T = static_cast<double*>(m.work.ensure((size_t)(n * n) * sizeof(double)));
    for (int64_t w = 1; w < nblk; w *= 2) {
      const int64_t half = w * NB, span = 2 * half, nfull = n / span, gs = span * (1 + n);
      if (nfull > 0) {
        // T_g = Q_g P_g^{-1}   (Q = L[mid:hi, lo:mid], P^{-1} = Linv[lo:mid, lo:mid]; T_g is half x half)
        gemm_f64<64, 64, false, false, false>(s, half, half, half, 1.0, L + half, n, Li, n, 0.0, T, half, nfull, gs,
                                              gs, half * half);
        // Linv[mid:hi, lo:mid] = -R_g^{-1} T_g   (R^{-1} lower triangular)
        gemm_f64<64, 64, false, true, false>(s, half, half, half, -1.0, Li + half * (1 + n), n, T, half, 0.0,
                                             Li + half, n, nfull, gs, half * half, gs);
      }
      const int64_t lo = nfull * span, mid = lo + half;
      if (mid < n) {  // ragged last group
        const int64_t sr = n - mid;
        gemm_f64<64, 64, false, false, false>(s, sr, half, half, 1.0, L + mid + n * lo, n, Li + lo + n * lo, n, 0.0,
                                              T, sr);
        gemm_f64<64, 64, false, true, false>(s, sr, half, sr, -1.0, Li + mid + n * mid, n, T, sr, 0.0,
                                             Li + mid + n * lo, n);
      }
    }
  }
  {
    // L^{-1} in A-fragment order for the transform GEMM
    const int64_t MB = round_up(n, TR_TM) / 16, KS = round_up(n, TR_KQ) / 4;
    double* Lp = static_cast<double*>(m.linv_packed.ensure(linv_packed_bytes(n)));
    hipLaunchKernelGGL(pack_linv_kernel, dim3(grid_blocks(KS * MB * 64, 256)), dim3(256), 0, s, Li, n, MB, KS,
                       Lp);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
}

// The context's metric: its own buffers and stream; c->naux records the size.
void metric_factor(Ctx* c, int64_t n, const double* dV_in) {
  metric_factor(ctx_metric_view(c), n, dV_in);
  c->naux = n;
}

void gemm_nt_f64(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  gemm_f64<128, 128, true, false, false>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// dB = Linv * X  (X: n x (nvirt*nocc) device, column-major): X is packed
// into B-fragment order, then one DMA-pipelined FP64 MFMA GEMM.
// Packs ni occupied orbitals of X (dXi: n x nvirt each) into Xp, one slab per
// orbital in the transform's operand order.
static void pack_x_into(const MetricView& m, hipStream_t s, int64_t n, int64_t nvirt, int64_t ni, const double* dXi,
                        void* Xp) {
  PackGeom gx;
  gx.mode = RIMP2_MODE_FP64;
  gx.rb = 16;
  gx.kk = 4;
  gx.planes = 1;
  gx.lane_bytes = 8;
  gx.nrb = round_up(nvirt, TR_TN) / 16;
  gx.nks = round_up(n, TR_KQ) / 4;
  gx.nvp = gx.nrb * 16;
  gx.nqp = gx.nks * 4;
  // on a side stream, a capped grid: it runs beside the Cholesky's chain of
  // small launches instead of competing with it for the whole GPU (512
  // workgroups measured best: a full grid slows the Cholesky, 128 outlasts it)
  pack_B(s == m.stream ? m.timer : nullptr, dXi, n, nvirt, ni, gx, Xp, nullptr, s, s == m.stream ? 0u : 512u);
}

void metric_pack_x(const MetricView& m, hipStream_t s, int64_t n, int64_t nvirt, int64_t nocc, const double* dX) {
  pack_x_into(m, s, n, nvirt, nocc, dX, m.x_packed.ensure(x_packed_bytes(n, nvirt, nocc)));
}

void metric_reserve_x(const MetricView& m, int64_t n, int64_t nvirt, int64_t nocc) {
  m.x_packed.ensure(x_packed_bytes(n, nvirt, nocc));
}

void metric_pack_x_range(const MetricView& m, hipStream_t s, int64_t n, int64_t nvirt, int64_t i0, int64_t ni,
                         const double* dXi) {
  if (m.x_packed.bytes < x_packed_bytes(n, nvirt, i0 + ni))
    fail(RIMP2_ERR_STATE, "packed X holds fewer than %lld orbitals", (long long)(i0 + ni));
  pack_x_into(m, s, n, nvirt, ni, dXi, m.x_packed.as<unsigned char>() + x_packed_bytes(n, nvirt, i0));
}

void metric_transform(const MetricView& m, int64_t n, int64_t nvirt, int64_t nocc, const double* dX, double* B) {
  metric_pack_x(m, m.stream, n, nvirt, nocc, dX);
  metric_transform_packed(m, n, nvirt, nocc, B);
}

// B (n x nvirt x ni) = L^{-1} X for the ni orbitals whose packed X starts at Xp.
static void transform_packed_at(const MetricView& m, int64_t n, int64_t nvirt, int64_t nocc,
                                const unsigned char* Xp, double* B) {
  TransformArgs p;
  p.naux = n;
  p.nvirt = nvirt;
  p.nocc = nocc;
  p.MB = round_up(n, TR_TM) / 16;
  p.KS = round_up(n, TR_KQ) / 4;
  p.NRB = round_up(nvirt, TR_TN) / 16;
  p.MT = round_up(n, TR_TM) / TR_TM;
  p.NT = p.NRB / 8;
  // the transform kernel addresses L^{-1} and one occupied orbital's X with 32-bit offsets
  if ((size_t)p.KS * p.MB * 512 > 0xFFFFFFFFull || (size_t)p.KS * p.NRB * 512 > 0xFFFFFFFFull)
    fail(RIMP2_ERR_ARG, "metric transform operands exceed 4 GiB per slab (naux=%lld nvirt=%lld)", (long long)n,
         (long long)nvirt);
  p.nwu = ((p.MT + 1) / 2) * p.NT * nocc;  // paired row tiles
  p.chunk = ceil_div(p.nwu, 8);
  ScopedTimer t(m.timer, RIMP2_K_TRANSFORM);
  p.Lp = m.linv_packed.as<unsigned char>();
  p.Xp = Xp;
  p.B = B;
  const int64_t nslot = p.chunk < kMaxGridBlocks / 8 ? p.chunk : kMaxGridBlocks / 8;
  hipLaunchKernelGGL(transform_kernel, dim3((unsigned)(8 * nslot)), dim3(512), 0, m.stream, p);
  RIMP2_HIP_CHECK(hipGetLastError());
}

void metric_transform_packed(const MetricView& m, int64_t n, int64_t nvirt, int64_t nocc, double* B) {
  transform_packed_at(m, n, nvirt, nocc, m.x_packed.as<unsigned char>(), B);
}

void metric_transform_range(const MetricView& m, int64_t n, int64_t nvirt, int64_t i0, int64_t ni, double* B) {
  transform_packed_at(m, n, nvirt, ni, m.x_packed.as<unsigned char>() + x_packed_bytes(n, nvirt, i0),
                      B + i0 * n * nvirt);
}



// ---------------------------------------------------------------------------
// Grouped metric pipeline (host side)
// ---------------------------------------------------------------------------
void MetricGroup::release() {
  for (DevBuf* b : {&L, &Linv, &T, &Lp, &Xp, &jobs, &Ls, &Xs, &sig, &ijobs, &sjobs}) b->release();
  blob.clear();
}

// Lays the arenas out, plans the job tables (metric_group.h), sends them to
// the device in one copy and queues one launch per step.
void metric_group_run(MetricGroup& G, hipStream_t s, const std::vector<mg::Item>& items,
                      const std::vector<mg::Copy>& copies, const std::vector<mg::Denom>& denoms, bool i8) {
  for (const mg::Item& it : items) {
    if (i8) break;  // the int8 transform checks its own slab sizes
    const TransformArgs a = mg::transform_args(it.n, it.nvirt, it.nocc);
    if ((size_t)a.KS * a.MB * 512 > 0xFFFFFFFFull || (size_t)a.KS * a.NRB * 512 > 0xFFFFFFFFull)
      fail(RIMP2_ERR_ARG, "metric transform operands exceed 4 GiB per slab (naux=%lld nvirt=%lld)", (long long)it.n,
           (long long)it.nvirt);
  }
  const mg::Layout g = mg::layout(items);
  mg::Arenas ar;
  ar.L = static_cast<unsigned char*>(G.L.ensure(g.ln ? g.ln : 256));
  ar.Linv = static_cast<unsigned char*>(G.Linv.ensure(g.ln ? g.ln : 256));
  ar.T = static_cast<unsigned char*>(G.T.ensure(g.ln ? g.ln : 256));
  ar.Lp = static_cast<unsigned char*>(G.Lp.ensure(g.lpn && !i8 ? g.lpn : 256));
  ar.Xp = static_cast<unsigned char*>(G.Xp.ensure(g.xpn && !i8 ? g.xpn : 256));
  std::vector<mg::GroupStep> steps;
  mg::plan(items, copies, denoms, g, ar, steps, G.blob, !i8);
  if (steps.empty()) return;
  unsigned char* dj = static_cast<unsigned char*>(G.jobs.ensure(G.blob.size()));
  RIMP2_HIP_CHECK(hipMemcpyAsync(dj, G.blob.data(), G.blob.size(), hipMemcpyHostToDevice, s));
  for (const mg::GroupStep& st : steps) {
    const void* jp = dj + st.off;
    if (st.kind >= mg::GL_PANEL && st.kind <= mg::GL_TRI_L && st.total > kMaxGridBlocks)
      fail(RIMP2_ERR_ARG, "grouped metric launch too large (%lld tiles)", (long long)st.total);
    switch (st.kind) {
      case mg::GL_COPY:
        hipLaunchKernelGGL(copy_check_group_kernel, dim3(grid_blocks(st.total, 64 * 256)), dim3(256), 0, s,
                           (const CopyJob*)jp, st.njobs, st.total);
        break;
      case mg::GL_DENOM:
        hipLaunchKernelGGL(denom_group_kernel, dim3((unsigned)st.njobs), dim3(64), 0, s, (const DenomJob*)jp);
        break;
      case mg::GL_ZERO_LINV: RIMP2_HIP_CHECK(hipMemsetAsync(ar.Linv, 0, (size_t)st.total, s)); break;
      case mg::GL_POTRF:
        hipLaunchKernelGGL(potrf_diag_group_kernel, dim3((unsigned)st.njobs), dim3(256), 0, s, (const PotrfJob*)jp);
        break;
      case mg::GL_PANEL:  // L21 = A21 L11^{-T}
        hipLaunchKernelGGL((gemm_f64_group_kernel<64, 64, true, false, false>), dim3((unsigned)st.total), dim3(256), 0,
                           s, (const GemmJob*)jp, st.njobs);
        break;
      case mg::GL_SYRK:  // A22 -= L21 L21^T, lower tiles
        hipLaunchKernelGGL((gemm_f64_group_kernel<64, 64, true, false, true>), dim3((unsigned)st.total), dim3(256), 0,
                           s, (const GemmJob*)jp, st.njobs);
        break;
      case mg::GL_TRI_T:  // T = Q P^{-1}
        hipLaunchKernelGGL((gemm_f64_group_kernel<64, 64, false, false, false>), dim3((unsigned)st.total), dim3(256),
                           0, s, (const GemmJob*)jp, st.njobs);
        break;
      case mg::GL_TRI_L:  // Linv block = -R^{-1} T, R^{-1} lower triangular
        hipLaunchKernelGGL((gemm_f64_group_kernel<64, 64, false, true, false>), dim3((unsigned)st.total), dim3(256),
                           0, s, (const GemmJob*)jp, st.njobs);
        break;
      case mg::GL_LPACK:
        hipLaunchKernelGGL(pack_linv_group_kernel, dim3((unsigned)std::min<int64_t>(st.total, kMaxGridBlocks)),
                           dim3(256), 0, s, (const LinvPackJob*)jp, st.njobs, st.total);
        break;
      case mg::GL_XPACK:  // one workgroup per 64 x 64 tile
        if (st.total > kMaxGridBlocks) fail(RIMP2_ERR_ARG, "grouped X packing too large (%lld tiles)", (long long)st.total);
        hipLaunchKernelGGL(pack_x_group_kernel, dim3((unsigned)st.total), dim3(256), 0, s, (const XPackJob*)jp, st.njobs);
        break;
      case mg::GL_TRANSFORM: {
        const mg::XcdRanges r = mg::transform_xcd_ranges(
            reinterpret_cast<const TransformJob*>(G.blob.data() + st.off), st.njobs, st.total);
        int64_t most = 0;
        for (int x = 0; x < 8; x++) most = std::max(most, r.b[x + 1] - r.b[x]);
        const int64_t nslot = std::max<int64_t>(1, std::min(most, kMaxGridBlocks / 8));
        hipLaunchKernelGGL(transform_group_kernel, dim3((unsigned)(8 * nslot)), dim3(512), 0, s,
                           (const TransformJob*)jp, st.njobs, r);
        break;
      }
    }
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  if (i8) {
    std::vector<I8GroupItem> gi;
    for (size_t k : g.order) {  // largest first, as the plan
      const mg::Item& it = items[k];
      gi.push_back(I8GroupItem{reinterpret_cast<const double*>(ar.Linv + g.lo[k]),
                               reinterpret_cast<double*>(ar.T + g.lo[k]), it.X, it.B, it.n, it.nvirt, it.nocc,
                               it.status});
    }
    transform_i8_group(G, s, gi);
  }
}

}  // namespace rimp2
