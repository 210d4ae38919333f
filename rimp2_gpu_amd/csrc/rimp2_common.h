// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rimp2_hip.h"
#include "metric_group.h"

namespace rimp2 {

// Thrown inside the library only; every extern "C" entry point catches it and
// turns it into a status code + last-error text.
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const char* fmt, ...);

#define RIMP2_HIP_CHECK(expr)                                                                   \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) ::rimp2::fail(RIMP2_ERR_HIP, "%s failed: %s (%s:%d)", #expr,          \
                                        hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

// Grow-only device buffer owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* ensure(size_t need);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Geometry of the packed ("MFMA-operand order") B for one precision mode.
// Global order: [i][kstep][rowblock][plane][64 lanes x lane_bytes].
struct PackGeom {
  int mode = -1;
  int rb = 0;          // rows per MFMA operand fragment (16 or 32)
  int kk = 0;          // k values per fragment (one MFMA K step)
  int planes = 0;      // 1 (fp64/fp32), 2 (bf16 hi/lo), S (int8 slices)
  int lane_bytes = 0;  // bytes each lane holds per fragment
  int64_t nvp = 0;     // padded nvirt (multiple of the pair tile)
  int64_t nqp = 0;     // padded naux (multiple of kk * ksteps per LDS stage)
  int64_t nks = 0;     // nqp / kk
  int64_t nrb = 0;     // nvp / rb
  size_t frag_bytes() const { return (size_t)64 * lane_bytes; }
  size_t bytes_per_occ() const { return (size_t)nks * nrb * planes * frag_bytes(); }
};

struct Timer {
  double total_ms = 0.0;
  int64_t count = 0;
};

// event pair recorded around one launch, resolved lazily by rimp2_timing_get
struct PendingTimer {
  int kid;
  hipEvent_t e0, e1;
};

struct BatchState;
struct MlpState;
struct EosState;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Side stream (default priority) for the X packing that overlaps the
  // Cholesky; two events hand work over (stream -> side, side -> stream).
  hipStream_t side_pack = nullptr;
  hipEvent_t ev_pack_go = nullptr, ev_pack_done = nullptr;
  std::string last_error;
  // metric state
  int64_t naux = 0;
  bool have_metric = false;
  DevBuf V, Linv, work, linv_packed;
  // transform state
  int64_t nvirt = 0, nocc = 0;
  int mode = -1;
  bool have_B = false;
  DevBuf X, B64, packed, scales, x_packed;
  // int8 transform operands of the exact-slice modes (transform_i8.hip)
  DevBuf lt_slices, lt_sigma, x_sigma;
  PackGeom geom;
  int tile = 0;  // pair tile (virtual block) used for the packed layout
  // pair state
  DevBuf eps_o, eps_v, partials, eps_pair, etot, status;
  // an asynchronous entry point queued input checks whose status word no
  // call has read yet (rimp2_sync, or the next blocking call, reads it)
  bool status_unread = false;
  // status word of the allocation-free entry point (in the device's static
  // scratch) while its checks are unread; null otherwise
  int* status_scratch = nullptr;
  // this device's static scratch array (set by rimp2_ctx_create)
  unsigned char* scratch = nullptr;
  // timing
  bool timing = false;
  Timer timers[RIMP2_NUM_KERNELS];
  std::vector<hipEvent_t> ev_pool;
  std::vector<PendingTimer> pending;
  // batch + surrogate
  BatchState* batch = nullptr;
  MlpState* mlp = nullptr;
  // DEC local-basis arrays: work buffers, occupied orbitals per amplitude
  // chunk (0: sized automatically)
  EosState* eos = nullptr;
  int64_t eos_chunk = 0;
  int eos_mode = RIMP2_MODE_FP64;  // amplitude slab precision: FP64 or I8S5
  // host-to-device staging: two pinned chunk buffers (allocated on first
  // use), per buffer an event recorded after its last DMA, and the buffer the
  // next chunk goes to (they alternate across calls too)
  char* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  bool pin_busy[2] = {false, false};
  int pin_next = 0;
};

// Copies a pageable host array to the device, in order on stream s. Copies
// of 1 MiB and more go through the pinned chunk buffers: host threads fill
// one buffer while the DMA engine drains the other. The host array may be
// reused as soon as the call returns.
void h2d_copy(Ctx* c, hipStream_t s, void* dst, const void* src, size_t bytes);

// Records hipEvents around a launch on ctx->stream when timing is on (no-op
// for a null ctx).
// A roctx range over a host scope: `rocprofv3 --marker-trace` shows it on the
// host timeline (an asynchronous call's range covers its enqueue, a blocking
// call's range its whole run).  With no tool attached, push and pop are a
// dispatch-table call each.
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

// Event timer of one kernel class (RIMP2_K_*) around the launches of a scope,
// on the context stream when the context's timing is on; the scope is also a
// roctx range named after the class.
struct ScopedTimer {
  Ctx* c;
  int kid;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ScopedTimer(Ctx* ctx, int k);
  ~ScopedTimer();
};

// --- metric (metric.hip) ---
// The stream and scratch buffers one metric pipeline (Cholesky, triangular
// inverse, transform) runs on.
struct MetricView {
  hipStream_t stream;
  DevBuf &L, &Linv, &work, &linv_packed, &x_packed;
  int* status;  // device: [0] error bits, [1] first failing pivot (1-based)
  Ctx* timer;   // context whose kernel timers record the launches (or null)
};
MetricView ctx_metric_view(Ctx* c);
// L L^T = V, L^{-1}, and L^{-1} in transform-operand order (dV_in: device, n x n)
void metric_factor(const MetricView& m, int64_t n, const double* dV_in);
// dB[n, nvirt, nocc] = L^{-1} X  (after metric_factor on the same view)
void metric_transform(const MetricView& m, int64_t n, int64_t nvirt, int64_t nocc, const double* dX, double* dB);
// The two halves of metric_transform: X into B-fragment order (m.x_packed) on
// stream s, then the product from the packed X on m.stream.
void metric_pack_x(const MetricView& m, hipStream_t s, int64_t n, int64_t nvirt, int64_t nocc, const double* dX);
void metric_transform_packed(const MetricView& m, int64_t n, int64_t nvirt, int64_t nocc, double* dB);
// The same for orbitals [i0, i0 + ni) only (dXi points at orbital i0 of X,
// dB is the whole B); metric_reserve_x sizes m.x_packed for all orbitals first.
void metric_reserve_x(const MetricView& m, int64_t n, int64_t nvirt, int64_t nocc);
void metric_pack_x_range(const MetricView& m, hipStream_t s, int64_t n, int64_t nvirt, int64_t i0, int64_t ni,
                         const double* dXi);
void metric_transform_range(const MetricView& m, int64_t n, int64_t nvirt, int64_t i0, int64_t ni, double* dB);
void metric_factor(Ctx* c, int64_t naux, const double* dV_in);  // dV_in: device
// Metric transform on int8 MFMA for the exact-slice modes (transform_i8.hip):
// L^{-1} (after metric_factor) and X each split into six int8 slices per
// element with one scale per column, B = L^{-1} X from the 21 slice products
// of order <= 5 with an FP64 epilogue (relative error ~1e-12 of B's scale).
// metric_slice_linv slices the context's L^{-1} on stream s (c->work is its
// scratch); metric_reserve_x_i8 sizes the X slices of nocc orbitals;
// metric_slice_x_range slices orbitals [i0, i0 + ni) (dXi points at orbital
// i0 of X) on stream s; metric_transform_i8_range writes those orbitals of
// dB (the whole B) on c->stream.
size_t metric_i8_x_bytes(int64_t n, int64_t nvirt, int64_t nocc);
size_t metric_i8_linv_bytes(int64_t n);
void metric_slice_linv(Ctx* c, hipStream_t s, int64_t n);
void metric_reserve_x_i8(Ctx* c, int64_t n, int64_t nvirt, int64_t nocc);
void metric_slice_x_range(Ctx* c, hipStream_t s, int64_t n, int64_t nvirt, int64_t i0, int64_t ni,
                          const double* dXi);
void metric_transform_i8_range(Ctx* c, int64_t n, int64_t nvirt, int64_t i0, int64_t ni, double* dB);
// Grouped metric pipeline: the Cholesky, inverse and transform of many
// fragments with one launch per step for the whole group (metric_group.h).
// Arenas and job tables of a grouped pipeline (grow-only, reused call after
// call; blob keeps the host copy of the job tables until the caller has
// synchronized the stream).
struct MetricGroup {
  DevBuf L, Linv, T, Lp, Xp, jobs;
  std::vector<unsigned char> blob;
  // the int8 transform of the exact-slice modes: slice arenas, scales, job tables
  DevBuf Ls, Xs, sig, ijobs, sjobs;
  std::vector<unsigned char> iblob, sblob;
  void release();
};
// Queues on s: the copies, the denominator checks, then the metric pipeline
// of every item.  The caller synchronizes s before the next call on G.
// i8: the transform runs on int8 MFMA with six slices per operand
// (transform_i8_group) instead of FP64.
void metric_group_run(MetricGroup& G, hipStream_t s, const std::vector<mg::Item>& items,
                      const std::vector<mg::Copy>& copies, const std::vector<mg::Denom>& denoms, bool i8 = false);
// One fragment of the grouped int8 transform: L^{-1} (device, n x n), scratch
// T (n x n, for L^{-T}), X and B [n, nvirt, nocc]; a NaN or Inf in X sets bit
// 4 of status.
struct I8GroupItem {
  const double* Li;
  double* T;
  const double* X;
  double* B;
  int64_t n, nvirt, nocc;
  int* status;
};
void transform_i8_group(MetricGroup& G, hipStream_t s, const std::vector<I8GroupItem>& items);
// C = alpha * A * B^T + beta * C (column-major FP64 MFMA GEMM)
void gemm_nt_f64(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc);

// --- packing (pack.hip) ---
PackGeom make_geom(int mode, int64_t naux, int64_t nvirt, int tile);
// max_blocks (0: no limit) caps the grid, for packing that shares the GPU
// with a latency-bound chain on another stream.
void pack_B(Ctx* c, const double* dB, int64_t naux, int64_t nvirt, int64_t nocc, const PackGeom& g, void* dst,
            double* dscales, hipStream_t s, unsigned max_blocks = 0);
int mode_slices(int mode);
// Six int8 slices of the columns of src [nq, nrows, norb] (nq fastest), each
// column scaled by its largest |value| (sigma[orb * nrp + row]; 0 on padding
// rows): the exact-slice packing (32 rows x 32 k-values per fragment,
// orbital-major) with nks 32-deep k-steps and nrp rows (multiple of 32).
// max_blocks (0: no cap) caps the slicing grid.
void pack_slices6(hipStream_t s, const double* src, int64_t nq, int64_t nrows, int64_t norb, int64_t nks,
                  int64_t nrp, double* sigma, void* dst, unsigned max_blocks = 0);
// The same for many arrays in two launches (row scales, then slices); a
// column holding a NaN or Inf sets bit 4 of its item's status (if not null).
// jobs / blob hold the job tables until the caller synchronizes s.
struct Slice6Item {
  const double* src;
  int64_t nq, nrows, norb, nks, nrp;
  double* sigma;
  void* dst;
  int* status;
};
void pack_slices6_group(const std::vector<Slice6Item>& items, hipStream_t s, DevBuf& jobs,
                        std::vector<unsigned char>& blob);
// One fragment of a grouped packing (device arrays).
struct PackItem {
  const double* B;      // FP64 B [naux, nvirt, nocc]
  int64_t naux, nvirt, nocc;
  PackGeom g;           // geometry of the mode and pair tile
  void* dst;            // packed operands
  double* sigma;        // exact-slice row scales [nocc * nvp] (else unused)
  const double* ev;     // virtual energies [nvirt]
  double* ev_padded;    // out: [nvp], 1.0 on padding rows
};
// Packs every item with three launches in all (row scales, packing, padded
// virtual energies); jobs / blob hold the job tables (the caller synchronizes
// s before the next call on them).
void pack_B_group(const std::vector<PackItem>& items, int mode, hipStream_t s, DevBuf& jobs,
                  std::vector<unsigned char>& blob);

// --- pair kernel (pair.hip) ---
int choose_tile(int mode, int64_t nvirt, int64_t nocc);
// Whether the pair kernel of (mode, tile) carries two diagonal blocks per unit.
bool pair_packs_diagonal(int mode, int tile);
void pair_energies_device(Ctx* c, const PackGeom& g, const void* packed, const double* scales, int64_t nocc,
                          int64_t nvirt, const double* d_eps_o, const double* d_eps_v, double* d_eps_pair,
                          double* d_etot, const int64_t* d_pair_list, int64_t npairs_listed);
// The same on stream s with caller-provided scratch (allocates nothing):
// evp holds nvp doubles, part npairs * units + npairs + 1 doubles;
// pair_scratch_bytes gives their total (partials first, then evp at
// *evp_offset).  timer may be null (no event timers).
size_t pair_scratch_bytes(const PackGeom& g, int tile, int64_t npairs, size_t* evp_offset);
void pair_energies_core(Ctx* timer, hipStream_t s, int tile, const PackGeom& g, const void* packed,
                        const double* scales, int64_t nocc, int64_t nvirt, const double* d_eps_o,
                        const double* d_eps_v, double* d_eps_pair, double* d_etot, const int64_t* d_pair_list,
                        int64_t npairs_listed, double* evp, double* part);
void check_inputs_device(Ctx* c, const double* p, int64_t n, const char* what);
// Device status word of the context (ctx.hip): check_status waits for the
// stream and turns set bits into an error (clearing them); check_pending does
// so only when an asynchronous call left them unread; reset_status clears the
// word at the start of a call (after check_pending).
// reset_status_async is the reset of an asynchronous call: it keeps bits no
// call has read yet.
void check_status(Ctx* c);
void check_pending(Ctx* c);
void reset_status(Ctx* c);
void reset_status_async(Ctx* c);
// Amplitude slab of occupied orbitals [i0, i0 + nchunk) from the FP64 or
// i8s5 packing (tile 64; sigma: i8s5 row scales, else null):
// T[a + nv b + nv^2 (j - i0) + nv^2 (nocc - i0) (i - i0)] =
// w_ij V^{ij}_ab / (e_i + e_j - e_b - e_a), j >= i, w = 1 (i < j), 1/2 (i = j).
void pair_amplitudes_device(hipStream_t s, const PackGeom& g, const void* packed, const double* sigma, int64_t nocc,
                            int64_t nvirt, const double* d_eo, const double* d_ev, int64_t i0, int64_t nchunk,
                            double* T);
// Status bits on a given stream: 4 = a non-finite value among p[0..n);
// 8 = max(eo) >= min(ev) (some pair denominator >= 0).
void check_finite_on(hipStream_t s, int* status, const double* p, int64_t n, unsigned max_blocks = 8192);
void check_denom_on(hipStream_t s, int* status, const double* eo, int64_t nocc, const double* ev, int64_t nvirt);

// --- allocation-free entry point (noalloc.hip) ---
// Device address of this device's static scratch array, looked up once per
// device (the calling thread must have made `device` current).
unsigned char* noalloc_scratch(int device);

// --- batch / mlp cleanup ---
void batch_destroy(BatchState* b);
void mlp_destroy(MlpState* m);
void eos_destroy(EosState* s);

// helpers
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// A dispatch counts its work-items in 32 bits, so a grid of blocks x threads
// >= 2^32 wraps and silently runs only part of the work.  Launches over data-
// sized ranges therefore cap the grid (grid_blocks) and their kernels loop
// grid-stride (RIMP2_GRID_STRIDE); work-unit kernels loop over their units.
constexpr int64_t kMaxGridBlocks = int64_t(1) << 20;
inline unsigned grid_blocks(int64_t items, int64_t threads) {
  const int64_t b = ceil_div(items, threads);
  return (unsigned)(b < 1 ? 1 : (b > kMaxGridBlocks ? kMaxGridBlocks : b));
}
#define RIMP2_GRID_STRIDE(t, n)                                                        \
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)(n); \
       t += (int64_t)gridDim.x * blockDim.x)

}  // namespace rimp2

// Wraps an extern "C" body: converts rimp2::Error into a status code.  The
// call is a roctx range named after the entry point.
#define RIMP2_API_BEGIN                             \
  ::rimp2::TraceRange rimp2_api_range_(__func__); \
  try {
#define RIMP2_API_END(ctxp)                                         \
  }                                                                 \
  catch (const ::rimp2::Error& e) {                                 \
    ::rimp2::set_error((ctxp), e.msg);                              \
    return e.code;                                                  \
  }                                                                 \
  catch (...) {                                                     \
    ::rimp2::set_error((ctxp), std::string("unknown C++ exception")); \
    return RIMP2_ERR_HIP;                                           \
  }                                                                 \
  return RIMP2_OK;

namespace rimp2 {
void set_error(void* ctx, const std::string& msg);
Ctx* as_ctx(void* ctx);
}  // namespace rimp2
