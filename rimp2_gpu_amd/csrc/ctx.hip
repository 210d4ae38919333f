// C-ABI entry points: context lifetime, device-memory query, errors, timers,
// and the blocking / device-resident RI-MP2 energy path.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "rimp2_common.h"

namespace rimp2 {

static std::string g_last_error;
static std::mutex g_err_mu;

[[noreturn]] void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error{code, std::string(buf)};
}

void set_error(void* ctx, const std::string& msg) {
  if (ctx) static_cast<Ctx*>(ctx)->last_error = msg;
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_last_error = msg;
}

Ctx* as_ctx(void* ctx) {
  if (!ctx) fail(RIMP2_ERR_ARG, "null context");
  Ctx* c = static_cast<Ctx*>(ctx);
  RIMP2_HIP_CHECK(hipSetDevice(c->device));
  return c;
}

// Fault injection for tests of the allocation-failure path: with
// RIMP2_FAULT_ALLOC_BYTES=n in the environment, every allocation of the
// library's own device buffers of at least n bytes fails as hipMalloc would.
// Read at each allocation, so a caller can set and clear it between calls.
static bool injected_alloc_failure(size_t want) {
  const char* v = getenv("RIMP2_FAULT_ALLOC_BYTES");
  if (!v || !*v) return false;
  char* end = nullptr;
  const unsigned long long n = strtoull(v, &end, 10);
  return end != v && want >= n;
}

void* DevBuf::ensure(size_t need) {
  if (need <= bytes && ptr) return ptr;
  release();
  size_t want = need < 256 ? 256 : need;
  const bool injected = injected_alloc_failure(want);
  hipError_t e = injected ? hipErrorOutOfMemory : hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    fail(RIMP2_ERR_NOMEM, "device allocation of %zu bytes failed: %s%s", want, hipGetErrorString(e),
         injected ? " (injected by RIMP2_FAULT_ALLOC_BYTES)" : "");
  }
  bytes = want;
  return ptr;
}

// Frees after the device is idle: work queued on any stream (side streams
// included) may still read the buffer.
void DevBuf::release() {
  if (ptr) {
    (void)hipDeviceSynchronize();
    (void)hipFree(ptr);
  }
  ptr = nullptr;
  bytes = 0;
}

static hipEvent_t get_event(Ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  RIMP2_HIP_CHECK(hipEventCreate(&e));
  return e;
}

static const char* const kTimerNames[RIMP2_NUM_KERNELS] = {"pair",  "transform", "pack",     "cholesky",
                                                            "trtri", "reduce",    "batch",    "mlp",
                                                            "features", "eos",    "eos_ampl"};

ScopedTimer::ScopedTimer(Ctx* ctx, int k) : c(ctx), kid(k) {
  roctxRangePushA(k >= 0 && k < RIMP2_NUM_KERNELS ? kTimerNames[k] : "kernel");
  if (!c || !c->timing) return;
  e0 = get_event(c);
  e1 = get_event(c);
  RIMP2_HIP_CHECK(hipEventRecord(e0, c->stream));
}

ScopedTimer::~ScopedTimer() {
  if (e0 && hipEventRecord(e1, c->stream) == hipSuccess) c->pending.push_back({kid, e0, e1});
  roctxRangePop();
}

static void resolve_timers(Ctx* c) {
  auto& q = c->pending;
  for (auto& t : q) {
    RIMP2_HIP_CHECK(hipEventSynchronize(t.e1));
    float ms = 0.f;
    RIMP2_HIP_CHECK(hipEventElapsedTime(&ms, t.e0, t.e1));
    c->timers[t.kid].total_ms += ms;
    c->timers[t.kid].count += 1;
    c->ev_pool.push_back(t.e0);
    c->ev_pool.push_back(t.e1);
  }
  q.clear();
}

static void check_host_finite(const double* x, int64_t n, const char* what) {
  for (int64_t k = 0; k < n; k++)
    if (!std::isfinite(x[k])) fail(RIMP2_ERR_NUMERIC, "%s[%lld] is not finite", what, (long long)k);
}

constexpr size_t kPinChunk = (size_t)16 << 20;

static void parallel_memcpy(char* d, const char* s, size_t n) {
  const int T = std::max(1, std::min(omp_get_max_threads(), (int)(n >> 18)));  // >= 256 KiB per thread
  const size_t piece = (n + T - 1) / T;
#pragma omp parallel for num_threads(T) schedule(static)
  for (int t = 0; t < T; t++) {
    const size_t o = (size_t)t * piece;
    if (o < n) memcpy(d + o, s + o, std::min(piece, n - o));
  }
}

void h2d_copy(Ctx* c, hipStream_t s, void* dst, const void* src, size_t bytes) {
  if (bytes < ((size_t)1 << 20)) {
    RIMP2_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  for (int b = 0; b < 2; b++) {
    if (!c->pin[b]) RIMP2_HIP_CHECK(hipHostMalloc((void**)&c->pin[b], kPinChunk, hipHostMallocDefault));
    if (!c->pin_ev[b]) RIMP2_HIP_CHECK(hipEventCreateWithFlags(&c->pin_ev[b], hipEventDisableTiming));
  }
  const char* from = static_cast<const char*>(src);
  char* to = static_cast<char*>(dst);
  for (size_t off = 0; off < bytes; off += kPinChunk) {
    const int b = c->pin_next;
    c->pin_next ^= 1;
    const size_t n = std::min(kPinChunk, bytes - off);
    if (c->pin_busy[b]) RIMP2_HIP_CHECK(hipEventSynchronize(c->pin_ev[b]));
    c->pin_busy[b] = false;
    parallel_memcpy(c->pin[b], from + off, n);
    RIMP2_HIP_CHECK(hipMemcpyAsync(to + off, c->pin[b], n, hipMemcpyHostToDevice, s));
    RIMP2_HIP_CHECK(hipEventRecord(c->pin_ev[b], s));
    c->pin_busy[b] = true;
  }
}

static void check_host_denominators(const double* eo, int64_t nocc, const double* ev, int64_t nvirt) {
  double mo = -INFINITY, mv = INFINITY;
  for (int64_t i = 0; i < nocc; i++) mo = std::fmax(mo, eo[i]);
  for (int64_t a = 0; a < nvirt; a++) mv = std::fmin(mv, ev[a]);
  if (!(2.0 * mo - 2.0 * mv < 0.0))
    fail(RIMP2_ERR_NUMERIC, "orbital-energy denominator e_i+e_j-e_a-e_b >= 0 (max e_occ=%g, min e_virt=%g)", mo, mv);
}

// Reads and clears the device status word; converts set bits to an error.
// A set bit also invalidates the metric and the transformed B the context
// holds: they were formed from the rejected inputs.
// The context's device status word, zeroed when first allocated (a read
// before any reset must see no error bits).
static int* status_word(Ctx* c) {
  const bool fresh = !c->status.ptr;
  int* d = static_cast<int*>(c->status.ensure(64));
  if (fresh) RIMP2_HIP_CHECK(hipMemsetAsync(d, 0, 64, c->stream));
  return d;
}

// The allocation-free entry point keeps its status word in the device's static
// scratch; both words are read while its checks are pending.
void check_status(Ctx* c) {
  int st[2] = {0, 0}, sx[2] = {0, 0};
  int* d = (c->status_scratch && !c->status.ptr) ? nullptr : status_word(c);
  int* dx = c->status_scratch;
  if (d) RIMP2_HIP_CHECK(hipMemcpyAsync(st, d, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  if (c->status_scratch)
    RIMP2_HIP_CHECK(hipMemcpyAsync(sx, c->status_scratch, sizeof(sx), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  c->status_unread = false;
  c->status_scratch = nullptr;
  if (!st[0] && sx[0]) {
    st[0] = sx[0];
    st[1] = sx[1];
  } else {
    st[0] |= sx[0];
  }
  if (st[0]) {
    if (d) RIMP2_HIP_CHECK(hipMemsetAsync(d, 0, 64, c->stream));
    if (dx && sx[0]) RIMP2_HIP_CHECK(hipMemsetAsync(dx, 0, 64, c->stream));
    c->have_metric = false;
    c->have_B = false;
  }
  if (st[0] & 4) fail(RIMP2_ERR_NUMERIC, "non-finite value in an input array");
  if (st[0] & 2) fail(RIMP2_ERR_NOT_SPD, "metric is not positive definite (pivot %d <= 0)", st[1]);
  if (st[0] & 8) fail(RIMP2_ERR_NUMERIC, "orbital-energy denominator e_i+e_j-e_a-e_b >= 0");
}

// An asynchronous entry point's input checks are read by rimp2_sync; a
// blocking call that follows it without that sync reads them first, so an
// error is reported (once) instead of being cleared unseen.
void check_pending(Ctx* c) {
  if (c->status_unread) check_status(c);
}

void denom_check_device(Ctx* c, const double* eo, int64_t nocc, const double* ev, int64_t nvirt);

// Clears the device status word at the start of a blocking call (after
// reporting any error an earlier asynchronous call left unread).
void reset_status(Ctx* c) {
  check_pending(c);
  RIMP2_HIP_CHECK(hipMemsetAsync(status_word(c), 0, 64, c->stream));
}

// The same for an asynchronous call: bits that no call has read yet are kept
// (the new checks add theirs), so rimp2_sync reports an error of any call
// queued since the last read, and queuing never waits for the stream.
void reset_status_async(Ctx* c) {
  if (c->status_unread) return;
  RIMP2_HIP_CHECK(hipMemsetAsync(status_word(c), 0, 64, c->stream));
}

static void validate_mode(int mode) {
  if (mode < 0 || mode >= RIMP2_NUM_MODES) fail(RIMP2_ERR_ARG, "unknown precision mode %d", mode);
}

static void validate_dims(int64_t naux, int64_t nvirt, int64_t nocc) {
  if (naux <= 0 || nvirt <= 0 || nocc <= 0)
    fail(RIMP2_ERR_ARG, "dimensions must be positive (naux=%lld nvirt=%lld nocc=%lld)", (long long)naux,
         (long long)nvirt, (long long)nocc);
  if (naux > 65536 || nvirt > 65536 || nocc > 65536) fail(RIMP2_ERR_ARG, "dimension above 65536");
}

// transform + pack from a device X; the metric must already be factored.
static void check_slice_bound(int64_t naux, int mode) {
  // padded to at most 16511 auxiliary rows: 5 * 127^2 * 16511 < 2^31
  if (mode_slices(mode) && naux > 16384)
    fail(RIMP2_ERR_ARG, "exact-slice modes support naux <= 16384 (int32 accumulation bound)");
}

// Refuses an energy whose device buffers cannot fit before anything is
// allocated or read: the buffers the path needs (metric factor, inverse and
// its packing, packed X, FP64 B, the mode's packed B and row scales, the
// staging ring of the host path) against hipMemGetInfo's free memory plus what
// the context already holds (grow-only buffers are reused or replaced).
static void require_device_memory(Ctx* c, int64_t naux, int64_t nvirt, int64_t nocc, int mode, bool from_host) {
  const size_t n2 = (size_t)naux * (size_t)naux * sizeof(double);
  const size_t b64 = (size_t)naux * (size_t)nvirt * (size_t)nocc * sizeof(double);
  const PackGeom g = make_geom(mode, naux, nvirt, choose_tile(mode, nvirt, nocc));
  size_t need = 3 * n2 + mg::linv_packed_bytes(naux) + mg::x_packed_bytes(naux, nvirt, nocc) + b64 +
                g.bytes_per_occ() * (size_t)nocc + (size_t)g.nvp * (size_t)nocc * sizeof(double);
  if (mode_slices(mode)) {  // the int8 transform's slices (X's share the packed-X buffer)
    need += metric_i8_linv_bytes(naux) + (size_t)(naux + 64) * sizeof(double) +
            (size_t)(nvirt + 64) * (size_t)nocc * sizeof(double);
    const size_t xi8 = metric_i8_x_bytes(naux, nvirt, nocc), x64 = mg::x_packed_bytes(naux, nvirt, nocc);
    if (xi8 > x64) need += xi8 - x64;
  }
  if (from_host) {
    const size_t per = (size_t)naux * (size_t)nvirt * sizeof(double);
    need += 2 * std::max(per, std::min(b64, (size_t)32 << 20));
  }
  size_t held = 0;
  for (const DevBuf* b : {&c->V, &c->Linv, &c->work, &c->linv_packed, &c->X, &c->B64, &c->packed, &c->scales,
                          &c->x_packed, &c->lt_slices, &c->lt_sigma, &c->x_sigma})
    held += b->bytes;
  size_t free_b = 0, total_b = 0;
  RIMP2_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  if (need > free_b + held)
    fail(RIMP2_ERR_NOMEM, "energy for naux=%lld nvirt=%lld nocc=%lld needs %.2f GiB of device memory; %.2f GiB are "
         "free and the context holds %.2f GiB", (long long)naux, (long long)nvirt, (long long)nocc,
         need / 1073741824.0, free_b / 1073741824.0, held / 1073741824.0);
}

// Packs the FP64 B held in c->B64 ([c->naux, nvirt, nocc]) for the pair kernel.
static void pack_current_B(Ctx* c, int64_t nvirt, int64_t nocc, int mode) {
  c->tile = choose_tile(mode, nvirt, nocc);
  PackGeom g = make_geom(mode, c->naux, nvirt, c->tile);
  const size_t bytes = g.bytes_per_occ() * (size_t)nocc;
  void* dst = c->packed.ensure(bytes);
  double* sc = nullptr;
  if (mode_slices(mode)) sc = static_cast<double*>(c->scales.ensure((size_t)(g.nvp * nocc) * sizeof(double)));
  pack_B(c, c->B64.as<double>(), c->naux, nvirt, nocc, g, dst, sc, c->stream);
  c->geom = g;
  c->nvirt = nvirt;
  c->nocc = nocc;
  c->mode = mode;
  c->have_B = true;
}

// Starts the side stream behind everything queued so far on the context
// stream (earlier readers of X and of the packed X, the status reset).
static void side_after_stream(Ctx* c) {
  RIMP2_HIP_CHECK(hipEventRecord(c->ev_pack_go, c->stream));
  RIMP2_HIP_CHECK(hipStreamWaitEvent(c->side_pack, c->ev_pack_go, 0));
}

// (P|ia) from host memory into c->B64 = L^{-1} X, by chunks of occupied
// orbitals (about 32 MiB each). Each chunk is copied, checked for NaN/Inf and
// packed (the exact-slice modes: sliced for the int8 transform) on the side stream (which side_after_stream has started), then
// transformed on the context stream, behind whatever that stream holds by then
// (the metric factorisation in rimp2_energy), while the host copies the next
// chunk. A non-finite value is reported by the next check_status.
// The raw chunks alternate between two slots of c->X: the side stream is in
// order, so the copy of chunk k + 2 starts only after chunk k has been checked
// and packed (the transform reads the packed copy), and the raw (P|ia) never
// needs more than two chunks of device memory.
static void transform_from_host(Ctx* c, int64_t nvirt, int64_t nocc, const double* Pia, int mode) {
  const int64_t n = c->naux;
  const int64_t per = n * nvirt;  // elements per occupied orbital
  const int64_t step = std::min(nocc, std::max<int64_t>(1, ((int64_t)32 << 20) / (per * (int64_t)sizeof(double))));
  const int64_t slots = std::min<int64_t>(2, ceil_div(nocc, step));
  double* dX = static_cast<double*>(c->X.ensure((size_t)(per * step * slots) * sizeof(double)));
  double* B = static_cast<double*>(c->B64.ensure((size_t)(per * nocc) * sizeof(double)));
  const MetricView m = ctx_metric_view(c);
  const bool i8 = mode_slices(mode) != 0;
  if (i8) {
    metric_reserve_x_i8(c, n, nvirt, nocc);
    metric_slice_linv(c, c->stream, n);
  } else {
    metric_reserve_x(m, n, nvirt, nocc);
  }
  int* status = static_cast<int*>(c->status.ensure(64));
  for (int64_t i0 = 0, k = 0; i0 < nocc; i0 += step, k++) {
    const int64_t ni = std::min(step, nocc - i0);
    double* dXi = dX + (k % slots) * step * per;
    h2d_copy(c, c->side_pack, dXi, Pia + i0 * per, (size_t)(ni * per) * sizeof(double));
    check_finite_on(c->side_pack, status, dXi, ni * per, 512);
    if (i8) metric_slice_x_range(c, c->side_pack, n, nvirt, i0, ni, dXi);
    else metric_pack_x_range(m, c->side_pack, n, nvirt, i0, ni, dXi);
    RIMP2_HIP_CHECK(hipEventRecord(c->ev_pack_done, c->side_pack));
    RIMP2_HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_pack_done, 0));
    if (i8) metric_transform_i8_range(c, n, nvirt, i0, ni, B);
    else metric_transform_range(m, n, nvirt, i0, ni, B);
  }
}

// Already-fitted coefficients (B = Calpha, no metric) from a device array.
// A factored metric of another naux no longer matches the context.
// With check_now, the copied coefficients are scanned on the device and a
// non-finite value fails the call before anything is packed.
static void load_fitted_device(Ctx* c, int64_t naux, int64_t nvirt, int64_t nocc, const double* dB, int mode,
                               hipMemcpyKind kind, bool check_now = false) {
  check_slice_bound(naux, mode);
  c->have_B = false;
  if (naux != c->naux) c->have_metric = false;
  c->naux = naux;
  const size_t bytes = (size_t)(naux * nvirt * nocc) * sizeof(double);
  double* B = static_cast<double*>(c->B64.ensure(bytes));
  if (kind == hipMemcpyHostToDevice) h2d_copy(c, c->stream, B, dB, bytes);
  else RIMP2_HIP_CHECK(hipMemcpyAsync(B, dB, bytes, kind, c->stream));
  if (check_now) {
    check_inputs_device(c, B, naux * nvirt * nocc, "Calpha");
    check_status(c);
  }
  pack_current_B(c, nvirt, nocc, mode);
}

}  // namespace rimp2

using namespace rimp2;

static void destroy_streams(Ctx* c) {
  for (hipEvent_t e : {c->ev_pack_go, c->ev_pack_done})
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t st : {c->stream, c->side_pack})
    if (st) (void)hipStreamDestroy(st);
}

extern "C" {

int rimp2_ctx_create(int device, void** ctx) {
  RIMP2_API_BEGIN
  if (!ctx) fail(RIMP2_ERR_ARG, "null ctx pointer");
  *ctx = nullptr;
  int ndev = 0;
  RIMP2_HIP_CHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) fail(RIMP2_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
  RIMP2_HIP_CHECK(hipSetDevice(device));
  Ctx* c = new Ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side_pack, hipStreamNonBlocking);
  for (hipEvent_t* ev : {&c->ev_pack_go, &c->ev_pack_done})
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  // the "done" event starts recorded (on the idle side stream), so a first wait on it passes
  if (e == hipSuccess) e = hipEventRecord(c->ev_pack_done, c->side_pack);
  if (e != hipSuccess) {
    destroy_streams(c);
    delete c;
    fail(RIMP2_ERR_HIP, "stream or event creation failed: %s", hipGetErrorString(e));
  }
  try {
    c->scratch = noalloc_scratch(device);
  } catch (...) {
    destroy_streams(c);
    delete c;
    throw;
  }
  *ctx = c;
  RIMP2_API_END(nullptr)
}

int rimp2_ctx_destroy(void* ctx) {
  RIMP2_API_BEGIN
  if (!ctx) return RIMP2_OK;
  Ctx* c = as_ctx(ctx);
  (void)hipStreamSynchronize(c->stream);
  if (c->side_pack) (void)hipStreamSynchronize(c->side_pack);
  resolve_timers(c);
  batch_destroy(c->batch);
  mlp_destroy(c->mlp);
  eos_destroy(c->eos);
  for (DevBuf* b : {&c->V, &c->Linv, &c->work, &c->linv_packed, &c->X, &c->B64, &c->packed, &c->scales, &c->x_packed,
                    &c->lt_slices, &c->lt_sigma, &c->x_sigma, &c->eps_o, &c->eps_v,
                    &c->partials, &c->eps_pair, &c->etot, &c->status})
    b->release();
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (int b = 0; b < 2; b++) {
    if (c->pin[b]) (void)hipHostFree(c->pin[b]);
    if (c->pin_ev[b]) (void)hipEventDestroy(c->pin_ev[b]);
  }
  destroy_streams(c);
  delete c;
  RIMP2_API_END(nullptr)
}

int rimp2_get_dev_mem(size_t* total, size_t* free_bytes) {
  RIMP2_API_BEGIN
  if (!total || !free_bytes) fail(RIMP2_ERR_ARG, "null output pointer");
  size_t f = 0, t = 0;
  RIMP2_HIP_CHECK(hipMemGetInfo(&f, &t));
  *total = t;
  *free_bytes = f;
  RIMP2_API_END(nullptr)
}

void get_dev_mem(size_t* total, size_t* free_bytes) {
  size_t t = 0, f = 0;
  if (rimp2_get_dev_mem(&t, &f) != RIMP2_OK) t = f = 0;
  if (total) *total = t;
  if (free_bytes) *free_bytes = f;
}

int rimp2_last_error(void* ctx, char* buf, size_t n) {
  if (!buf || n == 0) return RIMP2_ERR_ARG;
  std::string m;
  if (ctx) m = static_cast<Ctx*>(ctx)->last_error;
  else {
    std::lock_guard<std::mutex> lk(g_err_mu);
    m = g_last_error;
  }
  size_t k = m.size() < n - 1 ? m.size() : n - 1;
  memcpy(buf, m.data(), k);
  buf[k] = '\0';
  return RIMP2_OK;
}

int rimp2_metric(void* ctx, int64_t naux, const double* V) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (naux <= 0 || naux > 65536) fail(RIMP2_ERR_ARG, "naux=%lld out of range", (long long)naux);
  if (!V) fail(RIMP2_ERR_ARG, "null V");
  c->have_metric = false;
  c->have_B = false;
  reset_status(c);
  double* d = static_cast<double*>(c->V.ensure((size_t)(naux * naux) * sizeof(double)));
  h2d_copy(c, c->stream, d, V, (size_t)(naux * naux) * sizeof(double));
  // host arrays are scanned for NaN/Inf on the device once copied: a NaN
  // never reaches the factorization
  check_inputs_device(c, d, naux * naux, "V");
  check_status(c);
  metric_factor(c, naux, d);
  check_status(c);
  c->have_metric = true;
  RIMP2_API_END(ctx)
}

int rimp2_transform(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Pia, int mode) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  validate_dims(naux, nvirt, nocc);
  validate_mode(mode);
  if (!Pia) fail(RIMP2_ERR_ARG, "null Pia");
  if (!c->have_metric || c->naux != naux) fail(RIMP2_ERR_STATE, "rimp2_metric with naux=%lld must run first", (long long)naux);
  check_slice_bound(naux, mode);
  reset_status(c);
  c->have_B = false;
  side_after_stream(c);
  transform_from_host(c, nvirt, nocc, Pia, mode);
  check_status(c);
  pack_current_B(c, nvirt, nocc, mode);
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_load_fitted(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Calpha, int mode) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  validate_dims(naux, nvirt, nocc);
  validate_mode(mode);
  if (!Calpha) fail(RIMP2_ERR_ARG, "null Calpha");
  reset_status(c);
  load_fitted_device(c, naux, nvirt, nocc, Calpha, mode, hipMemcpyHostToDevice, true);
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_energy_fitted(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Calpha,
                        const double* eps_o, const double* eps_v, int mode, double* eps_pair, double* e_tot) {
  int st = rimp2_load_fitted(ctx, naux, nvirt, nocc, Calpha, mode);
  if (st) return st;
  return rimp2_pair_energies(ctx, nocc, nvirt, eps_o, eps_v, mode, eps_pair, e_tot);
}

int rimp2_energy_fitted_device(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* d_Calpha,
                               const double* d_eps_o, const double* d_eps_v, int mode, double* d_eps_pair,
                               double* d_e_tot) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  validate_dims(naux, nvirt, nocc);
  validate_mode(mode);
  if (!d_Calpha || !d_eps_o || !d_eps_v || !d_e_tot) fail(RIMP2_ERR_ARG, "null device pointer");
  reset_status_async(c);
  check_inputs_device(c, d_Calpha, naux * nvirt * nocc, "Calpha");
  check_inputs_device(c, d_eps_o, nocc, "eps_o");
  check_inputs_device(c, d_eps_v, nvirt, "eps_v");
  denom_check_device(c, d_eps_o, nocc, d_eps_v, nvirt);
  load_fitted_device(c, naux, nvirt, nocc, d_Calpha, mode, hipMemcpyDeviceToDevice);
  pair_energies_device(c, c->geom, c->packed.ptr, c->scales.as<double>(), nocc, nvirt, d_eps_o, d_eps_v, d_eps_pair,
                       d_e_tot, nullptr, 0);
  c->status_unread = true;
  RIMP2_API_END(ctx)
}

int rimp2_pair_energies(void* ctx, int64_t nocc, int64_t nvirt, const double* eps_o, const double* eps_v, int mode,
                        double* eps_pair, double* e_tot) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  check_pending(c);
  validate_mode(mode);
  if (!eps_o || !eps_v || !e_tot) fail(RIMP2_ERR_ARG, "null eps_o / eps_v / e_tot");
  if (!c->have_B) fail(RIMP2_ERR_STATE, "rimp2_transform or rimp2_load_fitted must run before rimp2_pair_energies");
  if (c->nocc != nocc || c->nvirt != nvirt || c->mode != mode)
    fail(RIMP2_ERR_STATE, "pair energies requested for (nocc=%lld, nvirt=%lld, mode=%d) but the transform holds "
         "(%lld, %lld, %d)", (long long)nocc, (long long)nvirt, mode, (long long)c->nocc, (long long)c->nvirt, c->mode);
  check_host_finite(eps_o, nocc, "eps_o");
  check_host_finite(eps_v, nvirt, "eps_v");
  check_host_denominators(eps_o, nocc, eps_v, nvirt);
  double* deo = static_cast<double*>(c->eps_o.ensure((size_t)(nocc + nvirt) * sizeof(double)));
  double* dev = deo + nocc;
  RIMP2_HIP_CHECK(hipMemcpyAsync(deo, eps_o, (size_t)nocc * sizeof(double), hipMemcpyHostToDevice, c->stream));
  RIMP2_HIP_CHECK(hipMemcpyAsync(dev, eps_v, (size_t)nvirt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  double* dpe = eps_pair ? static_cast<double*>(c->eps_pair.ensure((size_t)(nocc * nocc) * sizeof(double))) : nullptr;
  double* det = static_cast<double*>(c->etot.ensure(sizeof(double)));
  pair_energies_device(c, c->geom, c->packed.ptr, c->scales.as<double>(), nocc, nvirt, deo, dev, dpe, det, nullptr, 0);
  if (eps_pair)
    RIMP2_HIP_CHECK(hipMemcpyAsync(eps_pair, dpe, (size_t)(nocc * nocc) * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
  RIMP2_HIP_CHECK(hipMemcpyAsync(e_tot, det, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

// rimp2_metric + rimp2_transform + rimp2_pair_energies in one queue: the
// (P|ia) copy runs on the side stream under the Cholesky, each chunk's
// transform under the next chunk's copy, and every input check is read once,
// at the end, before any result is copied out.
int rimp2_energy(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Pia, const double* V,
                 const double* eps_o, const double* eps_v, int mode, double* eps_pair, double* e_tot) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  validate_dims(naux, nvirt, nocc);
  validate_mode(mode);
  if (!Pia || !V || !eps_o || !eps_v || !e_tot) fail(RIMP2_ERR_ARG, "null Pia / V / eps_o / eps_v / e_tot");
  check_slice_bound(naux, mode);
  require_device_memory(c, naux, nvirt, nocc, mode, true);
  check_host_finite(eps_o, nocc, "eps_o");
  check_host_finite(eps_v, nvirt, "eps_v");
  check_host_denominators(eps_o, nocc, eps_v, nvirt);
  c->have_metric = false;
  c->have_B = false;
  reset_status(c);
  // the (P|ia) chunks wait for the status reset only, not for the metric
  side_after_stream(c);
  double* dV = static_cast<double*>(c->V.ensure((size_t)(naux * naux) * sizeof(double)));
  h2d_copy(c, c->stream, dV, V, (size_t)(naux * naux) * sizeof(double));
  check_inputs_device(c, dV, naux * naux, "V");
  metric_factor(c, naux, dV);
  transform_from_host(c, nvirt, nocc, Pia, mode);
  double* deo = static_cast<double*>(c->eps_o.ensure((size_t)(nocc + nvirt) * sizeof(double)));
  double* dev = deo + nocc;
  RIMP2_HIP_CHECK(hipMemcpyAsync(deo, eps_o, (size_t)nocc * sizeof(double), hipMemcpyHostToDevice, c->stream));
  RIMP2_HIP_CHECK(hipMemcpyAsync(dev, eps_v, (size_t)nvirt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  pack_current_B(c, nvirt, nocc, mode);
  c->have_B = false;  // until the input checks are known to have passed
  double* dpe = eps_pair ? static_cast<double*>(c->eps_pair.ensure((size_t)(nocc * nocc) * sizeof(double))) : nullptr;
  double* det = static_cast<double*>(c->etot.ensure(sizeof(double)));
  pair_energies_device(c, c->geom, c->packed.ptr, c->scales.as<double>(), nocc, nvirt, deo, dev, dpe, det, nullptr, 0);
  check_status(c);
  c->have_metric = true;
  c->have_B = true;
  if (eps_pair)
    RIMP2_HIP_CHECK(hipMemcpyAsync(eps_pair, dpe, (size_t)(nocc * nocc) * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
  RIMP2_HIP_CHECK(hipMemcpyAsync(e_tot, det, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_energy_device(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* d_Pia, const double* d_V,
                        const double* d_eps_o, const double* d_eps_v, int mode, double* d_eps_pair,
                        double* d_e_tot) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  validate_dims(naux, nvirt, nocc);
  validate_mode(mode);
  if (!d_Pia || !d_V || !d_eps_o || !d_eps_v || !d_e_tot) fail(RIMP2_ERR_ARG, "null device pointer");
  check_slice_bound(naux, mode);
  require_device_memory(c, naux, nvirt, nocc, mode, false);
  reset_status_async(c);
  // (P|ia) depends on nothing computed here: its check and its packing into
  // transform order run on a side stream, under the Cholesky
  int* status = static_cast<int*>(c->status.ensure(64));
  RIMP2_HIP_CHECK(hipEventRecord(c->ev_pack_go, c->stream));
  RIMP2_HIP_CHECK(hipStreamWaitEvent(c->side_pack, c->ev_pack_go, 0));
  const bool i8 = mode_slices(mode) != 0;
  check_finite_on(c->side_pack, status, d_Pia, naux * nvirt * nocc, 512);
  if (i8) {
    metric_reserve_x_i8(c, naux, nvirt, nocc);
    metric_slice_x_range(c, c->side_pack, naux, nvirt, 0, nocc, d_Pia);
  } else {
    metric_pack_x(ctx_metric_view(c), c->side_pack, naux, nvirt, nocc, d_Pia);
  }
  RIMP2_HIP_CHECK(hipEventRecord(c->ev_pack_done, c->side_pack));
  check_inputs_device(c, d_V, naux * naux, "V");
  check_inputs_device(c, d_eps_o, nocc, "eps_o");
  check_inputs_device(c, d_eps_v, nvirt, "eps_v");
  denom_check_device(c, d_eps_o, nocc, d_eps_v, nvirt);
  metric_factor(c, naux, d_V);
  c->have_metric = true;
  if (i8) metric_slice_linv(c, c->stream, naux);
  RIMP2_HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_pack_done, 0));
  {
    double* B = static_cast<double*>(c->B64.ensure((size_t)(naux * nvirt * nocc) * sizeof(double)));
    if (i8) metric_transform_i8_range(c, naux, nvirt, 0, nocc, B);
    else metric_transform_packed(ctx_metric_view(c), naux, nvirt, nocc, B);
  }
  pack_current_B(c, nvirt, nocc, mode);
  pair_energies_device(c, c->geom, c->packed.ptr, c->scales.as<double>(), nocc, nvirt, d_eps_o, d_eps_v, d_eps_pair,
                       d_e_tot, nullptr, 0);
  c->status_unread = true;
  RIMP2_API_END(ctx)
}

int rimp2_get_cholesky(void* ctx, double* L) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  check_pending(c);
  if (!c->have_metric) fail(RIMP2_ERR_STATE, "no metric factored");
  if (!L) fail(RIMP2_ERR_ARG, "null L");
  RIMP2_HIP_CHECK(hipMemcpyAsync(L, c->V.ptr, (size_t)(c->naux * c->naux) * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_get_metric_inverse(void* ctx, double* Linv) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  check_pending(c);
  if (!c->have_metric) fail(RIMP2_ERR_STATE, "no metric factored");
  if (!Linv) fail(RIMP2_ERR_ARG, "null Linv");
  RIMP2_HIP_CHECK(hipMemcpyAsync(Linv, c->Linv.ptr, (size_t)(c->naux * c->naux) * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_get_B(void* ctx, double* B) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  check_pending(c);
  if (!c->have_B) fail(RIMP2_ERR_STATE, "no transform done");
  if (!B) fail(RIMP2_ERR_ARG, "null B");
  RIMP2_HIP_CHECK(hipMemcpyAsync(B, c->B64.ptr, (size_t)(c->naux * c->nvirt * c->nocc) * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_device_alloc(void* ctx, size_t bytes, void** dptr) {
  RIMP2_API_BEGIN
  as_ctx(ctx);
  if (!dptr) fail(RIMP2_ERR_ARG, "null dptr");
  *dptr = nullptr;
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    fail(RIMP2_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  RIMP2_API_END(ctx)
}

int rimp2_device_free(void* ctx, void* dptr) {
  RIMP2_API_BEGIN
  as_ctx(ctx);
  if (dptr) RIMP2_HIP_CHECK(hipFree(dptr));
  RIMP2_API_END(ctx)
}

int rimp2_copy_h2d(void* ctx, void* dst, const void* src, size_t bytes) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if ((!dst || !src) && bytes) fail(RIMP2_ERR_ARG, "null pointer");
  h2d_copy(c, c->stream, dst, src, bytes);
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_copy_d2h(void* ctx, void* dst, const void* src, size_t bytes) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if ((!dst || !src) && bytes) fail(RIMP2_ERR_ARG, "null pointer");
  RIMP2_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  RIMP2_API_END(ctx)
}

int rimp2_sync(void* ctx) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  RIMP2_HIP_CHECK(hipStreamSynchronize(c->stream));
  check_status(c);
  RIMP2_API_END(ctx)
}

int rimp2_timing_enable(void* ctx, int on) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  c->timing = (on != 0);
  RIMP2_API_END(ctx)
}

int rimp2_timing_get(void* ctx, int kid, double* total_ms, int64_t* count) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (kid < 0 || kid >= RIMP2_NUM_KERNELS) fail(RIMP2_ERR_ARG, "kernel id %d out of range", kid);
  resolve_timers(c);
  if (total_ms) *total_ms = c->timers[kid].total_ms;
  if (count) *count = c->timers[kid].count;
  RIMP2_API_END(ctx)
}

int rimp2_timing_reset(void* ctx) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  resolve_timers(c);
  for (auto& t : c->timers) t = Timer{};
  RIMP2_API_END(ctx)
}

}  // extern "C"
