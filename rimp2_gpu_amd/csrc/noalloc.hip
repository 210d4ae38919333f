// Allocation-free variant of rimp2_energy_fitted_device for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// Every buffer beyond the caller's inputs and outputs (packed operands,
// exact-slice row scales, padded virtual energies, per-unit partials, the
// status word) lives in one static device array of RIMP2_NOALLOC_SCRATCH_BYTES
// per GPU, held until the process ends.  Its device address is looked up once
// per device at context creation (hipGetSymbolAddress with that device
// current); host code uses only that address, never the array's name, which
// in host code is a stand-in and not a device address.  Every call on a GPU
// shares that GPU's array and leaves its contents behind, so calls that
// overlap on one GPU (two streams, two contexts, two threads) race.
#include <mutex>

#include "rimp2_common.h"

namespace rimp2 {

static __device__ __align__(256) unsigned char scratch[RIMP2_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The status word must start at zero: a kernel clears it (no memset node, so
// the call stays capturable).
__global__ void clear_words_kernel(int* w, int n) {
  if ((int)threadIdx.x < n) w[threadIdx.x] = 0;
}

struct Layout {
  size_t status, evp, part, sigma, packed, end;
};

Layout layout(const PackGeom& g, int tile, int64_t nocc, bool slices) {
  Layout L;
  L.status = 0;
  size_t evp_in_part = 0;
  const size_t pair = pair_scratch_bytes(g, tile, nocc * (nocc + 1) / 2, &evp_in_part);
  L.part = 256;
  L.evp = L.part + evp_in_part;
  L.sigma = L.part + pair;
  const size_t sig = slices ? (size_t)round_up(g.nvp * nocc * (int64_t)sizeof(double), 256) : 0;
  L.packed = L.sigma + sig;
  L.end = L.packed + g.bytes_per_occ() * (size_t)nocc;
  return L;
}

}  // namespace

unsigned char* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) fail(RIMP2_ERR_ARG, "device %d beyond the scratch table", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fail(RIMP2_ERR_HIP, "hipGetSymbolAddress of the scratch array on device %d failed: %s", device,
           hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace rimp2

using namespace rimp2;

extern "C" {

int rimp2_energy_fitted_device_noalloc(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* d_Calpha,
                                       const double* d_eps_o, const double* d_eps_v, int mode, double* d_eps_pair,
                                       double* d_e_tot) {
  RIMP2_API_BEGIN
  Ctx* c = as_ctx(ctx);
  if (naux <= 0 || nvirt <= 0 || nocc <= 0 || naux > 65536 || nvirt > 65536 || nocc > 65536)
    fail(RIMP2_ERR_ARG, "dimensions out of range (naux=%lld nvirt=%lld nocc=%lld)", (long long)naux,
         (long long)nvirt, (long long)nocc);
  if (mode < 0 || mode >= RIMP2_NUM_MODES) fail(RIMP2_ERR_ARG, "unknown precision mode %d", mode);
  if (!d_Calpha || !d_eps_o || !d_eps_v || !d_e_tot) fail(RIMP2_ERR_ARG, "null device pointer");
  const bool slices = mode_slices(mode) != 0;
  if (slices && naux > 16384) fail(RIMP2_ERR_ARG, "exact-slice modes support naux <= 16384");
  if (!c->scratch) fail(RIMP2_ERR_STATE, "context has no scratch array");
  const int tile = choose_tile(mode, nvirt, nocc);
  const PackGeom g = make_geom(mode, naux, nvirt, tile);
  const Layout L = layout(g, tile, nocc, slices);
  if (L.end > (size_t)RIMP2_NOALLOC_SCRATCH_BYTES)
    fail(RIMP2_ERR_NOMEM, "naux=%lld nvirt=%lld nocc=%lld in mode %d needs %zu bytes of scratch; the "
         "allocation-free entry point holds %zu (RIMP2_NOALLOC_SCRATCH_BYTES)", (long long)naux, (long long)nvirt,
         (long long)nocc, mode, L.end, (size_t)RIMP2_NOALLOC_SCRATCH_BYTES);
  unsigned char* S = c->scratch;
  int* status = reinterpret_cast<int*>(S + L.status);
  hipStream_t s = c->stream;
  // unread bits of an earlier call of this entry point are kept (rimp2_sync
  // reports them); otherwise the word starts at zero
  if (!(c->status_unread && c->status_scratch == status)) {
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, s, status, 16);
    RIMP2_HIP_CHECK(hipGetLastError());
  }
  check_finite_on(s, status, d_Calpha, naux * nvirt * nocc);
  check_finite_on(s, status, d_eps_o, nocc);
  check_finite_on(s, status, d_eps_v, nvirt);
  check_denom_on(s, status, d_eps_o, nocc, d_eps_v, nvirt);
  double* sigma = slices ? reinterpret_cast<double*>(S + L.sigma) : nullptr;
  pack_B(nullptr, d_Calpha, naux, nvirt, nocc, g, S + L.packed, sigma, s);
  pair_energies_core(nullptr, s, tile, g, S + L.packed, sigma, nocc, nvirt, d_eps_o, d_eps_v, d_eps_pair, d_e_tot,
                     nullptr, 0, reinterpret_cast<double*>(S + L.evp), reinterpret_cast<double*>(S + L.part));
  c->status_unread = true;
  c->status_scratch = status;
  RIMP2_API_END(ctx)
}

}  // extern "C"
