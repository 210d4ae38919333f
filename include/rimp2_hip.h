/*
 * rimp2_hip.h -- C ABI of the MI355X RI-MP2 energy library (librimp2_hip.so).
 *
 * This is the drop-in boundary that replaces the reference's Fortran FFI module
 * gpu_interfaces.F90 (cuBLAS create/destroy/Dgemm/SetStream + get_dev_mem,
 * gpu_interfaces.F90:8-55) and its native helper cuda_api.cu:4-9.  The Fortran
 * binding is rimp2_gpu_amd/fortran/rimp2_hip_interfaces.F90; the Python host
 * mirror is rimp2_gpu_amd/__init__.py (ctypes).
 *
 * Conventions (all entry points):
 *   - every dimension is int64_t, every array is column-major (Fortran order);
 *     (P|ia) and B are [naux, nvirt, nocc] with the auxiliary index fastest,
 *     as Calpha(NBA,nvirt,nocc) in RIMP2.F90:48;
 *   - every function returns int status, 0 = RIMP2_OK;  the text of the last
 *     failure is available from rimp2_last_error;  no C++ exception crosses the
 *     ABI;
 *   - host-pointer entry points block until results are in host memory;
 *     *_device entry points take device pointers and are asynchronous on the
 *     context's stream (rimp2_sync waits);
 *   - the library never frees caller memory;  one host thread per context.
 */
#ifndef RIMP2_HIP_H
#define RIMP2_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define RIMP2_OK 0
#define RIMP2_ERR_ARG 1        /* bad dimension, null pointer, unknown mode */
#define RIMP2_ERR_HIP 2        /* HIP runtime failure (message has the HIP error) */
#define RIMP2_ERR_NOT_SPD 3    /* metric not positive definite (Cholesky pivot <= 0) */
#define RIMP2_ERR_STATE 4      /* call order: e.g. pair energies before a transform */
#define RIMP2_ERR_NOMEM 5      /* device allocation failed / request exceeds free memory */
#define RIMP2_ERR_NUMERIC 6    /* NaN/Inf input, or an orbital-energy denominator >= 0 */

/* ---- precision modes of the pair contraction ---- */
#define RIMP2_MODE_FP64 0      /* FP64 MFMA (v_mfma_f64_16x16x4), the plain path */
#define RIMP2_MODE_FP32 1      /* FP32 MFMA (v_mfma_f32_32x32x2), FP32 accumulation */
#define RIMP2_MODE_BF16X3 2    /* split BF16 hi+lo, 3 products, FP32 accumulation */
#define RIMP2_MODE_I8S4 3      /* exact-slice: 4 int8 slices, 13 products, int32 acc */
#define RIMP2_MODE_I8S5 4      /* exact-slice: 5 int8 slices, 15 products, int32 acc */
#define RIMP2_NUM_MODES 5

/* ---- kernel ids for the built-in event timers ---- */
#define RIMP2_K_PAIR 0
#define RIMP2_K_TRANSFORM 1
#define RIMP2_K_PACK 2
#define RIMP2_K_CHOLESKY 3
#define RIMP2_K_TRTRI 4
#define RIMP2_K_REDUCE 5
#define RIMP2_K_BATCH 6
#define RIMP2_K_MLP 7
#define RIMP2_K_FEATURES 8
#define RIMP2_K_EOS 9        /* a whole rimp2_eos_arrays or rimp2_first_order call */
#define RIMP2_K_EOS_AMPL 10  /* its amplitude GEMMs (1/Delta epilogue) */
#define RIMP2_NUM_KERNELS 11

/* ---- context, device memory query, errors ---- */

/* Replaces cublasCreate_v2 (gpu_interfaces.F90:11-15, called RIMP2.GPU.F90:81):
 * creates a context on `device` owning a HIP stream, device buffers, graph
 * cache and tile-shape table. */
int rimp2_ctx_create(int device, void** ctx);
/* Replaces cublasDestroy_v2 (gpu_interfaces.F90:18-22, called RIMP2.GPU.F90:347). */
int rimp2_ctx_destroy(void* ctx);
/* Replaces get_dev_mem (cuda_api.cu:4-9, gpu_interfaces.F90:47-53); same
 * argument order (total, free), built on hipMemGetInfo for the current device. */
int rimp2_get_dev_mem(size_t* total, size_t* free_bytes);
/* The reference's own symbol, same name and ABI (cuda_api.cu:4-9 declares
 * `void get_dev_mem(size_t& total, size_t& free)` under extern "C"; bound by
 * gpu_interfaces.F90:47-53), so that interface links against this library
 * unchanged.  Both outputs are 0 when the query fails. */
void get_dev_mem(size_t* total, size_t* free_bytes);
/* Copies the last error text of ctx (or the library-global one when ctx is
 * NULL) into buf (NUL-terminated, at most n bytes). */
int rimp2_last_error(void* ctx, char* buf, size_t n);

/* ---- the RI-MP2 energy path (host pointers, blocking) ---- */

/* FP64 Cholesky V = L L^T and inverse L^{-1} of the SPD metric V[naux,naux] on
 * the GPU.  The reference receives its 3-index tensor already fitted
 * (Calpha, RIMP2.F90:48-49); this is the fitting step in front of it. */
int rimp2_metric(void* ctx, int64_t naux, const double* V);
/* B = L^{-1} (P|ia) as an FP64 MFMA GEMM, then packed for `mode`.  Requires
 * rimp2_metric with the same naux.  Pia is [naux, nvirt, nocc]. */
int rimp2_transform(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Pia, int mode);
/* Fused pair kernel on the transformed B of the last rimp2_transform:
 *   eps_ij = sum_ab V_ab (2 V_ab - V_ba) / (e_i + e_j - e_a - e_b),
 *   V_ab = sum_Q B^Q_ia B^Q_jb,   E = sum_ij eps_ij.
 * This is RIMP2_calc_toccA (RIMP2.F90:381-417, t = g/Delta) combined with
 * RIMP2_calc_gocc (RIMP2.F90:562-587, g) at U = identity, contracted to
 * energies.  eps_pair is [nocc,nocc] (may be NULL); e_tot receives E. */
int rimp2_pair_energies(void* ctx, int64_t nocc, int64_t nvirt, const double* eps_o, const double* eps_v, int mode,
                        double* eps_pair, double* e_tot);
/* Whole path from host inputs to host energies: the three calls above as one
 * queue ((P|ia) is copied by chunks of occupied orbitals under the Cholesky).
 * A NaN/Inf input (RIMP2_ERR_NUMERIC) or a non-positive pivot
 * (RIMP2_ERR_NOT_SPD) is reported before eps_pair / e_tot are written. */
int rimp2_energy(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Pia, const double* V,
                 const double* eps_o, const double* eps_v, int mode, double* eps_pair, double* e_tot);
/* Same path with every pointer in device memory; asynchronous on the context
 * stream.  d_eps_pair [nocc,nocc] may be NULL; d_e_tot is one double. */
int rimp2_energy_device(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* d_Pia, const double* d_V,
                        const double* d_eps_o, const double* d_eps_v, int mode, double* d_eps_pair, double* d_e_tot);

/* ---- the same path from already-fitted coefficients (no metric) ---- */

/* Loads the fitted 3-index tensor Calpha(NBA,nvirt,nocc) exactly as the
 * reference reads it (RIMP2.F90:47-49, read(fileLu) Calpha) and packs it for
 * `mode`; B = Calpha, no metric step.  rimp2_pair_energies then runs on it. */
int rimp2_load_fitted(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Calpha, int mode);
/* rimp2_load_fitted + rimp2_pair_energies: the reference's per-fragment calls
 * of RIMP2_calc_toccA and RIMP2_calc_gocc (RIMP2.F90:73, 146; GPU build
 * RIMP2.GPU.F90:109, 228) reduced to the canonical energy, from host arrays. */
int rimp2_energy_fitted(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* Calpha,
                        const double* eps_o, const double* eps_v, int mode, double* eps_pair, double* e_tot);
/* Device-pointer, asynchronous variant of rimp2_energy_fitted. */
int rimp2_energy_fitted_device(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* d_Calpha,
                               const double* d_eps_o, const double* d_eps_v, int mode, double* d_eps_pair,
                               double* d_e_tot);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rimp2_energy_fitted_device with every buffer beyond the
 * inputs and outputs (packed operands, row scales, partials, status word) in
 * one static device array of RIMP2_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; the call allocates nothing.  A fragment whose buffers do
 * not fit is refused with RIMP2_ERR_NOMEM before anything is queued (the
 * 16/128/512 fragment needs ~9 MiB in fp64, a 32/256/1024 one ~67 MiB).
 * Asynchronous on the context stream; input checks are reported by
 * rimp2_sync.  All calls on one GPU share its array and the last call's
 * contents stay there: calls that overlap on one GPU (other streams, contexts
 * or threads) race. */
#define RIMP2_NOALLOC_SCRATCH_BYTES (128ull << 20)
int rimp2_energy_fitted_device_noalloc(void* ctx, int64_t naux, int64_t nvirt, int64_t nocc, const double* d_Calpha,
                                       const double* d_eps_o, const double* d_eps_v, int mode, double* d_eps_pair,
                                       double* d_e_tot);

/* ---- inspection of intermediate results (host pointers, blocking) ---- */
int rimp2_get_cholesky(void* ctx, double* L);      /* L [naux,naux], lower, zeros above */
int rimp2_get_metric_inverse(void* ctx, double* Linv); /* L^{-1} [naux,naux] */
int rimp2_get_B(void* ctx, double* B);             /* FP64 B [naux,nvirt,nocc] */

/* ---- batched DEC fragments ---- */
typedef struct rimp2_frag_desc {
  int64_t nocc, nvirt, naux;
  const double* Pia;   /* [naux, nvirt, nocc] */
  const double* V;     /* [naux, naux]; NULL: Pia is already fitted (B = Pia) */
  const double* eps_o; /* [nocc] */
  const double* eps_v; /* [nvirt] */
} rimp2_frag_desc;
/* Many fragments of uneven shape through one persistent, graph-captured pair
 * launch (the analogue of the reference's serial fragment loop
 * `do ifile=1,2`, RIMP2.F90:36).  e_frag[f] receives E of fragment f;
 * eps_pair_packed (may be NULL) receives each fragment's [nocc,nocc] pair
 * matrix, concatenated in fragment order. */
int rimp2_batch(void* ctx, int64_t nfrag, const rimp2_frag_desc* frags, int mode, double* e_frag,
                double* eps_pair_packed);
/* Device-resident variant: metric, transform and pack for every fragment are
 * done beforehand by rimp2_batch_prepare (host inputs); rimp2_batch_run then
 * replays the captured graph (pair kernel + reductions) and leaves energies in
 * device memory (d_e_frag [nfrag]). */
int rimp2_batch_prepare(void* ctx, int64_t nfrag, const rimp2_frag_desc* frags, int mode);
/* As rimp2_batch_prepare, with every array of the descriptors in device memory
 * (inputs already in HBM: no host copies).  Input checks (non-finite values,
 * denominators, positive definiteness) run on the device and are reported per
 * fragment when the call returns.  Serves the reference's fragment loop
 * (`do ifile=1,2`, RIMP2.GPU.F90:52) when the caller already holds each
 * fragment's arrays on the device, where the reference reads them from file
 * into host memory per fragment (RIMP2.GPU.F90:83-88). */
int rimp2_batch_prepare_device(void* ctx, int64_t nfrag, const rimp2_frag_desc* frags, int mode);
int rimp2_batch_run(void* ctx, double* d_e_frag);
int rimp2_batch_result(void* ctx, double* e_frag, double* eps_pair_packed);
/* Tile candidate of the prepared batch: 32 or 64 (every fragment with that
 * pair tile, i.e. virtual block width) or -1 (mixed: per fragment the 32-wide
 * tile where the 64-wide one's padding executes at least 1.5x the MFMA work,
 * i.e. ceil64(nvirt)^2 >= 1.5 ceil32(nvirt)^2, else the 64-wide; one launch
 * carries both).  fp64 and fp32 have all three candidates, bf16x3 and
 * the exact-slice modes 64 only. */
int rimp2_batch_tile(void* ctx, int64_t* tile);
/* The autotuner's measurements for the prepared batch: *n (<= 4) candidates
 * (encoded as rimp2_batch_tile) and the time of one pair-stage launch with
 * each (ms).  *n = 0 when the choice came from the per-(mode, size class)
 * cache or from rimp2_batch_set_tile. */
int rimp2_batch_tuning(void* ctx, int* n, int64_t* tiles, double* ms);
/* Fixes the tile candidate of the following rimp2_batch_prepare(_device)
 * calls on ctx (32, 64 or -1 as rimp2_batch_tile; 0 returns to the
 * autotuner).  A candidate the mode does not have makes the prepare fail with
 * RIMP2_ERR_ARG.  No counterpart in the reference (its fragment loop,
 * RIMP2.F90:36, has no tiles); for callers that need a reproducible layout. */
int rimp2_batch_set_tile(void* ctx, int64_t tile);

/* ---- pair-screening surrogate (BF16 MLP 16 -> 64 -> 64 -> 1, GELU) ---- */
#define RIMP2_NFEAT 16
/* Per-pair features of the last transform: [npairs_unique, 16] float32,
 * pairs ordered j-major (i <= j, index j*(j+1)/2 + i).  Feature 15 is the
 * base estimate log(sum_Q ||B_Qi||^2 ||B_Qj||^2 / (2 mean(e_a) - e_i - e_j));
 * the others are dimensionless ratios of row-norm sums and orbital-energy
 * gaps (mlp.hip lists them). */
int rimp2_pair_features(void* ctx, int64_t nocc, int64_t nvirt, const double* eps_o, const double* eps_v,
                        float* feats);
int rimp2_mlp_init(void* ctx, uint64_t seed);
/* Standardisation used by every MLP call: feature f enters as
 * (x_f - mean[f]) / stdev[f]; the network output y gives the prediction
 * log(-eps_ij) = feature 15 + y tstd + tmean (a residual on the base
 * estimate), so tmean / tstd are the statistics of log(-eps_ij) - feature 15. */
int rimp2_mlp_set_norm(void* ctx, const float* mean, const float* stdev, float tmean, float tstd);
/* One Adam step on a minibatch (host arrays: raw features [n,16], raw targets
 * log(-eps_ij)); loss receives the mean squared error of the standardised
 * target before the step. */
int rimp2_mlp_train_step(void* ctx, int64_t n, const float* feats, const float* target, float lr, float* loss);
/* pred[k] = predicted log(-eps) of sample k (raw units). */
int rimp2_mlp_predict(void* ctx, int64_t n, const float* feats, float* pred);
int rimp2_mlp_get_params(void* ctx, float* params, int64_t n);
int rimp2_mlp_set_params(void* ctx, const float* params, int64_t n);
int64_t rimp2_mlp_num_params(void);
/* Single forward+backward of the MLP on one minibatch: returns the loss and the
 * gradient of every parameter (float32, same order as the parameter vector). */
int rimp2_mlp_grad(void* ctx, int64_t n, const float* feats, const float* target, float* loss, float* grad);
/* Screened energy: pairs whose predicted |eps_ij| < thresh are skipped, the
 * rest run through the exact pair kernel.  nskipped receives the count. */
int rimp2_screened_energy(void* ctx, int64_t nocc, int64_t nvirt, const double* eps_o, const double* eps_v, int mode,
                          double thresh, double* eps_pair, double* e_tot, int64_t* nskipped);

/* ---- device memory and stream helpers (for device-resident callers) ---- */
int rimp2_device_alloc(void* ctx, size_t bytes, void** dptr);
int rimp2_device_free(void* ctx, void* dptr);
int rimp2_copy_h2d(void* ctx, void* dst, const void* src, size_t bytes);
int rimp2_copy_d2h(void* ctx, void* dst, const void* src, size_t bytes);
int rimp2_sync(void* ctx);

/* ---- event timers on the context stream ---- */
int rimp2_timing_enable(void* ctx, int on);
/* Total milliseconds and launch count of kernel id `kid` since the last reset. */
int rimp2_timing_get(void* ctx, int kid, double* total_ms, int64_t* count);
int rimp2_timing_reset(void* ctx);

/* ---- seeded synthetic inputs (host) ---- */
/* (P|ia) ~ N(0,1)*0.1/sqrt(1+a/16), V = A A^T/naux + I with A ~ N(0,1),
 * eps_o ~ U[-2,-0.3], eps_v ~ U[0.1,3.0]; any output pointer may be NULL. */
int rimp2_gen_fragment(uint64_t seed, int64_t naux, int64_t nvirt, int64_t nocc, double* Pia, double* V,
                       double* eps_o, double* eps_v);
/* Locality-decay variant (lam > 0): orbitals on a 1-D chain, (P|ia) scaled by
 * exp(-(|x_Q-x_i| + |x_a-x_i|)/lam), A[p,q] by exp(-|x_p-x_q|/lam), so pair
 * energies decay with |i-j|.  lam <= 0 gives rimp2_gen_fragment exactly. */
int rimp2_gen_fragment_local(uint64_t seed, int64_t naux, int64_t nvirt, int64_t nocc, double lam, double* Pia,
                             double* V, double* eps_o, double* eps_v);
/* The same generator evaluated on the GPU into device buffers (any may be
 * NULL), asynchronous on the context stream; agrees with the host generator to
 * a few ulps (device math library). */
int rimp2_gen_fragment_device(void* ctx, uint64_t seed, int64_t naux, int64_t nvirt, int64_t nocc, double lam,
                              double* d_Pia, double* d_V, double* d_eps_o, double* d_eps_v);
/* nocc ~ U{4..32}, nvirt ~ U{32..256}, naux = 4 nvirt per fragment. */
int rimp2_gen_batch_shapes(uint64_t seed, int64_t nfrag, int64_t* nocc, int64_t* nvirt, int64_t* naux);

/* ---- general FP64 GEMM and the DEC local-basis 4-index arrays ---- */
/* Replaces cublasDgemm_v2 as bound in gpu_interfaces.F90:24-32 and called at
 * RIMP2.GPU.F90:119,135,172,186,213,244: C = alpha op(A) op(B) + beta C,
 * column-major, device pointers, asynchronous on the context stream (the
 * binding's handle + cublasSetStream pair becomes the context).  transa /
 * transb: 0 = N, 1 = T (2 = C, the same for real data; 'N'/'T'/'C' also
 * accepted).  alpha and beta are host pointers (cuBLAS host pointer mode);
 * beta = 0 does not read C, alpha = 0 does not read A or B.  Leading
 * dimensions are checked as BLAS does. */
int rimp2_dgemm(void* ctx, int transa, int transb, int64_t m, int64_t n, int64_t k, const double* alpha,
                const double* d_A, int64_t lda, const double* d_B, int64_t ldb, const double* beta, double* d_C,
                int64_t ldc);
/* Replaces Major Steps 5-8 of the GPU host (RIMP2.GPU.F90:103-260; CPU form
 * RIMP2.F90:67-170): from Calpha(NBA,nvirt,nocc), EVocc(nocc), EVvirt(nvirt),
 * UoccEOST(nocc,noccEOS), UvirtT(nvirt,nvirt), UoccT(nocc,nocc) and
 * UvirtEOST(nvirt,nvirtEOS) forms toccEOS(nvirt,noccEOS,nvirt,noccEOS),
 * tvirtEOS(nvirtEOS,nocc,nvirtEOS,nocc), goccEOS(nvirt,noccEOS,nvirt,noccEOS)
 * and gvirtEOS(nvirtEOS,nocc,nvirtEOS,nocc), all FP64 on MFMA.  As in the
 * reference the denominators e_I + e_J - e_A - e_B are not sign-checked (its
 * own fixture has them > 0).  Blocking host-pointer form; the _device form
 * takes device pointers and is asynchronous on the context stream. */
int rimp2_eos_arrays(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba,
                     const double* Calpha, const double* EVocc, const double* EVvirt, const double* UoccEOST,
                     const double* UvirtT, const double* UoccT, const double* UvirtEOST, double* toccEOS,
                     double* tvirtEOS, double* goccEOS, double* gvirtEOS);
int rimp2_eos_arrays_device(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba,
                            const double* d_Calpha, const double* d_EVocc, const double* d_EVvirt,
                            const double* d_UoccEOST, const double* d_UvirtT, const double* d_UoccT,
                            const double* d_UvirtEOST, double* d_toccEOS, double* d_tvirtEOS, double* d_goccEOS,
                            double* d_gvirtEOS);
/* Replaces the first-order block of the reference host (RIMP2.F90:172-251;
 * the GPU build skips it, RIMP2.GPU.F90:100): from Calpha(NBA,nvirt,nocc),
 * CalphaOcc(NBA,nocc,nocctot), CalphaVV(NBA,nvirt,nvirt) and the four U
 * matrices forms djik(nvirt,noccEOS,noccEOS,nocctot) and
 * blad(nvirtEOS,nocc,nvirtEOS,nvirt) (RIMP2_calc_gen4DimFO, RIMP2.F90:616-641)
 * in FP64 on MFMA, with the reference's own dgemm arguments: the CalphaVV
 * product has K = ldb = nocc (RIMP2.F90:231-235), so it sums the first nocc
 * virtual columns of CalphaVV against UvirtT's storage read as an nocc x nvirt
 * matrix.  The reference is in bounds only for nocctot = nocc and nocc <=
 * nvirt; other sizes return RIMP2_ERR_ARG.  Blocking host form; _device is
 * asynchronous on the context stream. */
int rimp2_first_order(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS, int64_t nba,
                      int64_t nocctot, const double* Calpha, const double* CalphaOcc, const double* CalphaVV,
                      const double* UoccEOST, const double* UvirtT, const double* UoccT, const double* UvirtEOST,
                      double* djik, double* blad);
int rimp2_first_order_device(void* ctx, int64_t nvirt, int64_t nocc, int64_t noccEOS, int64_t nvirtEOS,
                              int64_t nba, int64_t nocctot, const double* d_Calpha, const double* d_CalphaOcc,
                              const double* d_CalphaVV, const double* d_UoccEOST, const double* d_UvirtT,
                              const double* d_UoccT, const double* d_UvirtEOST, double* d_djik, double* d_blad);
/* Occupied orbitals per amplitude chunk of rimp2_eos_arrays.  0 (default):
 * automatic, the chunk's amplitude slab (nvirt^2 nocc doubles per orbital)
 * and its tvirtEOS partial within 8 GiB, and the toccEOS partial of every
 * orbital (nvirt^2 noccEOS nocc doubles) kept whole when it fits 32 GiB.  An
 * explicit chunk bounds the work memory: the toccEOS partial then holds one
 * chunk.  Has no counterpart in the reference, which holds every amplitude at
 * once. */
int rimp2_eos_set_chunk(void* ctx, int64_t nocc_per_chunk);
/* Precision of rimp2_eos_arrays' amplitude slab V^{IJ}_AB (the rest is FP64
 * either way): RIMP2_MODE_FP64 (default; the reference's arithmetic) or
 * RIMP2_MODE_I8S5 (the exact-slice int8 pair kernel, ~2x faster, V to ~1e-12
 * relative: within the reference's 1e-10 absolute check for arrays of O(1)
 * values, not for its constant fixtures of values ~10^2).  Other modes return
 * RIMP2_ERR_ARG.  No counterpart in the reference, which is FP64 only. */
int rimp2_eos_set_mode(void* ctx, int mode);

#ifdef __cplusplus
}
#endif
#endif /* RIMP2_HIP_H */
