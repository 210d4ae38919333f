!> @file
!> Fortran bindings of the MI355X RI-MP2 library (include/rimp2_hip.h).
!> Replaces the reference's gpu_interfaces.F90 (cublasCreate_v2,
!> cublasDestroy_v2, cublasDgemm_v2, cublasSetStream_v2 and get_dev_mem,
!> gpu_interfaces.F90:8-55).  All dimensions are C_INT64_T by value, all
!> arrays are passed by reference (column-major, as allocated by the host),
!> every function returns a C_INT status (0 = ok).
module rimp2_hip_interfaces
  use iso_c_binding
  implicit none

  integer(C_INT), parameter :: RIMP2_OK = 0
  integer(C_INT), parameter :: RIMP2_ERR_ARG = 1, RIMP2_ERR_HIP = 2, RIMP2_ERR_NOT_SPD = 3, RIMP2_ERR_STATE = 4
  integer(C_INT), parameter :: RIMP2_ERR_NOMEM = 5, RIMP2_ERR_NUMERIC = 6
  integer(C_INT), parameter :: RIMP2_MODE_FP64 = 0, RIMP2_MODE_FP32 = 1, RIMP2_MODE_BF16X3 = 2
  integer(C_INT), parameter :: RIMP2_MODE_I8S4 = 3, RIMP2_MODE_I8S5 = 4
  ! kernel ids of the event timers (rimp2_timing_get)
  integer(C_INT), parameter :: RIMP2_K_PAIR = 0, RIMP2_K_TRANSFORM = 1, RIMP2_K_PACK = 2, RIMP2_K_CHOLESKY = 3
  integer(C_INT), parameter :: RIMP2_K_TRTRI = 4, RIMP2_K_REDUCE = 5, RIMP2_K_BATCH = 6, RIMP2_K_MLP = 7
  integer(C_INT), parameter :: RIMP2_K_FEATURES = 8, RIMP2_K_EOS = 9, RIMP2_K_EOS_AMPL = 10
  integer(C_INT), parameter :: RIMP2_NFEAT = 16

  type, bind(C) :: rimp2_frag_desc
     integer(C_INT64_T) :: nocc, nvirt, naux
     type(C_PTR) :: Pia, V, eps_o, eps_v
  end type rimp2_frag_desc

  interface

     ! replaces cublasCreate_v2 (gpu_interfaces.F90:11-15)
     integer(C_INT) function rimp2_ctx_create(device, ctx) bind(C, name="rimp2_ctx_create")
       import :: C_INT, C_PTR
       integer(C_INT), value :: device
       type(C_PTR) :: ctx
     end function rimp2_ctx_create

     ! replaces cublasDestroy_v2 (gpu_interfaces.F90:18-22)
     integer(C_INT) function rimp2_ctx_destroy(ctx) bind(C, name="rimp2_ctx_destroy")
       import :: C_INT, C_PTR
       type(C_PTR), value :: ctx
     end function rimp2_ctx_destroy

     ! replaces get_dev_mem (gpu_interfaces.F90:47-53); same (total, free) order
     integer(C_INT) function rimp2_get_dev_mem(total, free) bind(C, name="rimp2_get_dev_mem")
       import :: C_INT, C_SIZE_T
       integer(C_SIZE_T) :: total, free
     end function rimp2_get_dev_mem

     integer(C_INT) function rimp2_last_error(ctx, buf, n) bind(C, name="rimp2_last_error")
       import :: C_INT, C_PTR, C_CHAR, C_SIZE_T
       type(C_PTR), value :: ctx
       character(kind=C_CHAR) :: buf(*)
       integer(C_SIZE_T), value :: n
     end function rimp2_last_error

     integer(C_INT) function rimp2_metric(ctx, naux, V) bind(C, name="rimp2_metric")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux
       real(C_DOUBLE) :: V(*)
     end function rimp2_metric

     integer(C_INT) function rimp2_transform(ctx, naux, nvirt, nocc, Pia, mode) bind(C, name="rimp2_transform")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux, nvirt, nocc
       real(C_DOUBLE) :: Pia(*)
       integer(C_INT), value :: mode
     end function rimp2_transform

     integer(C_INT) function rimp2_pair_energies(ctx, nocc, nvirt, eps_o, eps_v, mode, eps_pair, e_tot) &
          & bind(C, name="rimp2_pair_energies")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nocc, nvirt
       real(C_DOUBLE) :: eps_o(*), eps_v(*)
       integer(C_INT), value :: mode
       real(C_DOUBLE) :: eps_pair(*), e_tot
     end function rimp2_pair_energies

     integer(C_INT) function rimp2_energy(ctx, naux, nvirt, nocc, Pia, V, eps_o, eps_v, mode, eps_pair, e_tot) &
          & bind(C, name="rimp2_energy")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux, nvirt, nocc
       real(C_DOUBLE) :: Pia(*), V(*), eps_o(*), eps_v(*)
       integer(C_INT), value :: mode
       real(C_DOUBLE) :: eps_pair(*), e_tot
     end function rimp2_energy

     integer(C_INT) function rimp2_load_fitted(ctx, naux, nvirt, nocc, Calpha, mode) bind(C, name="rimp2_load_fitted")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux, nvirt, nocc
       real(C_DOUBLE) :: Calpha(*)
       integer(C_INT), value :: mode
     end function rimp2_load_fitted

     integer(C_INT) function rimp2_energy_fitted(ctx, naux, nvirt, nocc, Calpha, eps_o, eps_v, mode, eps_pair, e_tot) &
          & bind(C, name="rimp2_energy_fitted")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux, nvirt, nocc
       real(C_DOUBLE) :: Calpha(*), eps_o(*), eps_v(*)
       integer(C_INT), value :: mode
       real(C_DOUBLE) :: eps_pair(*), e_tot
     end function rimp2_energy_fitted

     integer(C_INT) function rimp2_batch(ctx, nfrag, frags, mode, e_frag, eps_pair_packed) bind(C, name="rimp2_batch")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE, rimp2_frag_desc
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nfrag
       type(rimp2_frag_desc) :: frags(*)
       integer(C_INT), value :: mode
       real(C_DOUBLE) :: e_frag(*)
       type(C_PTR), value :: eps_pair_packed
     end function rimp2_batch

     integer(C_INT) function rimp2_gen_fragment(seed, naux, nvirt, nocc, Pia, V, eps_o, eps_v) &
          & bind(C, name="rimp2_gen_fragment")
       import :: C_INT, C_INT64_T, C_DOUBLE
       integer(C_INT64_T), value :: seed, naux, nvirt, nocc
       real(C_DOUBLE) :: Pia(*), V(*), eps_o(*), eps_v(*)
     end function rimp2_gen_fragment

     !> Major Steps 5-8 (RIMP2.F90:67-170) on the GPU, host arrays, blocking.
     integer(C_INT) function rimp2_eos_arrays(ctx, nvirt, nocc, noccEOS, nvirtEOS, nba, Calpha, EVocc, EVvirt, &
          & UoccEOST, UvirtT, UoccT, UvirtEOST, toccEOS, tvirtEOS, goccEOS, gvirtEOS) bind(C, name="rimp2_eos_arrays")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nvirt, nocc, noccEOS, nvirtEOS, nba
       real(C_DOUBLE) :: Calpha(*), EVocc(*), EVvirt(*), UoccEOST(*), UvirtT(*), UoccT(*), UvirtEOST(*)
       real(C_DOUBLE) :: toccEOS(*), tvirtEOS(*), goccEOS(*), gvirtEOS(*)
     end function rimp2_eos_arrays

     !> Occupied orbitals per amplitude chunk of rimp2_eos_arrays (0: automatic).
     integer(C_INT) function rimp2_eos_set_chunk(ctx, nocc_per_chunk) bind(C, name="rimp2_eos_set_chunk")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nocc_per_chunk
     end function rimp2_eos_set_chunk

     !> Amplitude slab precision of rimp2_eos_arrays: RIMP2_MODE_FP64 (default)
     !> or RIMP2_MODE_I8S5.
     integer(C_INT) function rimp2_eos_set_mode(ctx, mode) bind(C, name="rimp2_eos_set_mode")
       import :: C_INT, C_PTR
       type(C_PTR), value :: ctx
       integer(C_INT), value :: mode
     end function rimp2_eos_set_mode

     !> First-order block (RIMP2.F90:172-251) on the GPU, host arrays, blocking.
     integer(C_INT) function rimp2_first_order(ctx, nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot, Calpha, &
          & CalphaOcc, CalphaVV, UoccEOST, UvirtT, UoccT, UvirtEOST, djik, blad) bind(C, name="rimp2_first_order")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot
       real(C_DOUBLE) :: Calpha(*), CalphaOcc(*), CalphaVV(*), UoccEOST(*), UvirtT(*), UoccT(*), UvirtEOST(*)
       real(C_DOUBLE) :: djik(*), blad(*)
     end function rimp2_first_order

     !> cublasDgemm_v2 as bound in gpu_interfaces.F90:24-32: device pointers,
     !> asynchronous on the context stream, alpha/beta by reference.
     integer(C_INT) function rimp2_dgemm(ctx, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) &
          & bind(C, name="rimp2_dgemm")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx, A, B, C
       integer(C_INT), value :: transa, transb
       integer(C_INT64_T), value :: m, n, k, lda, ldb, ldc
       real(C_DOUBLE) :: alpha, beta
     end function rimp2_dgemm

     ! ---- device-resident energy path (device pointers, asynchronous) ----

     integer(C_INT) function rimp2_energy_device(ctx, naux, nvirt, nocc, d_Pia, d_V, d_eps_o, d_eps_v, mode, &
          & d_eps_pair, d_e_tot) bind(C, name="rimp2_energy_device")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux, nvirt, nocc
       type(C_PTR), value :: d_Pia, d_V, d_eps_o, d_eps_v
       integer(C_INT), value :: mode
       type(C_PTR), value :: d_eps_pair, d_e_tot
     end function rimp2_energy_device

     integer(C_INT) function rimp2_energy_fitted_device(ctx, naux, nvirt, nocc, d_Calpha, d_eps_o, d_eps_v, mode, &
          & d_eps_pair, d_e_tot) bind(C, name="rimp2_energy_fitted_device")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux, nvirt, nocc
       type(C_PTR), value :: d_Calpha, d_eps_o, d_eps_v
       integer(C_INT), value :: mode
       type(C_PTR), value :: d_eps_pair, d_e_tot
     end function rimp2_energy_fitted_device

     !> No counterpart in the reference: rimp2_energy_fitted_device with no
     !> device allocation (scratch in a static array of 128 MiB per GPU), for
     !> callers that cannot allocate (graph capture, serving loops).
     integer(C_INT) function rimp2_energy_fitted_device_noalloc(ctx, naux, nvirt, nocc, d_Calpha, d_eps_o, d_eps_v, &
          & mode, d_eps_pair, d_e_tot) bind(C, name="rimp2_energy_fitted_device_noalloc")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: naux, nvirt, nocc
       type(C_PTR), value :: d_Calpha, d_eps_o, d_eps_v
       integer(C_INT), value :: mode
       type(C_PTR), value :: d_eps_pair, d_e_tot
     end function rimp2_energy_fitted_device_noalloc

     ! ---- intermediate results (host arrays, blocking) ----

     integer(C_INT) function rimp2_get_cholesky(ctx, L) bind(C, name="rimp2_get_cholesky")
       import :: C_INT, C_PTR, C_DOUBLE
       type(C_PTR), value :: ctx
       real(C_DOUBLE) :: L(*)
     end function rimp2_get_cholesky

     integer(C_INT) function rimp2_get_metric_inverse(ctx, Linv) bind(C, name="rimp2_get_metric_inverse")
       import :: C_INT, C_PTR, C_DOUBLE
       type(C_PTR), value :: ctx
       real(C_DOUBLE) :: Linv(*)
     end function rimp2_get_metric_inverse

     integer(C_INT) function rimp2_get_B(ctx, B) bind(C, name="rimp2_get_B")
       import :: C_INT, C_PTR, C_DOUBLE
       type(C_PTR), value :: ctx
       real(C_DOUBLE) :: B(*)
     end function rimp2_get_B

     ! ---- batched fragments ----

     integer(C_INT) function rimp2_batch_prepare(ctx, nfrag, frags, mode) bind(C, name="rimp2_batch_prepare")
       import :: C_INT, C_PTR, C_INT64_T, rimp2_frag_desc
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nfrag
       type(rimp2_frag_desc) :: frags(*)
       integer(C_INT), value :: mode
     end function rimp2_batch_prepare

     integer(C_INT) function rimp2_batch_prepare_device(ctx, nfrag, frags, mode) &
          & bind(C, name="rimp2_batch_prepare_device")
       import :: C_INT, C_PTR, C_INT64_T, rimp2_frag_desc
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nfrag
       type(rimp2_frag_desc) :: frags(*)
       integer(C_INT), value :: mode
     end function rimp2_batch_prepare_device

     !> Replays the captured pair stage; d_e_frag (device, may be C_NULL_PTR)
     integer(C_INT) function rimp2_batch_run(ctx, d_e_frag) bind(C, name="rimp2_batch_run")
       import :: C_INT, C_PTR
       type(C_PTR), value :: ctx, d_e_frag
     end function rimp2_batch_run

     integer(C_INT) function rimp2_batch_result(ctx, e_frag, eps_pair_packed) bind(C, name="rimp2_batch_result")
       import :: C_INT, C_PTR, C_DOUBLE
       type(C_PTR), value :: ctx
       real(C_DOUBLE) :: e_frag(*)
       type(C_PTR), value :: eps_pair_packed
     end function rimp2_batch_result

     integer(C_INT) function rimp2_batch_tile(ctx, tile) bind(C, name="rimp2_batch_tile")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T) :: tile
     end function rimp2_batch_tile

     integer(C_INT) function rimp2_batch_tuning(ctx, n, tiles, ms) bind(C, name="rimp2_batch_tuning")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT) :: n
       integer(C_INT64_T) :: tiles(*)
       real(C_DOUBLE) :: ms(*)
     end function rimp2_batch_tuning

     !> tile candidate of later prepares: 32, 64, -1 (mixed per fragment); 0 = autotune
     integer(C_INT) function rimp2_batch_set_tile(ctx, tile) bind(C, name="rimp2_batch_set_tile")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: tile
     end function rimp2_batch_set_tile

     ! ---- pair-screening surrogate ----

     !> feats(16, npairs): the C array [npairs, 16] seen column-major
     integer(C_INT) function rimp2_pair_features(ctx, nocc, nvirt, eps_o, eps_v, feats) &
          & bind(C, name="rimp2_pair_features")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE, C_FLOAT
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nocc, nvirt
       real(C_DOUBLE) :: eps_o(*), eps_v(*)
       real(C_FLOAT) :: feats(*)
     end function rimp2_pair_features

     integer(C_INT) function rimp2_mlp_init(ctx, seed) bind(C, name="rimp2_mlp_init")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: seed
     end function rimp2_mlp_init

     integer(C_INT) function rimp2_mlp_set_norm(ctx, mean, stdev, tmean, tstd) bind(C, name="rimp2_mlp_set_norm")
       import :: C_INT, C_PTR, C_FLOAT
       type(C_PTR), value :: ctx
       real(C_FLOAT) :: mean(*), stdev(*)
       real(C_FLOAT), value :: tmean, tstd
     end function rimp2_mlp_set_norm

     integer(C_INT) function rimp2_mlp_train_step(ctx, n, feats, target, lr, loss) bind(C, name="rimp2_mlp_train_step")
       import :: C_INT, C_PTR, C_INT64_T, C_FLOAT
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: n
       real(C_FLOAT) :: feats(*), target(*)
       real(C_FLOAT), value :: lr
       real(C_FLOAT) :: loss
     end function rimp2_mlp_train_step

     integer(C_INT) function rimp2_mlp_predict(ctx, n, feats, pred) bind(C, name="rimp2_mlp_predict")
       import :: C_INT, C_PTR, C_INT64_T, C_FLOAT
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: n
       real(C_FLOAT) :: feats(*), pred(*)
     end function rimp2_mlp_predict

     integer(C_INT) function rimp2_mlp_get_params(ctx, params, n) bind(C, name="rimp2_mlp_get_params")
       import :: C_INT, C_PTR, C_INT64_T, C_FLOAT
       type(C_PTR), value :: ctx
       real(C_FLOAT) :: params(*)
       integer(C_INT64_T), value :: n
     end function rimp2_mlp_get_params

     integer(C_INT) function rimp2_mlp_set_params(ctx, params, n) bind(C, name="rimp2_mlp_set_params")
       import :: C_INT, C_PTR, C_INT64_T, C_FLOAT
       type(C_PTR), value :: ctx
       real(C_FLOAT) :: params(*)
       integer(C_INT64_T), value :: n
     end function rimp2_mlp_set_params

     integer(C_INT64_T) function rimp2_mlp_num_params() bind(C, name="rimp2_mlp_num_params")
       import :: C_INT64_T
     end function rimp2_mlp_num_params

     integer(C_INT) function rimp2_mlp_grad(ctx, n, feats, target, loss, grad) bind(C, name="rimp2_mlp_grad")
       import :: C_INT, C_PTR, C_INT64_T, C_FLOAT
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: n
       real(C_FLOAT) :: feats(*), target(*), loss, grad(*)
     end function rimp2_mlp_grad

     integer(C_INT) function rimp2_screened_energy(ctx, nocc, nvirt, eps_o, eps_v, mode, thresh, eps_pair, e_tot, &
          & nskipped) bind(C, name="rimp2_screened_energy")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nocc, nvirt
       real(C_DOUBLE) :: eps_o(*), eps_v(*)
       integer(C_INT), value :: mode
       real(C_DOUBLE), value :: thresh
       real(C_DOUBLE) :: eps_pair(*), e_tot
       integer(C_INT64_T) :: nskipped
     end function rimp2_screened_energy

     ! ---- device memory, copies, stream ----

     integer(C_INT) function rimp2_device_alloc(ctx, bytes, dptr) bind(C, name="rimp2_device_alloc")
       import :: C_INT, C_PTR, C_SIZE_T
       type(C_PTR), value :: ctx
       integer(C_SIZE_T), value :: bytes
       type(C_PTR) :: dptr
     end function rimp2_device_alloc

     integer(C_INT) function rimp2_device_free(ctx, dptr) bind(C, name="rimp2_device_free")
       import :: C_INT, C_PTR
       type(C_PTR), value :: ctx, dptr
     end function rimp2_device_free

     integer(C_INT) function rimp2_copy_h2d(ctx, dst, src, bytes) bind(C, name="rimp2_copy_h2d")
       import :: C_INT, C_PTR, C_SIZE_T
       type(C_PTR), value :: ctx, dst, src
       integer(C_SIZE_T), value :: bytes
     end function rimp2_copy_h2d

     integer(C_INT) function rimp2_copy_d2h(ctx, dst, src, bytes) bind(C, name="rimp2_copy_d2h")
       import :: C_INT, C_PTR, C_SIZE_T
       type(C_PTR), value :: ctx, dst, src
       integer(C_SIZE_T), value :: bytes
     end function rimp2_copy_d2h

     integer(C_INT) function rimp2_sync(ctx) bind(C, name="rimp2_sync")
       import :: C_INT, C_PTR
       type(C_PTR), value :: ctx
     end function rimp2_sync

     ! ---- event timers ----

     integer(C_INT) function rimp2_timing_enable(ctx, on) bind(C, name="rimp2_timing_enable")
       import :: C_INT, C_PTR
       type(C_PTR), value :: ctx
       integer(C_INT), value :: on
     end function rimp2_timing_enable

     integer(C_INT) function rimp2_timing_get(ctx, kid, total_ms, count) bind(C, name="rimp2_timing_get")
       import :: C_INT, C_PTR, C_DOUBLE, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT), value :: kid
       real(C_DOUBLE) :: total_ms
       integer(C_INT64_T) :: count
     end function rimp2_timing_get

     integer(C_INT) function rimp2_timing_reset(ctx) bind(C, name="rimp2_timing_reset")
       import :: C_INT, C_PTR
       type(C_PTR), value :: ctx
     end function rimp2_timing_reset

     ! ---- seeded synthetic inputs ----

     integer(C_INT) function rimp2_gen_fragment_local(seed, naux, nvirt, nocc, lam, Pia, V, eps_o, eps_v) &
          & bind(C, name="rimp2_gen_fragment_local")
       import :: C_INT, C_INT64_T, C_DOUBLE
       integer(C_INT64_T), value :: seed, naux, nvirt, nocc
       real(C_DOUBLE), value :: lam
       real(C_DOUBLE) :: Pia(*), V(*), eps_o(*), eps_v(*)
     end function rimp2_gen_fragment_local

     integer(C_INT) function rimp2_gen_fragment_device(ctx, seed, naux, nvirt, nocc, lam, d_Pia, d_V, d_eps_o, &
          & d_eps_v) bind(C, name="rimp2_gen_fragment_device")
       import :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: seed, naux, nvirt, nocc
       real(C_DOUBLE), value :: lam
       type(C_PTR), value :: d_Pia, d_V, d_eps_o, d_eps_v
     end function rimp2_gen_fragment_device

     integer(C_INT) function rimp2_gen_batch_shapes(seed, nfrag, nocc, nvirt, naux) &
          & bind(C, name="rimp2_gen_batch_shapes")
       import :: C_INT, C_INT64_T
       integer(C_INT64_T), value :: seed, nfrag
       integer(C_INT64_T) :: nocc(*), nvirt(*), naux(*)
     end function rimp2_gen_batch_shapes

     ! ---- 4-index arrays and first order, device pointers ----

     integer(C_INT) function rimp2_eos_arrays_device(ctx, nvirt, nocc, noccEOS, nvirtEOS, nba, d_Calpha, d_EVocc, &
          & d_EVvirt, d_UoccEOST, d_UvirtT, d_UoccT, d_UvirtEOST, d_toccEOS, d_tvirtEOS, d_goccEOS, d_gvirtEOS) &
          & bind(C, name="rimp2_eos_arrays_device")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nvirt, nocc, noccEOS, nvirtEOS, nba
       type(C_PTR), value :: d_Calpha, d_EVocc, d_EVvirt, d_UoccEOST, d_UvirtT, d_UoccT, d_UvirtEOST
       type(C_PTR), value :: d_toccEOS, d_tvirtEOS, d_goccEOS, d_gvirtEOS
     end function rimp2_eos_arrays_device

     integer(C_INT) function rimp2_first_order_device(ctx, nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot, d_Calpha, &
          & d_CalphaOcc, d_CalphaVV, d_UoccEOST, d_UvirtT, d_UoccT, d_UvirtEOST, d_djik, d_blad) &
          & bind(C, name="rimp2_first_order_device")
       import :: C_INT, C_PTR, C_INT64_T
       type(C_PTR), value :: ctx
       integer(C_INT64_T), value :: nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot
       type(C_PTR), value :: d_Calpha, d_CalphaOcc, d_CalphaVV, d_UoccEOST, d_UvirtT, d_UoccT, d_UvirtEOST
       type(C_PTR), value :: d_djik, d_blad
     end function rimp2_first_order_device

  end interface

contains

  !> Aborts the host program with the library's last error text on a non-zero
  !> status (the reference assigned `stat` and never checked it,
  !> RIMP2.GPU.F90:81,118-121).
  subroutine rimp2_check(stat, ctx, what)
    integer(C_INT), intent(in) :: stat
    type(C_PTR), intent(in) :: ctx
    character(len=*), intent(in) :: what
    character(kind=C_CHAR) :: buf(1024)
    character(len=1024) :: msg
    integer :: k, ierr
    if (stat == RIMP2_OK) return
    ierr = rimp2_last_error(ctx, buf, int(1024, C_SIZE_T))
    msg = ' '
    do k = 1, 1024
       if (buf(k) == C_NULL_CHAR) exit
       msg(k:k) = buf(k)
    enddo
    print '(A,A,A,I0,A,A)', 'rimp2 error in ', what, ' (status ', stat, '): ', trim(msg)
    error stop 'ERROR3'
  end subroutine rimp2_check

end module rimp2_hip_interfaces
