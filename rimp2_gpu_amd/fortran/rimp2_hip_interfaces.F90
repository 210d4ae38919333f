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
  integer(C_INT), parameter :: RIMP2_MODE_FP64 = 0, RIMP2_MODE_FP32 = 1, RIMP2_MODE_BF16X3 = 2
  integer(C_INT), parameter :: RIMP2_MODE_I8S4 = 3, RIMP2_MODE_I8S5 = 4

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
