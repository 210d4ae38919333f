!> Host driver in the role of the reference's PROGRAM MAIN (RIMP2.F90:1-377):
!> obtains a fragment, runs the RI-MP2 energy on the GPU through the C ABI and
!> verifies the selected precision mode against the FP64 path.
!>
!> usage:
!>   rimp2_driver.x gen  <seed> <naux> <nvirt> <nocc> <mode>
!>       seeded synthetic fragment: raw (P|ia), metric V, orbital energies
!>   rimp2_driver.x file <Calpha.data> <mode>
!>       fragment file in the reference's record order (RIMP2.F90:45-63):
!>       header (nvirt,nocc,noccEOS,nvirtEOS,NBA,nocctot), Calpha(NBA,nvirt,nocc),
!>       EVocc(nocc), EVvirt(nvirt).  Calpha is already fitted (B = Calpha) and
!>       goes through rimp2_energy_fitted / rimp2_load_fitted, no metric.
!> mode: 0 fp64, 1 fp32, 2 bf16x3, 3 i8s4, 4 i8s5.
!> Prints "E_MP2 <mode> <E>" and "dE_uEh <mode-vs-fp64>"; stops with
!> 'ERROR3' when the energy is not finite (the reference's failure word,
!> RIMP2.F90:266).
program rimp2_driver
  use iso_c_binding
  use rimp2_hip_interfaces
  implicit none
  character(len=512) :: arg, fname
  integer(C_INT64_T) :: seed, naux, nvirt, nocc
  integer :: hdr(6), lu, ios, imode
  integer(C_INT) :: mode, stat
  type(C_PTR) :: ctx
  real(C_DOUBLE), allocatable :: Pia(:,:,:), V(:,:), eo(:), ev(:), eps(:,:), eps64(:,:)
  real(C_DOUBLE) :: E, E64
  integer(C_SIZE_T) :: total, free
  logical :: fitted

  fitted = .false.
  call get_command_argument(1, arg)
  if (trim(arg) == 'gen') then
     call get_command_argument(2, arg); read(arg, *) seed
     call get_command_argument(3, arg); read(arg, *) naux
     call get_command_argument(4, arg); read(arg, *) nvirt
     call get_command_argument(5, arg); read(arg, *) nocc
     call get_command_argument(6, arg); read(arg, *) imode
     allocate(Pia(naux, nvirt, nocc), V(naux, naux), eo(nocc), ev(nvirt))
     stat = rimp2_gen_fragment(seed, naux, nvirt, nocc, Pia, V, eo, ev)
     call rimp2_check(stat, C_NULL_PTR, 'rimp2_gen_fragment')
  else if (trim(arg) == 'file') then
     call get_command_argument(2, fname)
     call get_command_argument(3, arg); read(arg, *) imode
     lu = 1345
     open(unit=lu, file=trim(fname), status='OLD', form='UNFORMATTED', iostat=ios)
     if (ios /= 0) then
        print *, 'cannot open ', trim(fname), ' iostat=', ios
        error stop 'ERROR3'
     endif
     read(lu, iostat=ios) hdr
     if (ios /= 0) error stop 'ERROR3'
     nvirt = hdr(1); nocc = hdr(2); naux = hdr(5)
     print *, 'nvirt,nocc,noccEOS,nvirtEOS,NBA,nocctot', hdr
     allocate(Pia(naux, nvirt, nocc), eo(nocc), ev(nvirt))
     fitted = .true.
     read(lu, iostat=ios) Pia
     if (ios /= 0) error stop 'ERROR3'
     read(lu, iostat=ios) eo
     if (ios /= 0) error stop 'ERROR3'
     read(lu, iostat=ios) ev
     if (ios /= 0) error stop 'ERROR3'
     close(lu)
  else
     print *, 'usage: rimp2_driver.x gen <seed> <naux> <nvirt> <nocc> <mode> | file <path> <mode>'
     error stop 'ERROR3'
  endif
  mode = int(imode, C_INT)

  stat = rimp2_ctx_create(0_C_INT, ctx)
  call rimp2_check(stat, C_NULL_PTR, 'rimp2_ctx_create')
  stat = rimp2_get_dev_mem(total, free)
  call rimp2_check(stat, ctx, 'rimp2_get_dev_mem')
  print '(A,I0,A,I0)', 'device memory total ', total, ' free ', free

  allocate(eps(nocc, nocc), eps64(nocc, nocc))
  if (fitted) then
     stat = rimp2_energy_fitted(ctx, naux, nvirt, nocc, Pia, eo, ev, RIMP2_MODE_FP64, eps64, E64)
     call rimp2_check(stat, ctx, 'rimp2_energy_fitted(fp64)')
     stat = rimp2_load_fitted(ctx, naux, nvirt, nocc, Pia, mode)
     call rimp2_check(stat, ctx, 'rimp2_load_fitted')
  else
     stat = rimp2_energy(ctx, naux, nvirt, nocc, Pia, V, eo, ev, RIMP2_MODE_FP64, eps64, E64)
     call rimp2_check(stat, ctx, 'rimp2_energy(fp64)')
     ! the metric is factored once; the other mode reuses it
     stat = rimp2_transform(ctx, naux, nvirt, nocc, Pia, mode)
     call rimp2_check(stat, ctx, 'rimp2_transform')
  endif
  stat = rimp2_pair_energies(ctx, nocc, nvirt, eo, ev, mode, eps, E)
  call rimp2_check(stat, ctx, 'rimp2_pair_energies')
  if (E /= E .or. E64 /= E64) error stop 'ERROR3'
  print '(A,I0,1X,ES25.16)', 'E_MP2 ', mode, E
  print '(A,ES25.16)', 'E_MP2_fp64 ', E64
  print '(A,ES14.6)', 'dE_uEh ', (E - E64) * 1.0e6_C_DOUBLE
  print '(A,ES25.16)', 'eps_11 ', eps(1, 1)
  stat = rimp2_ctx_destroy(ctx)
  call rimp2_check(stat, C_NULL_PTR, 'rimp2_ctx_destroy')
end program rimp2_driver
