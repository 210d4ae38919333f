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
!>       When the file goes on with UoccEOST, UvirtT, UoccT, UvirtEOST
!>       (RIMP2.F90:55-63) the local-basis arrays toccEOS, tvirtEOS, goccEOS,
!>       gvirtEOS are formed on the GPU (Major Steps 5-8); when it then holds
!>       the expected arrays as generate.F90:333-336 writes them (toccEOS,
!>       tvirtEOS, gvirtEOS, goccEOS) every element is checked as
!>       RIMP2.F90:253-323 does (|test - got| > 1e-10 stops with ERROR3).
!>       The energy runs only when every denominator e_i + e_j - e_a - e_b is
!>       negative (the reference's fixtures have them positive).
!>   rimp2_driver.x file <Calpha.data> <mode> fo
!>       the same with the first-order record layout the reference reads when
!>       first_order is on (RIMP2.F90:65, 189, 230; generate.F90:262, 308,
!>       338-339): CalphaOcc(NBA,nocc,nocctot) and CalphaVV(NBA,nvirt,nvirt)
!>       after the U matrices, djik and blad after the four arrays; djik and
!>       blad are formed on the GPU and checked as RIMP2.F90:327-364 does.
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
  logical :: fitted, have_u, have_ref, energy_ok, first_order
  integer(C_INT64_T) :: nocctot
  real(C_DOUBLE), allocatable :: CalphaOcc(:,:,:), CalphaVV(:,:,:), djik(:,:,:,:), blad(:,:,:,:)
  real(C_DOUBLE), allocatable :: dref(:,:,:,:), bref(:,:,:,:)
  integer(C_INT64_T) :: noccEOS, nvirtEOS
  real(C_DOUBLE), allocatable :: UoccEOST(:,:), UvirtT(:,:), UoccT(:,:), UvirtEOST(:,:)
  real(C_DOUBLE), allocatable :: toccEOS(:,:,:,:), tvirtEOS(:,:,:,:), goccEOS(:,:,:,:), gvirtEOS(:,:,:,:)
  real(C_DOUBLE), allocatable :: tref(:,:,:,:), vref(:,:,:,:), gvref(:,:,:,:), goref(:,:,:,:)

  fitted = .false.
  have_u = .false.
  have_ref = .false.
  first_order = .false.
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
     call get_command_argument(4, arg)
     first_order = (trim(arg) == 'fo')
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
     noccEOS = hdr(3); nvirtEOS = hdr(4)
     allocate(UoccEOST(nocc, noccEOS), UvirtT(nvirt, nvirt), UoccT(nocc, nocc), UvirtEOST(nvirt, nvirtEOS))
     read(lu, iostat=ios) UoccEOST
     if (ios == 0) read(lu, iostat=ios) UvirtT
     if (ios == 0) read(lu, iostat=ios) UoccT
     if (ios == 0) read(lu, iostat=ios) UvirtEOST
     have_u = (ios == 0)
     nocctot = hdr(6)
     if (first_order) then
        if (.not. have_u) error stop 'ERROR3'
        allocate(CalphaOcc(naux, nocc, nocctot), CalphaVV(naux, nvirt, nvirt))
        read(lu, iostat=ios) CalphaOcc
        if (ios /= 0) error stop 'ERROR3'
        read(lu, iostat=ios) CalphaVV
        if (ios /= 0) error stop 'ERROR3'
     endif
     if (have_u) then
        allocate(tref(nvirt, noccEOS, nvirt, noccEOS), vref(nvirtEOS, nocc, nvirtEOS, nocc))
        allocate(gvref(nvirtEOS, nocc, nvirtEOS, nocc), goref(nvirt, noccEOS, nvirt, noccEOS))
        read(lu, iostat=ios) tref
        if (ios == 0) read(lu, iostat=ios) vref
        if (ios == 0) read(lu, iostat=ios) gvref
        if (ios == 0) read(lu, iostat=ios) goref
        have_ref = (ios == 0)
     endif
     if (first_order .and. have_ref) then
        allocate(dref(nvirt, noccEOS, noccEOS, nocctot), bref(nvirtEOS, nocc, nvirtEOS, nvirt))
        read(lu, iostat=ios) dref
        if (ios == 0) read(lu, iostat=ios) bref
        if (ios /= 0) error stop 'ERROR3'
     endif
     close(lu)
  else
     print *, 'usage: rimp2_driver.x gen <seed> <naux> <nvirt> <nocc> <mode> | file <path> <mode> [fo]'
     error stop 'ERROR3'
  endif
  mode = int(imode, C_INT)

  stat = rimp2_ctx_create(0_C_INT, ctx)
  call rimp2_check(stat, C_NULL_PTR, 'rimp2_ctx_create')
  stat = rimp2_get_dev_mem(total, free)
  call rimp2_check(stat, ctx, 'rimp2_get_dev_mem')
  print '(A,I0,A,I0)', 'device memory total ', total, ' free ', free

  if (have_u) then
     allocate(toccEOS(nvirt, noccEOS, nvirt, noccEOS), tvirtEOS(nvirtEOS, nocc, nvirtEOS, nocc))
     allocate(goccEOS(nvirt, noccEOS, nvirt, noccEOS), gvirtEOS(nvirtEOS, nocc, nvirtEOS, nocc))
     stat = rimp2_eos_arrays(ctx, nvirt, nocc, noccEOS, nvirtEOS, naux, Pia, eo, ev, UoccEOST, UvirtT, UoccT, &
          & UvirtEOST, toccEOS, tvirtEOS, goccEOS, gvirtEOS)
     call rimp2_check(stat, ctx, 'rimp2_eos_arrays')
     print '(A,4ES25.16)', 'EOS_sums ', sum(toccEOS), sum(tvirtEOS), sum(goccEOS), sum(gvirtEOS)
     if (have_ref) then
        call verify('toccEOS', tref, toccEOS)
        call verify('tvirtEOS', vref, tvirtEOS)
        call verify('gvirtEOS', gvref, gvirtEOS)
        call verify('goccEOS', goref, goccEOS)
        print '(A)', 'EOS arrays match the file within 1e-10'
     endif
  endif
  if (first_order) then
     allocate(djik(nvirt, noccEOS, noccEOS, nocctot), blad(nvirtEOS, nocc, nvirtEOS, nvirt))
     stat = rimp2_first_order(ctx, nvirt, nocc, noccEOS, nvirtEOS, naux, nocctot, Pia, CalphaOcc, CalphaVV, &
          & UoccEOST, UvirtT, UoccT, UvirtEOST, djik, blad)
     call rimp2_check(stat, ctx, 'rimp2_first_order')
     print '(A,2ES25.16)', 'FO_sums ', sum(djik), sum(blad)
     if (have_ref) then
        call verify('djik', dref, djik)
        call verify('blad', bref, blad)
        print '(A)', 'first-order arrays match the file within 1e-10'
     endif
  endif

  energy_ok = maxval(eo) < minval(ev)
  if (.not. energy_ok) then
     print '(A)', 'E_MP2 skipped: some denominator e_i + e_j - e_a - e_b >= 0'
     stat = rimp2_ctx_destroy(ctx)
     call rimp2_check(stat, C_NULL_PTR, 'rimp2_ctx_destroy')
     stop
  endif
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

contains

  !> The reference's element check (RIMP2.F90:257-271): the first element
  !> with |test - got| > 1e-10 is printed and the run stops with ERROR3.
  !> Written as .not. (|test - got| <= 1e-10) so that a NaN on either side
  !> fails too (the reference's ">" test lets a NaN pass).
  subroutine verify(what, test, got)
    character(len=*), intent(in) :: what
    real(C_DOUBLE), intent(in) :: test(:,:,:,:), got(:,:,:,:)
    integer :: a, b, c, d
    do d = 1, size(test, 4)
       do c = 1, size(test, 3)
          do b = 1, size(test, 2)
             do a = 1, size(test, 1)
                if (.not. (abs(test(a, b, c, d) - got(a, b, c, d)) <= 1.0e-10_C_DOUBLE)) then
                   print *, 'Test(A,B,C,D)', test(a, b, c, d)
                   print *, what, '(A,B,C,D)', got(a, b, c, d)
                   print *, 'Error ', a, b, c, d, 'DIFF', abs(test(a, b, c, d) - got(a, b, c, d))
                   error stop 'ERROR3'
                endif
             enddo
          enddo
       enddo
    enddo
  end subroutine verify

end program rimp2_driver
