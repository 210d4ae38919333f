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
!>   rimp2_driver.x batch <seed> <nfrag> <mode>
!>       nfrag seeded DEC fragments of uneven shape (rimp2_gen_batch_shapes)
!>       through the batched path: rimp2_batch (host arrays, one call), then
!>       rimp2_batch_prepare_device on copies in device memory with
!>       rimp2_batch_run / rimp2_batch_result; prints "E_FRAG k E" per fragment
!>       (the analogue of the reference's fragment loop, RIMP2.F90:36)
!>   rimp2_driver.x screen <seed> <naux> <nvirt> <nocc> <mode> <lam> <ntrain> <thresh>
!>       trains the pair-screening MLP on the exact pair energies and features
!>       of ntrain seeded fragments (locality-decay generator, lam), then runs a
!>       test fragment exactly (rimp2_energy) and screened
!>       (rimp2_screened_energy); prints "E_EXACT", "E_SCREENED", "NSKIPPED",
!>       "SKIP i j" per skipped pair and stops with ERROR3 when a kept pair's
!>       energy is not the exact kernel's
!>   rimp2_driver.x device <seed> <naux> <nvirt> <nocc> <mode>
!>       the device-resident path: buffers from rimp2_device_alloc, inputs from
!>       rimp2_gen_fragment_device, rimp2_energy_device + rimp2_sync, then the
!>       transformed B through rimp2_energy_fitted_device; prints "E_DEVICE"
!>       and "E_FITTED_DEVICE" and the pair kernel's event time
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
  if (trim(arg) == 'batch') then
     call run_batch()
     stop
  else if (trim(arg) == 'screen') then
     call run_screen()
     stop
  else if (trim(arg) == 'device') then
     call run_device()
     stop
  endif
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

  integer(C_INT64_T) function iarg(k)
    integer, intent(in) :: k
    character(len=128) :: a
    call get_command_argument(k, a)
    read(a, *) iarg
  end function iarg

  real(C_DOUBLE) function rarg(k)
    integer, intent(in) :: k
    character(len=128) :: a
    call get_command_argument(k, a)
    read(a, *) rarg
  end function rarg

  !> Batched fragments from host arrays and from device copies.
  subroutine run_batch()
    type frag_t
       real(C_DOUBLE), allocatable :: P(:,:,:), V(:,:), eo(:), ev(:)
       type(C_PTR) :: d(4)
    end type frag_t
    type(frag_t), allocatable, target :: fr(:)
    type(rimp2_frag_desc), allocatable :: desc(:), ddesc(:)
    integer(C_INT64_T), allocatable :: no(:), nv(:), nq(:)
    real(C_DOUBLE), allocatable :: ef(:), ef2(:)
    integer(C_INT64_T) :: sd, nf, k, tile, npp
    integer(C_INT) :: md, st, ntun
    integer(C_INT64_T) :: tiles(4)
    real(C_DOUBLE) :: tms(4)
    real(C_DOUBLE), allocatable, target :: pp(:)
    type(C_PTR) :: cx
    integer :: j
    sd = iarg(2)
    nf = iarg(3)
    md = int(iarg(4), C_INT)
    allocate(no(nf), nv(nf), nq(nf), fr(nf), desc(nf), ddesc(nf), ef(nf), ef2(nf))
    st = rimp2_gen_batch_shapes(sd, nf, no, nv, nq)
    call rimp2_check(st, C_NULL_PTR, 'rimp2_gen_batch_shapes')
    st = rimp2_ctx_create(0_C_INT, cx)
    call rimp2_check(st, C_NULL_PTR, 'rimp2_ctx_create')
    npp = 0
    do k = 1, nf
       allocate(fr(k)%P(nq(k), nv(k), no(k)), fr(k)%V(nq(k), nq(k)), fr(k)%eo(no(k)), fr(k)%ev(nv(k)))
       st = rimp2_gen_fragment(sd + k - 1, nq(k), nv(k), no(k), fr(k)%P, fr(k)%V, fr(k)%eo, fr(k)%ev)
       call rimp2_check(st, cx, 'rimp2_gen_fragment')
       desc(k) = rimp2_frag_desc(no(k), nv(k), nq(k), c_loc(fr(k)%P), c_loc(fr(k)%V), c_loc(fr(k)%eo), &
            & c_loc(fr(k)%ev))
       npp = npp + no(k) * no(k)
    enddo
    allocate(pp(npp))
    st = rimp2_batch(cx, nf, desc, md, ef, c_loc(pp))
    call rimp2_check(st, cx, 'rimp2_batch')
    ! the same fragments copied to device memory, prepared there, the graph replayed
    do k = 1, nf
       call to_device(cx, c_loc(fr(k)%P), 8 * size(fr(k)%P, kind=C_SIZE_T), fr(k)%d(1))
       call to_device(cx, c_loc(fr(k)%V), 8 * size(fr(k)%V, kind=C_SIZE_T), fr(k)%d(2))
       call to_device(cx, c_loc(fr(k)%eo), 8 * size(fr(k)%eo, kind=C_SIZE_T), fr(k)%d(3))
       call to_device(cx, c_loc(fr(k)%ev), 8 * size(fr(k)%ev, kind=C_SIZE_T), fr(k)%d(4))
       ddesc(k) = rimp2_frag_desc(no(k), nv(k), nq(k), fr(k)%d(1), fr(k)%d(2), fr(k)%d(3), fr(k)%d(4))
    enddo
    st = rimp2_batch_prepare_device(cx, nf, ddesc, md)
    call rimp2_check(st, cx, 'rimp2_batch_prepare_device')
    st = rimp2_batch_run(cx, C_NULL_PTR)
    call rimp2_check(st, cx, 'rimp2_batch_run')
    st = rimp2_batch_result(cx, ef2, C_NULL_PTR)
    call rimp2_check(st, cx, 'rimp2_batch_result')
    st = rimp2_batch_tile(cx, tile)
    call rimp2_check(st, cx, 'rimp2_batch_tile')
    st = rimp2_batch_tuning(cx, ntun, tiles, tms)
    call rimp2_check(st, cx, 'rimp2_batch_tuning')
    print '(A,I0,A,I0)', 'BATCH_TILE ', tile, ' TUNED ', ntun
    do j = 1, ntun
       print '(A,I0,1X,ES14.6)', 'TUNING ', tiles(j), tms(j)
    enddo
    do k = 1, nf
       if (ef(k) /= ef(k)) error stop 'ERROR3'
       print '(A,I0,1X,ES25.16,1X,ES25.16)', 'E_FRAG ', k - 1, ef(k), ef2(k)
       do j = 1, 4
          st = rimp2_device_free(cx, fr(k)%d(j))
          call rimp2_check(st, cx, 'rimp2_device_free')
       enddo
    enddo
    print '(A,ES25.16)', 'PAIR_SUM ', sum(pp)
    st = rimp2_ctx_destroy(cx)
    call rimp2_check(st, C_NULL_PTR, 'rimp2_ctx_destroy')
  end subroutine run_batch

  subroutine to_device(cx, src, bytes, d)
    type(C_PTR), intent(in) :: cx, src
    integer(C_SIZE_T), intent(in) :: bytes
    type(C_PTR), intent(out) :: d
    integer(C_INT) :: st
    st = rimp2_device_alloc(cx, bytes, d)
    call rimp2_check(st, cx, 'rimp2_device_alloc')
    st = rimp2_copy_h2d(cx, d, src, bytes)
    call rimp2_check(st, cx, 'rimp2_copy_h2d')
  end subroutine to_device

  !> Surrogate training on exact pair energies, then a screened energy.
  subroutine run_screen()
    integer(C_INT64_T) :: sd, nq, nv, no, ntr, np, kf, p, i, j, n, nskip, b0, nb, nparam
    integer(C_INT) :: md, st
    real(C_DOUBLE) :: lam, thresh, Ex, Es, Ef
    real(C_DOUBLE), allocatable :: P3(:,:,:), V2(:,:), eo(:), ev(:), eps(:,:), epss(:,:)
    real(C_FLOAT), allocatable :: F(:,:), T(:), mean(:), sdev(:), pred(:), params(:), grad(:)
    real(C_FLOAT) :: loss, tm, ts, lr
    integer :: ep
    type(C_PTR) :: cx
    sd = iarg(2); nq = iarg(3); nv = iarg(4); no = iarg(5); md = int(iarg(6), C_INT)
    lam = rarg(7); ntr = iarg(8); thresh = rarg(9)
    np = no * (no + 1) / 2
    allocate(P3(nq, nv, no), V2(nq, nq), eo(no), ev(nv), eps(no, no), epss(no, no))
    allocate(F(RIMP2_NFEAT, np * ntr), T(np * ntr), mean(RIMP2_NFEAT), sdev(RIMP2_NFEAT), pred(np))
    st = rimp2_ctx_create(0_C_INT, cx)
    call rimp2_check(st, C_NULL_PTR, 'rimp2_ctx_create')
    ! training set: features of each fragment's transformed B, targets log(-eps_ij)
    n = 0
    do kf = 0, ntr - 1
       st = rimp2_gen_fragment_local(sd + 1000 + kf, nq, nv, no, lam, P3, V2, eo, ev)
       call rimp2_check(st, cx, 'rimp2_gen_fragment_local')
       st = rimp2_energy(cx, nq, nv, no, P3, V2, eo, ev, md, eps, Ef)
       call rimp2_check(st, cx, 'rimp2_energy')
       st = rimp2_pair_features(cx, no, nv, eo, ev, F(:, n + 1:n + np))
       call rimp2_check(st, cx, 'rimp2_pair_features')
       do j = 1, no
          do i = 1, j
             n = n + 1
             T(n) = real(log(max(-eps(i, j), 1.0e-300_C_DOUBLE)), C_FLOAT)
          enddo
       enddo
    enddo
    ! standardisation: features, and the residual target T - feature 16
    do i = 1, RIMP2_NFEAT
       mean(i) = sum(F(i, 1:n)) / real(n, C_FLOAT)
       sdev(i) = sqrt(max(sum((F(i, 1:n) - mean(i))**2) / real(n, C_FLOAT), 0.0_C_FLOAT))
       if (sdev(i) < 1.0e-6_C_FLOAT) sdev(i) = 1.0_C_FLOAT
    enddo
    tm = sum(T(1:n) - F(RIMP2_NFEAT, 1:n)) / real(n, C_FLOAT)
    ts = sqrt(sum((T(1:n) - F(RIMP2_NFEAT, 1:n) - tm)**2) / real(n, C_FLOAT))
    ts = max(ts, 1.0e-6_C_FLOAT)
    st = rimp2_mlp_init(cx, 0_C_INT64_T)
    call rimp2_check(st, cx, 'rimp2_mlp_init')
    st = rimp2_mlp_set_norm(cx, mean, sdev, tm, ts)
    call rimp2_check(st, cx, 'rimp2_mlp_set_norm')
    nparam = rimp2_mlp_num_params()
    allocate(params(nparam), grad(nparam))
    st = rimp2_mlp_grad(cx, min(n, 256_C_INT64_T), F, T, loss, grad)
    call rimp2_check(st, cx, 'rimp2_mlp_grad')
    print '(A,I0,1X,ES14.6,1X,ES14.6)', 'MLP_PARAMS ', nparam, loss, sqrt(sum(grad**2))
    nb = 512
    do ep = 1, 6
       lr = 3.0e-3_C_FLOAT * 0.5_C_FLOAT**(ep - 1)
       do b0 = 1, n, nb
          st = rimp2_mlp_train_step(cx, min(nb, n - b0 + 1), F(:, b0:), T(b0:), lr, loss)
          call rimp2_check(st, cx, 'rimp2_mlp_train_step')
       enddo
       print '(A,I0,1X,ES14.6)', 'EPOCH_LOSS ', ep, loss
    enddo
    ! the trained parameters survive a get / set round trip
    st = rimp2_mlp_get_params(cx, params, nparam)
    call rimp2_check(st, cx, 'rimp2_mlp_get_params')
    st = rimp2_mlp_set_params(cx, params, nparam)
    call rimp2_check(st, cx, 'rimp2_mlp_set_params')
    ! test fragment: exact, then screened on the same transformed B
    st = rimp2_gen_fragment_local(sd, nq, nv, no, lam, P3, V2, eo, ev)
    call rimp2_check(st, cx, 'rimp2_gen_fragment_local')
    st = rimp2_energy(cx, nq, nv, no, P3, V2, eo, ev, md, eps, Ex)
    call rimp2_check(st, cx, 'rimp2_energy')
    st = rimp2_pair_features(cx, no, nv, eo, ev, F(:, 1:np))
    call rimp2_check(st, cx, 'rimp2_pair_features')
    st = rimp2_mlp_predict(cx, np, F, pred)
    call rimp2_check(st, cx, 'rimp2_mlp_predict')
    st = rimp2_screened_energy(cx, no, nv, eo, ev, md, thresh, epss, Es, nskip)
    call rimp2_check(st, cx, 'rimp2_screened_energy')
    print '(A,ES25.16)', 'E_EXACT ', Ex
    print '(A,ES25.16)', 'E_SCREENED ', Es
    print '(A,I0)', 'NSKIPPED ', nskip
    p = 0
    do j = 1, no
       do i = 1, j
          p = p + 1
          if (epss(i, j) == 0.0_C_DOUBLE) then
             print '(A,I0,1X,I0,1X,ES14.6)', 'SKIP ', i - 1, j - 1, pred(p)
          else if (epss(i, j) /= eps(i, j)) then
             print *, 'kept pair', i, j, 'screened', epss(i, j), 'exact', eps(i, j)
             error stop 'ERROR3'
          endif
       enddo
    enddo
    st = rimp2_ctx_destroy(cx)
    call rimp2_check(st, C_NULL_PTR, 'rimp2_ctx_destroy')
  end subroutine run_screen

  !> The device-resident path and its inspection calls.
  subroutine run_device()
    integer(C_INT64_T) :: sd, nq, nv, no, cnt
    integer(C_INT) :: md, st
    type(C_PTR) :: cx, dP, dV, deo, dev, dpe, det, dB
    real(C_DOUBLE), allocatable, target :: B(:,:,:), L(:,:), Li(:,:), eo(:), ev(:), eps(:,:)
    real(C_DOUBLE), target :: E, Ef
    real(C_DOUBLE) :: ms
    sd = iarg(2); nq = iarg(3); nv = iarg(4); no = iarg(5); md = int(iarg(6), C_INT)
    allocate(B(nq, nv, no), L(nq, nq), Li(nq, nq), eo(no), ev(nv), eps(no, no))
    st = rimp2_ctx_create(0_C_INT, cx)
    call rimp2_check(st, C_NULL_PTR, 'rimp2_ctx_create')
    st = rimp2_device_alloc(cx, int(8 * nq * nv * no, C_SIZE_T), dP)
    if (st == RIMP2_OK) st = rimp2_device_alloc(cx, int(8 * nq * nq, C_SIZE_T), dV)
    if (st == RIMP2_OK) st = rimp2_device_alloc(cx, int(8 * no, C_SIZE_T), deo)
    if (st == RIMP2_OK) st = rimp2_device_alloc(cx, int(8 * nv, C_SIZE_T), dev)
    if (st == RIMP2_OK) st = rimp2_device_alloc(cx, int(8 * no * no, C_SIZE_T), dpe)
    if (st == RIMP2_OK) st = rimp2_device_alloc(cx, 8_C_SIZE_T, det)
    if (st == RIMP2_OK) st = rimp2_device_alloc(cx, int(8 * nq * nv * no, C_SIZE_T), dB)
    call rimp2_check(st, cx, 'rimp2_device_alloc')
    st = rimp2_gen_fragment_device(cx, sd, nq, nv, no, 0.0_C_DOUBLE, dP, dV, deo, dev)
    call rimp2_check(st, cx, 'rimp2_gen_fragment_device')
    st = rimp2_timing_enable(cx, 1_C_INT)
    call rimp2_check(st, cx, 'rimp2_timing_enable')
    st = rimp2_timing_reset(cx)
    call rimp2_check(st, cx, 'rimp2_timing_reset')
    st = rimp2_energy_device(cx, nq, nv, no, dP, dV, deo, dev, md, dpe, det)
    call rimp2_check(st, cx, 'rimp2_energy_device')
    st = rimp2_sync(cx)
    call rimp2_check(st, cx, 'rimp2_sync')
    st = rimp2_timing_get(cx, RIMP2_K_PAIR, ms, cnt)
    call rimp2_check(st, cx, 'rimp2_timing_get')
    st = rimp2_copy_d2h(cx, c_loc(E), det, 8_C_SIZE_T)
    call rimp2_check(st, cx, 'rimp2_copy_d2h')
    st = rimp2_copy_d2h(cx, c_loc(eps), dpe, int(8 * no * no, C_SIZE_T))
    call rimp2_check(st, cx, 'rimp2_copy_d2h')
    st = rimp2_copy_d2h(cx, c_loc(eo), deo, int(8 * no, C_SIZE_T))
    call rimp2_check(st, cx, 'rimp2_copy_d2h')
    st = rimp2_copy_d2h(cx, c_loc(ev), dev, int(8 * nv, C_SIZE_T))
    call rimp2_check(st, cx, 'rimp2_copy_d2h')
    st = rimp2_get_cholesky(cx, L)
    call rimp2_check(st, cx, 'rimp2_get_cholesky')
    st = rimp2_get_metric_inverse(cx, Li)
    call rimp2_check(st, cx, 'rimp2_get_metric_inverse')
    st = rimp2_get_B(cx, B)
    call rimp2_check(st, cx, 'rimp2_get_B')
    print '(A,ES14.6)', 'MAX_LINV_L_MINUS_I ', maxval(abs(matmul(Li, L) - eye(nq)))
    ! the transformed B, fitted, through the device form of the fitted path
    st = rimp2_copy_h2d(cx, dB, c_loc(B), int(8 * nq * nv * no, C_SIZE_T))
    call rimp2_check(st, cx, 'rimp2_copy_h2d')
    st = rimp2_energy_fitted_device(cx, nq, nv, no, dB, deo, dev, md, C_NULL_PTR, det)
    call rimp2_check(st, cx, 'rimp2_energy_fitted_device')
    st = rimp2_sync(cx)
    call rimp2_check(st, cx, 'rimp2_sync')
    st = rimp2_copy_d2h(cx, c_loc(Ef), det, 8_C_SIZE_T)
    call rimp2_check(st, cx, 'rimp2_copy_d2h')
    if (E /= E .or. Ef /= Ef) error stop 'ERROR3'
    print '(A,ES25.16)', 'E_DEVICE ', E
    print '(A,ES25.16)', 'E_FITTED_DEVICE ', Ef
    print '(A,ES25.16)', 'PAIR_SUM ', sum(eps)
    print '(A,I0,1X,ES14.6)', 'PAIR_KERNEL_LAUNCHES_MS ', cnt, ms
    st = rimp2_device_free(cx, dP)
    if (st == RIMP2_OK) st = rimp2_device_free(cx, dV)
    if (st == RIMP2_OK) st = rimp2_device_free(cx, deo)
    if (st == RIMP2_OK) st = rimp2_device_free(cx, dev)
    if (st == RIMP2_OK) st = rimp2_device_free(cx, dpe)
    if (st == RIMP2_OK) st = rimp2_device_free(cx, det)
    if (st == RIMP2_OK) st = rimp2_device_free(cx, dB)
    call rimp2_check(st, cx, 'rimp2_device_free')
    st = rimp2_ctx_destroy(cx)
    call rimp2_check(st, C_NULL_PTR, 'rimp2_ctx_destroy')
  end subroutine run_device

  function eye(n) result(m)
    integer(C_INT64_T), intent(in) :: n
    real(C_DOUBLE) :: m(n, n)
    integer(C_INT64_T) :: k
    m = 0.0_C_DOUBLE
    do k = 1, n
       m(k, k) = 1.0_C_DOUBLE
    enddo
  end function eye

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
