! pnamp_driver.f90 -- Fortran PNAMP host on the MI355X library.
!
! Mirrors PNAMP_Main's pair loop (Codes/PNAMP/PNAMP_Main.f90:59-121): one
! batched library call replaces the rank-strided Kernel4Triaxial calls, and
! each pair is written in the post_general layout (Codes/PNAMP/PNAMP_IO.f90:39-94,
! formats (10i5) / (8e15.8)) with the CHOP rule of PNAMP_Kernel3D.f90:729-737.
!
! usage: pnamp_driver <case.bin> <outdir> [plain|factored]
!   case.bin is an unformatted stream file (see sccm_amd/driver.py:write_case).
!   factored: the factored FP64 form when the states have no proton-neutron
!   mixing (same kernels to rounding), else the plain form.
!   When the case file carries the Q_{2mu} matrices (iE2 = 1), the E2 kernels
!   fill the record's reduced Q2 fields: qpred(J,K1,K2,dJ) =
!   <J+dJ K1 q1 || e Q2 || J K2 q2> with e = (e_p, e_n) weighting the
!   proton and neutron parts, and qcpred the same with q1 and q2 exchanged,
!   (-1)^dJ conj(qpred(J+dJ, K2, K1, -dJ)) (zero where J + dJ > jmax).
program pnamp_driver
  use iso_c_binding
  use pnamp_gpu
  implicit none
  integer(c_int), target, allocatable :: t(:), nq(:), l(:), twoj(:), twom(:)
  integer(c_int), target, allocatable :: ka(:), kb(:), kc(:), kd(:), q0(:), q1(:)
  real(c_double), target, allocatable :: me(:), t1b(:), alp(:), walp(:), bet(:), wbet(:), gam(:), wgam(:)
  real(c_double), target, allocatable :: u(:), v(:), coords(:,:)
  complex(c_double_complex), target, allocatable :: kern(:,:,:,:,:), qred(:,:,:,:,:,:)
  real(c_double), target, allocatable :: q2(:)
  integer(c_int) :: hdr(16)
  real(c_double) :: dh(2), echarge(2)
  integer :: n, nme, nalp, nbet, ngam, nphi, jmax, fold, iscale, nstates, npairs, zval, nval, ncp, ncn
  type(pnamp_basis_t) :: basis
  type(pnamp_interaction_t) :: inter
  type(pnamp_mesh_t) :: mesh
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  real(c_float) :: ms_pts, ms_tot
  character(len=512) :: casefile, outdir, formarg
  integer :: ip, u_in, has_q2

  call get_command_argument(1, casefile)
  call get_command_argument(2, outdir)
  formarg = 'plain'
  if (command_argument_count() >= 3) call get_command_argument(3, formarg)
  open(newunit=u_in, file=trim(casefile), access='stream', form='unformatted', status='old')
  read(u_in) hdr, dh
  n = hdr(1); nme = hdr(2); nalp = hdr(3); nbet = hdr(4); ngam = hdr(5); nphi = hdr(6)
  jmax = hdr(7); fold = hdr(8); iscale = hdr(9); nstates = hdr(10); npairs = hdr(11)
  zval = hdr(12); nval = hdr(13); ncp = hdr(14); ncn = hdr(15); has_q2 = hdr(16)
  allocate(t(n), nq(n), l(n), twoj(n), twom(n), ka(nme), kb(nme), kc(nme), kd(nme))
  allocate(q0(npairs), q1(npairs), me(nme), t1b(n*n), alp(nalp), walp(nalp), bet(nbet), wbet(nbet))
  allocate(gam(ngam), wgam(ngam), u(n*n*nstates), v(n*n*nstates), coords(2, nstates))
  read(u_in) t, nq, l, twoj, twom, ka, kb, kc, kd, q0, q1
  read(u_in) me, t1b, alp, walp, bet, wbet, gam, wgam, u, v, coords
  if (has_q2 /= 0) then
    allocate(q2(5*n*n))
    read(u_in) echarge, q2
  end if
  close(u_in)

  basis = pnamp_basis_t(n, c_loc(t), c_loc(nq), c_loc(l), c_loc(twoj), c_loc(twom))
  inter = pnamp_interaction_t(int(nme, c_int64_t), c_loc(ka), c_loc(kb), c_loc(kc), c_loc(kd), &
                              c_loc(me), c_loc(t1b), dh(2))
  mesh = pnamp_mesh_t(nalp, nbet, ngam, c_loc(alp), c_loc(walp), c_loc(bet), c_loc(wbet), &
                      c_loc(gam), c_loc(wgam), nphi, jmax, fold, dh(1), iscale)
  rc = pnamp_create(basis, inter, mesh, 0_c_int, ctx)
  call check(rc, 'pnamp_create')
  rc = pnamp_set_states(ctx, int(nstates, c_int), u, v)
  call check(rc, 'pnamp_set_states')
  if (trim(formarg) == 'factored' .and. pnamp_states_separable(ctx) == 1) then
    rc = pnamp_set_form(ctx, PNAMP_FORM_FACTORED)
    call check(rc, 'pnamp_set_form')
    write(*, '(a)') ' pnamp_driver: factored form'
  end if
  allocate(kern(-jmax:jmax, -jmax:jmax, 0:jmax, PNAMP_NOBS, npairs))
  if (has_q2 /= 0) then
    rc = pnamp_set_q2(ctx, q2)
    call check(rc, 'pnamp_set_q2')
    ! C layout [pair][species][dJ + 2][J2][K1 + jmax + 2][K2 + jmax]
    allocate(qred(-jmax:jmax, -jmax-2:jmax+2, 0:jmax, -2:2, 2, npairs))
    rc = pnamp_kernels_e2(ctx, int(npairs, c_int), q0, q1, int(zval, c_int), int(nval, c_int), &
                          c_null_ptr, c_loc(kern), c_null_ptr, c_loc(qred), c_null_ptr)
    call check(rc, 'pnamp_kernels_e2')
  else
    rc = pnamp_kernels_batched(ctx, int(npairs, c_int), q0, q1, int(zval, c_int), int(nval, c_int), &
                               c_null_ptr, c_loc(kern))
    call check(rc, 'pnamp_kernels_batched')
  end if
  rc = pnamp_last_timing(ctx, ms_pts, ms_tot)
  write(*, '(a,i6,a,f12.3,a)') ' pnamp_driver: ', npairs, ' pairs, mesh-point kernel ', ms_pts, ' ms'
  do ip = 1, npairs
    call post_general(ip)
  end do
  call pnamp_destroy(ctx)

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(*, '(a,a,a,i6,a,a)') ' ', what, ' failed: ', code, ' ', pnamp_last_error_f(ctx)
      stop 1
    end if
  end subroutine

  ! e-weighted reduced Q2 element <J+dJ K1 q1 || e Q2 || J K2 q2> of pair ip, 0 outside the ranges
  complex(c_double_complex) function qp(ip, J, k1, k2, dJ)
    integer, intent(in) :: ip, J, k1, k2, dJ
    qp = (0.d0, 0.d0)
    if (has_q2 == 0 .or. J < 0 .or. J > jmax .or. J + dJ < 0) return
    if (abs(k1) > J + dJ .or. abs(k2) > J .or. abs(k1) > jmax + 2 .or. dJ < -2 .or. dJ > 2) return
    qp = echarge(1) * qred(k2, k1, J, dJ, 1, ip) + echarge(2) * qred(k2, k1, J, dJ, 2, ip)
  end function

  ! the same with the two states exchanged: (-1)^dJ conj(<J K2 q1 || e Q2 || J+dJ K1 q2>)
  complex(c_double_complex) function qcp(ip, J, k1, k2, dJ)
    integer, intent(in) :: ip, J, k1, k2, dJ
    qcp = (0.d0, 0.d0)
    if (J + dJ > jmax) return
    qcp = (-1)**abs(dJ) * conjg(qp(ip, J + dJ, k2, k1, -dJ))
  end function

  ! post_general (PNAMP_IO.f90:39-94): per (K2, K1) norm, norm-weighted H,
  ! <N>, <Z>, <J^2>, qpred/qcpred(dJ = 0) and q0p/q0pc (not set by the
  ! reference's Kernel4Triaxial text: zeros); then per K2 the dJ = 1 and 2
  ! reduced Q2 blocks over K1.  |N| < CHOP = 1e-8 -> zeros (N, H, <N>, <Z>, <J^2>).
  subroutine post_general(ip)
    integer, intent(in) :: ip
    integer :: jproj, k1, k2, k2max, lou
    complex(c_double_complex) :: zn, zh, znn, zzz, zj2
    character(len=600) :: fname
    real(c_double), parameter :: chop = 1.d-8
    write(fname, '(a,a,i4.4,a,i4.4,a)') trim(outdir), '/kern.', q0(ip), '.', q1(ip), '.elem'
    open(newunit=lou, file=trim(fname), status='replace')
    do jproj = 0, jmax
      if (jproj == 1) cycle
      if (fold == 1 .and. mod(jproj, 2) == 1) cycle
      write(lou, '(10i5)') jproj
      write(lou, '(8e15.8)') coords(1, q0(ip) + 1), coords(2, q0(ip) + 1), &
                             coords(1, q1(ip) + 1), coords(2, q1(ip) + 1)
      k2max = jproj
      if (mod(jproj, 2) == 1) k2max = abs(jproj - 1)
      do k2 = -k2max, k2max
        do k1 = -k2max, k2max
          write(lou, '(10i5)') k1, k2
          zn = kern(k2, k1, jproj, 1, ip)
          zh = (0.d0, 0.d0); znn = (0.d0, 0.d0); zzz = (0.d0, 0.d0); zj2 = (0.d0, 0.d0)
          if (abs(zn) >= chop) then
            zh = kern(k2, k1, jproj, 5, ip)
            zzz = kern(k2, k1, jproj, 2, ip) / zn + ncp
            znn = kern(k2, k1, jproj, 3, ip) / zn + ncn
            zj2 = kern(k2, k1, jproj, 6, ip) / zn
          else
            zn = (0.d0, 0.d0)
          end if
          write(lou, '(8e15.8)') zn, zh
          write(lou, '(8e15.8)') znn, zzz
          write(lou, '(8e15.8)') zj2
          write(lou, '(8e15.8)') qp(ip, jproj, k1, k2, 0), qcp(ip, jproj, k1, k2, 0)
          write(lou, '(8e15.8)') (0.d0, 0.d0), (0.d0, 0.d0)
        end do
      end do
      do k2 = -k2max, k2max
        do k1 = -k2max - 1, k2max + 1
          write(lou, '(8e15.8)') qp(ip, jproj, k1, k2, 1), qcp(ip, jproj, k1, k2, 1)
        end do
        do k1 = -k2max - 2, k2max + 2
          write(lou, '(8e15.8)') qp(ip, jproj, k1, k2, 2), qcp(ip, jproj, k1, k2, 2)
        end do
      end do
    end do
    close(lou)
  end subroutine

end program pnamp_driver
