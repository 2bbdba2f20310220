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
program pnamp_driver
  use iso_c_binding
  use pnamp_gpu
  implicit none
  integer(c_int), target, allocatable :: t(:), nq(:), l(:), twoj(:), twom(:)
  integer(c_int), target, allocatable :: ka(:), kb(:), kc(:), kd(:), q0(:), q1(:)
  real(c_double), target, allocatable :: me(:), t1b(:), alp(:), walp(:), bet(:), wbet(:), gam(:), wgam(:)
  real(c_double), target, allocatable :: u(:), v(:), coords(:,:)
  complex(c_double_complex), target, allocatable :: kern(:,:,:,:,:)
  integer(c_int) :: hdr(15)
  real(c_double) :: dh(2)
  integer :: n, nme, nalp, nbet, ngam, nphi, jmax, fold, iscale, nstates, npairs, zval, nval, ncp, ncn
  type(pnamp_basis_t) :: basis
  type(pnamp_interaction_t) :: inter
  type(pnamp_mesh_t) :: mesh
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  real(c_float) :: ms_pts, ms_tot
  character(len=512) :: casefile, outdir, formarg
  integer :: ip, u_in

  call get_command_argument(1, casefile)
  call get_command_argument(2, outdir)
  formarg = 'plain'
  if (command_argument_count() >= 3) call get_command_argument(3, formarg)
  open(newunit=u_in, file=trim(casefile), access='stream', form='unformatted', status='old')
  read(u_in) hdr, dh
  n = hdr(1); nme = hdr(2); nalp = hdr(3); nbet = hdr(4); ngam = hdr(5); nphi = hdr(6)
  jmax = hdr(7); fold = hdr(8); iscale = hdr(9); nstates = hdr(10); npairs = hdr(11)
  zval = hdr(12); nval = hdr(13); ncp = hdr(14); ncn = hdr(15)
  allocate(t(n), nq(n), l(n), twoj(n), twom(n), ka(nme), kb(nme), kc(nme), kd(nme))
  allocate(q0(npairs), q1(npairs), me(nme), t1b(n*n), alp(nalp), walp(nalp), bet(nbet), wbet(nbet))
  allocate(gam(ngam), wgam(ngam), u(n*n*nstates), v(n*n*nstates), coords(2, nstates))
  read(u_in) t, nq, l, twoj, twom, ka, kb, kc, kd, q0, q1
  read(u_in) me, t1b, alp, walp, bet, wbet, gam, wgam, u, v, coords
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
  rc = pnamp_kernels_batched(ctx, int(npairs, c_int), q0, q1, int(zval, c_int), int(nval, c_int), &
                             c_null_ptr, c_loc(kern))
  call check(rc, 'pnamp_kernels_batched')
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

  ! post_general (PNAMP_IO.f90:39-94): norm, norm-weighted H, <N>, <Z>, <J^2>,
  ! and the not-computed Q slots as zeros; |N| < CHOP = 1e-8 -> zeros.
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
          write(lou, '(8e15.8)') (0.d0, 0.d0), (0.d0, 0.d0)
          write(lou, '(8e15.8)') (0.d0, 0.d0), (0.d0, 0.d0)
        end do
      end do
    end do
    close(lou)
  end subroutine

end program pnamp_driver
