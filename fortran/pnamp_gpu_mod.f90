! pnamp_gpu_mod.f90 -- ISO_C_BINDING interface to libpnamp_gpu.so
! (include/pnamp_gpu.h).  A Fortran PNAMP host uses it in place of the
! per-pair Kernel4Triaxial call (Codes/PNAMP/PNAMP_Main.f90:100-118): the
! module arrays that Kernel4Triaxial reads (tnljm, H, AMP, PNP, Input) are
! handed over once through pnamp_create, the HFB states through
! pnamp_set_states, and each pair through pnamp_kernel_pair (or all pairs at
! once through pnamp_kernels_batched, with the E2 kernels through
! pnamp_kernels_e2).
module pnamp_gpu
  use iso_c_binding
  implicit none
  private
  public :: pnamp_basis_t, pnamp_interaction_t, pnamp_mesh_t
  public :: pnamp_create, pnamp_set_states, pnamp_kernel_pair, pnamp_kernels_batched
  public :: pnamp_hw_solve, pnamp_last_timing, pnamp_last_error_f, pnamp_destroy
  public :: pnamp_set_form, pnamp_states_separable, pnamp_project_zres
  public :: pnamp_kernels_batched_noalloc, pnamp_set_q2, pnamp_norb, pnamp_kernels_e2, pnamp_hw_presolve
  public :: pnamp_autotune, pnamp_set_tune, pnamp_graph_info
  public :: PNAMP_NOBS, PNAMP_FORM_PLAIN, PNAMP_FORM_FACTORED, PNAMP_FORM_MIXED_FP32, PNAMP_FORM_MIXED_BF16

  integer(c_int), parameter :: PNAMP_NOBS = 6
  integer(c_int), parameter :: PNAMP_FORM_PLAIN = 0, PNAMP_FORM_FACTORED = 1
  integer(c_int), parameter :: PNAMP_FORM_MIXED_FP32 = 2, PNAMP_FORM_MIXED_BF16 = 3

  type, bind(C) :: pnamp_basis_t
    integer(c_int) :: n
    type(c_ptr) :: t, nq, l, twoj, twom
  end type

  type, bind(C) :: pnamp_interaction_t
    integer(c_int64_t) :: nme
    type(c_ptr) :: ka, kb, kc, kd
    type(c_ptr) :: me2b, t1b
    real(c_double) :: e0
  end type

  type, bind(C) :: pnamp_mesh_t
    integer(c_int) :: nalp, nbet, ngam
    type(c_ptr) :: alp, walp, bet, wbet, gam, wgam
    integer(c_int) :: nphi, jmax, fold
    real(c_double) :: cscale
    integer(c_int) :: iscale
  end type

  interface
    integer(c_int) function pnamp_create(basis, inter, mesh, device, ctx) bind(C, name="pnamp_create")
      import :: c_int, c_ptr, pnamp_basis_t, pnamp_interaction_t, pnamp_mesh_t
      type(pnamp_basis_t), intent(in) :: basis
      type(pnamp_interaction_t), intent(in) :: inter
      type(pnamp_mesh_t), intent(in) :: mesh
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function pnamp_set_states(ctx, nstates, u, v) bind(C, name="pnamp_set_states")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstates
      real(c_double), intent(in) :: u(*), v(*)
    end function

    integer(c_int) function pnamp_set_form(ctx, form) bind(C, name="pnamp_set_form")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: form
    end function

    integer(c_int) function pnamp_states_separable(ctx) bind(C, name="pnamp_states_separable")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function pnamp_project_zres(ctx, npairs, zres, kern) bind(C, name="pnamp_project_zres")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: npairs
      type(c_ptr), value :: zres, kern
    end function

    integer(c_int) function pnamp_kernel_pair(ctx, q0, q1, zval, nval, zres, kern) bind(C, name="pnamp_kernel_pair")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: q0, q1, zval, nval
      type(c_ptr), value :: zres, kern
    end function

    integer(c_int) function pnamp_kernels_batched(ctx, npairs, q0, q1, zval, nval, zres, kern) &
        bind(C, name="pnamp_kernels_batched")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: npairs
      integer(c_int32_t), intent(in) :: q0(*), q1(*)
      integer(c_int), value :: zval, nval
      type(c_ptr), value :: zres, kern
    end function

    integer(c_int) function pnamp_hw_solve(ctx, dim, nmat, hmat, eps, nos, en, e, f, nkept) &
        bind(C, name="pnamp_hw_solve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: dim
      real(c_double), intent(in) :: nmat(*), hmat(*)
      real(c_double), value :: eps
      integer(c_int), value :: nos
      real(c_double), intent(out) :: en(*), e(*)
      type(c_ptr), value :: f
      integer(c_int), intent(out) :: nkept
    end function

    integer(c_int) function pnamp_kernels_batched_noalloc(ctx, npairs, q0, q1, zval, nval, zres, kern) &
        bind(C, name="pnamp_kernels_batched_noalloc")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: npairs
      integer(c_int32_t), intent(in) :: q0(*), q1(*)
      integer(c_int), value :: zval, nval
      type(c_ptr), value :: zres, kern
    end function

    integer(c_int) function pnamp_set_q2(ctx, q2) bind(C, name="pnamp_set_q2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q2(*)
    end function

    integer(c_int) function pnamp_norb(ctx) bind(C, name="pnamp_norb")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function pnamp_kernels_e2(ctx, npairs, q0, q1, zval, nval, zres, kern, ze2, qred, rho1b) &
        bind(C, name="pnamp_kernels_e2")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: npairs
      integer(c_int32_t), intent(in) :: q0(*), q1(*)
      integer(c_int), value :: zval, nval
      type(c_ptr), value :: zres, kern, ze2, qred, rho1b
    end function

    integer(c_int) function pnamp_hw_presolve(ctx, dim, nmat, hmat, eps, kmax, en, m, e, var, nos_best, e_best) &
        bind(C, name="pnamp_hw_presolve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: dim
      real(c_double), intent(in) :: nmat(*), hmat(*)
      real(c_double), value :: eps
      integer(c_int), value :: kmax
      real(c_double), intent(out) :: en(*)
      integer(c_int), intent(out) :: m
      real(c_double), intent(out) :: e(*), var(*)
      integer(c_int), intent(out) :: nos_best(*)
      real(c_double), intent(out) :: e_best(*)
    end function

    integer(c_int) function pnamp_autotune(ctx, npairs, q0, q1, zval, nval, choice, cand, ms, ncand) &
        bind(C, name="pnamp_autotune")
      import :: c_int, c_int32_t, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: npairs
      integer(c_int32_t), intent(in) :: q0(*), q1(*)
      integer(c_int), value :: zval, nval
      integer(c_int), intent(out) :: choice(3), cand(*)
      real(c_float), intent(out) :: ms(*)
      integer(c_int), intent(out) :: ncand
    end function

    integer(c_int) function pnamp_set_tune(ctx, pf_waves, rest_waves, chunk_div) bind(C, name="pnamp_set_tune")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: pf_waves, rest_waves, chunk_div
    end function

    integer(c_int) function pnamp_graph_info(ctx, captures, replays, nodes) bind(C, name="pnamp_graph_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: captures, replays, nodes
    end function

    integer(c_int) function pnamp_last_timing(ctx, ms_points, ms_total) bind(C, name="pnamp_last_timing")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(out) :: ms_points, ms_total
    end function

    type(c_ptr) function pnamp_last_error_c(ctx) bind(C, name="pnamp_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    subroutine pnamp_destroy(ctx) bind(C, name="pnamp_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  function pnamp_last_error_f(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: k, nlen
    p = pnamp_last_error_c(ctx)
    nlen = int(c_strlen(p))
    call c_f_pointer(p, chars, [nlen])
    allocate(character(len=nlen) :: msg)
    do k = 1, nlen
      msg(k:k) = chars(k)
    end do
  end function

end module pnamp_gpu
