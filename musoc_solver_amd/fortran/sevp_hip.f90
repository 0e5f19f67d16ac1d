! sevp_hip.f90 -- Fortran drop-in layer over the SEVP C ABI (include/sevp.h).
!
! Module sevp_hip binds the C entry points with ISO_C_BINDING.  After the
! module, external subroutines REPBIR, REP1, REP, REPBIR_PERIOD, PRE, PRE1
! and MATINV carry the reference's names and dummy-argument lists
! (MAIN/template/GLOBAL/src/sevp.f90:11,109,218,335 and
! PREP/template/GLOBAL/src/prep_GLOBAL.f90:327,457,516), so the unchanged
! TIMCOM callers (FS_SOLVER, INITFS) link against them in place of the CPU
! routines.  Like the reference, they have no status argument: a failure of
! the device path stops the program with the library's message.
!
! The plan interface (sevp_plan_create ... sevp_solve) is what a caller uses
! to keep the operator and influence matrices resident on the GPU across
! time steps instead of paying a host-device round trip per call.

module sevp_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: SEVP_FORM_FULL_PERIODIC = 0, SEVP_FORM_BIR = 1, &
                                       SEVP_FORM_FULL = 2
  integer(c_int), parameter, public :: SEVP_PREC_FP64 = 0, SEVP_PREC_FP32 = 1, SEVP_PREC_BF16 = 2, &
                                       SEVP_PREC_FP64_MFMA = 3
  integer(c_int), parameter, public :: SEVP_STORE_FP64 = 0, SEVP_STORE_PRODUCTS = 1

  type, bind(c), public :: sevp_desc
    integer(c_int) :: i2, j2, nb0
    type(c_ptr) :: ie
    integer(c_int) :: form, nbir, nsweep, periodic, nrhs_max, precision, seam_fix
  end type

  ! Step 14 of FS_GLOBAL on the device (sevp_step14_dev): all array members
  ! are device addresses of the reference's arrays (see include/sevp.h).
  type, bind(c), public :: sevp_step14_desc
    integer(c_int) :: i2, j2, k1, free_surface, periodic_we, lopenw, harness, mxmask
    real(c_float) :: tolerance, odt
  end type

  type, bind(c), public :: sevp_step14_grid
    type(c_ptr) :: kb, iu0, iv0, iu, iv, odx, ody, ocs, csv, odyv, odz
  end type

  type, bind(c), public :: sevp_step14_state
    type(c_ptr) :: u, v, w, p0, scr, s, x
  end type

  ! Krylov cross-check solvers (sevp_krylov_opts / sevp_krylov_stats).
  type, bind(c), public :: sevp_krylov_opts
    integer(c_int) :: periodic, ref_order, max_iter, restart
    real(c_double) :: tol, rtol
  end type

  type, bind(c), public :: sevp_krylov_stats
    integer(c_int) :: iterations, converged
    real(c_double) :: residual, seconds
  end type

  ! Settings of the drop-in P_BICGSTAB_LE, P_CSI and LANCZOS, whose argument
  ! lists cannot carry them: the reference's compile-time FLAG_PERIODIC_WE,
  ! and whether inner products are summed in the reference's sequential
  ! order (bitwise the reference, slow) or in the fast fixed parallel order.
  integer(c_int), public :: sevp_krylov_periodic = 1
  integer(c_int), public :: sevp_krylov_ref_order = 0

  public :: c_sevp_repbir, c_sevp_rep1, c_sevp_rep, c_sevp_repbir_period, c_sevp_pre, &
            c_sevp_pre1, c_sevp_matinv, sevp_plan_create, sevp_plan_destroy, &
            sevp_set_storage, sevp_build_influence, sevp_load_influence, sevp_get_influence, &
            sevp_solve, &
            sevp_reset_seam, sevp_step14_dev, sevp_evp_write, sevp_evp_read, sevp_check, &
            sevp_krylov_defaults, c_sevp_p_bicgstab_le, c_sevp_p_csi, c_sevp_lanczos

  interface
    integer(c_int) function c_sevp_repbir(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, &
                                          f, h, x, ie, i0, i2, m0, m2, nblk) &
        bind(c, name="sevp_repbir")
      import :: c_int, c_float, c_double
      real(c_float), intent(in) :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
      real(c_double), intent(in) :: rinv(*), rinv1(*)
      real(c_double), intent(inout) :: dum0(*), dum1(*), dum2(*), h(*), x(*)
      integer(c_int), intent(in) :: ie(*)
      integer(c_int), value :: i0, i2, m0, m2, nblk
    end function

    integer(c_int) function c_sevp_repbir_period(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, &
                                                 dum2, f, h, x, ie, i0, i2, m0, m2, nblk) &
        bind(c, name="sevp_repbir_period")
      import :: c_int, c_float, c_double
      real(c_float), intent(in) :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
      real(c_double), intent(in) :: rinv(*), rinv1(*)
      real(c_double), intent(inout) :: dum0(*), dum1(*), dum2(*), h(*), x(*)
      integer(c_int), intent(in) :: ie(*)
      integer(c_int), value :: i0, i2, m0, m2, nblk
    end function

    integer(c_int) function c_sevp_rep1(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, &
                                        f, h, x, ie, i0, i2, nblk) bind(c, name="sevp_rep1")
      import :: c_int, c_float, c_double
      real(c_float), intent(in) :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
      real(c_double), intent(in) :: rinv(*), rinv1(*)
      real(c_double), intent(inout) :: dum0(*), dum1(*), dum2(*), h(*), x(*)
      integer(c_int), intent(in) :: ie(*)
      integer(c_int), value :: i0, i2, nblk
    end function

    integer(c_int) function c_sevp_rep(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, &
                                       f, h, x, ie, i0, i2, nblk) bind(c, name="sevp_rep")
      import :: c_int, c_float, c_double
      real(c_float), intent(in) :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
      real(c_double), intent(in) :: rinv(*), rinv1(*)
      real(c_double), intent(inout) :: dum0(*), dum1(*), dum2(*), h(*), x(*)
      integer(c_int), intent(in) :: ie(*)
      integer(c_int), value :: i0, i2, nblk
    end function

    integer(c_int) function c_sevp_pre(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i2, m0, m2, &
                                       nblk) bind(c, name="sevp_pre")
      import :: c_int, c_float, c_double
      real(c_float), intent(in) :: ax(*), ay(*), bb(*), cx(*), cy(*)
      real(c_double), intent(inout) :: rinv(*), rinv1(*), h(*)
      integer(c_int), intent(in) :: ie(*)
      integer(c_int), value :: i2, m0, m2, nblk
    end function

    integer(c_int) function c_sevp_pre1(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i0, i2, &
                                        nblk) bind(c, name="sevp_pre1")
      import :: c_int, c_float, c_double
      real(c_float), intent(in) :: ax(*), ay(*), bb(*), cx(*), cy(*)
      real(c_double), intent(inout) :: rinv(*), rinv1(*), h(*)
      integer(c_int), intent(in) :: ie(*)
      integer(c_int), value :: i0, i2, nblk
    end function

    integer(c_int) function c_sevp_matinv(b, n) bind(c, name="sevp_matinv")
      import :: c_int, c_double
      real(c_double), intent(inout) :: b(*)
      integer(c_int), value :: n
    end function

    integer(c_int) function sevp_plan_create(desc, al, ab, ac, ar, at, plan) &
        bind(c, name="sevp_plan_create")
      import :: c_int, c_float, c_ptr, sevp_desc
      type(sevp_desc), intent(in) :: desc
      real(c_float), intent(in) :: al(*), ab(*), ac(*), ar(*), at(*)
      type(c_ptr), intent(out) :: plan
    end function

    integer(c_int) function sevp_plan_destroy(plan) bind(c, name="sevp_plan_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function

    integer(c_int) function sevp_set_storage(plan, storage) bind(c, name="sevp_set_storage")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: storage
    end function

    integer(c_int) function sevp_build_influence(plan, zero_pivots) &
        bind(c, name="sevp_build_influence")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), intent(out) :: zero_pivots
    end function

    integer(c_int) function sevp_load_influence(plan, rinv, rinv1) &
        bind(c, name="sevp_load_influence")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: rinv(*), rinv1(*)
    end function

    integer(c_int) function sevp_get_influence(plan, rinv, rinv1) &
        bind(c, name="sevp_get_influence")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(out) :: rinv(*), rinv1(*)
    end function

    integer(c_int) function sevp_solve(plan, s, x, nrhs) bind(c, name="sevp_solve")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: plan
      real(c_float), intent(in) :: s(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), value :: nrhs
    end function

    integer(c_int) function sevp_reset_seam(plan) bind(c, name="sevp_reset_seam")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function

    integer(c_int) function sevp_step14_dev(plan, desc, grid, state, itmask, tmp) &
        bind(c, name="sevp_step14_dev")
      import :: c_int, c_float, c_ptr, sevp_step14_desc, sevp_step14_grid, sevp_step14_state
      type(c_ptr), value :: plan
      type(sevp_step14_desc), intent(in) :: desc
      type(sevp_step14_grid), intent(in) :: grid
      type(sevp_step14_state), intent(in) :: state
      integer(c_int), intent(out) :: itmask
      real(c_float), intent(out) :: tmp
    end function

    ! The EVP record (include/sevp.h): path is a NUL-terminated string
    ! (trim(name)//c_null_char); oac is c_loc(OAC) or c_null_ptr.
    integer(c_int) function sevp_evp_write(path, desc, al, ab, ac, ar, at, rinv, rinv1, oac, &
                                           max_subrecord) bind(c, name="sevp_evp_write")
      import :: c_int, c_char, c_float, c_double, c_long, c_ptr, sevp_desc
      character(kind=c_char), intent(in) :: path(*)
      type(sevp_desc), intent(in) :: desc
      real(c_float), intent(in) :: al(*), ab(*), ac(*), ar(*), at(*)
      real(c_double), intent(in) :: rinv(*), rinv1(*)
      type(c_ptr), value :: oac
      integer(c_long), value :: max_subrecord
    end function

    integer(c_int) function sevp_evp_read(path, desc, al, ab, ac, ar, at, rinv, rinv1, ie, oac, &
                                          with_oac, had_oac) bind(c, name="sevp_evp_read")
      import :: c_int, c_char, c_float, c_double, c_ptr, sevp_desc
      character(kind=c_char), intent(in) :: path(*)
      type(sevp_desc), intent(in) :: desc
      real(c_float), intent(out) :: al(*), ab(*), ac(*), ar(*), at(*)
      real(c_double), intent(out) :: rinv(*), rinv1(*)
      integer(c_int), intent(out) :: ie(*)
      type(c_ptr), value :: oac
      integer(c_int), value :: with_oac
      integer(c_int), intent(out) :: had_oac
    end function

    integer(c_int) function sevp_krylov_defaults(opts, solver) bind(c, name="sevp_krylov_defaults")
      import :: c_int, sevp_krylov_opts
      type(sevp_krylov_opts), intent(out) :: opts
      integer(c_int), value :: solver
    end function

    integer(c_int) function c_sevp_p_bicgstab_le(al, ab, ac, ar, at, b, x, i2, j2, opts, stats) &
        bind(c, name="sevp_p_bicgstab_le")
      import :: c_int, c_float, c_double, sevp_krylov_opts, sevp_krylov_stats
      real(c_float), intent(in) :: al(*), ab(*), ac(*), ar(*), at(*), b(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), value :: i2, j2
      type(sevp_krylov_opts), intent(in) :: opts
      type(sevp_krylov_stats), intent(out) :: stats
    end function

    integer(c_int) function c_sevp_p_csi(al, ab, ac, ar, at, mineig, maxeig, b, x, i2, j2, opts, &
                                         stats) bind(c, name="sevp_p_csi")
      import :: c_int, c_float, c_double, sevp_krylov_opts, sevp_krylov_stats
      real(c_float), intent(in) :: al(*), ab(*), ac(*), ar(*), at(*), b(*)
      real(c_double), value :: mineig, maxeig
      real(c_double), intent(inout) :: x(*)
      integer(c_int), value :: i2, j2
      type(sevp_krylov_opts), intent(in) :: opts
      type(sevp_krylov_stats), intent(out) :: stats
    end function

    integer(c_int) function c_sevp_lanczos(al, ab, ac, ar, at, i2, j2, m, opts, mineig, maxeig, &
                                           alpha, sigma) bind(c, name="sevp_lanczos")
      import :: c_int, c_float, c_double, c_ptr, sevp_krylov_opts
      real(c_float), intent(in) :: al(*), ab(*), ac(*), ar(*), at(*)
      integer(c_int), value :: i2, j2, m
      type(sevp_krylov_opts), intent(in) :: opts
      real(c_double), intent(out) :: mineig, maxeig
      type(c_ptr), value :: alpha, sigma
    end function

    type(c_ptr) function sevp_last_error_c() bind(c, name="sevp_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! Stop with the library's message if rc is not SEVP_OK.
  subroutine sevp_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: k
    if (rc == 0) return
    call c_f_pointer(sevp_last_error_c(), msg, [256])
    text = ' '
    do k = 1, 256
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    write (*, '(A,A,A,I0,A,A)') 'sevp: ', where, ' failed (', rc, '): ', trim(text)
    stop 'sevp: device solver failed'
  end subroutine

end module sevp_hip

! ---------------------------------------------------------------------------
! Drop-in external procedures with the reference's names and argument lists.

subroutine repbir(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, i2, m0, &
                  m2, nblk)
  use sevp_hip
  implicit none
  integer, intent(in) :: i0, i2, m0, m2, nblk
  real(8) :: rinv(*), rinv1(*), dum0(*), dum1(*), dum2(*), h(*), x(*)
  real :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
  integer :: ie(*)
  call sevp_check(c_sevp_repbir(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, &
                                i0, i2, m0, m2, nblk), 'REPBIR')
end subroutine

subroutine repbir_period(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, &
                         i2, m0, m2, nblk)
  use sevp_hip
  implicit none
  integer, intent(in) :: i0, i2, m0, m2, nblk
  real(8) :: rinv(*), rinv1(*), dum0(*), dum1(*), dum2(*), h(*), x(*)
  real :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
  integer :: ie(*)
  call sevp_check(c_sevp_repbir_period(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, &
                                       x, ie, i0, i2, m0, m2, nblk), 'REPBIR_PERIOD')
end subroutine

subroutine rep1(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, i2, nblk)
  use sevp_hip
  implicit none
  integer, intent(in) :: i0, i2, nblk
  real(8) :: rinv(*), rinv1(*), dum0(*), dum1(*), dum2(*), h(*), x(*)
  real :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
  integer :: ie(*)
  call sevp_check(c_sevp_rep1(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, &
                              i0, i2, nblk), 'REP1')
end subroutine

subroutine rep(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, i0, i2, nblk)
  use sevp_hip
  implicit none
  integer, intent(in) :: i0, i2, nblk
  real(8) :: rinv(*), rinv1(*), dum0(*), dum1(*), dum2(*), h(*), x(*)
  real :: ax(*), ay(*), bb(*), cx(*), cy(*), f(*)
  integer :: ie(*)
  call sevp_check(c_sevp_rep(ax, ay, bb, cx, cy, rinv, rinv1, dum0, dum1, dum2, f, h, x, ie, &
                             i0, i2, nblk), 'REP')
end subroutine

subroutine pre(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i2, m0, m2, nblk)
  use sevp_hip
  implicit none
  integer, intent(in) :: i2, m0, m2, nblk
  real(8) :: rinv(*), rinv1(*), h(*)
  real :: ax(*), ay(*), bb(*), cx(*), cy(*)
  integer :: ie(*)
  call sevp_check(c_sevp_pre(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i2, m0, m2, nblk), 'PRE')
end subroutine

subroutine pre1(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i0, i2, nblk)
  use sevp_hip
  implicit none
  integer, intent(in) :: i0, i2, nblk
  real(8) :: rinv(*), rinv1(*), h(*)
  real :: ax(*), ay(*), bb(*), cx(*), cy(*)
  integer :: ie(*)
  call sevp_check(c_sevp_pre1(ax, ay, bb, cx, cy, rinv, rinv1, h, ie, i0, i2, nblk), 'PRE1')
end subroutine

subroutine matinv(b, n)
  use sevp_hip
  implicit none
  integer :: n
  real(8) :: b(*)
  call sevp_check(c_sevp_matinv(b, n), 'MATINV')
end subroutine

! ---------------------------------------------------------------------------
! The Krylov cross-check solvers with the reference's names and argument
! lists (MAIN/template/GLOBAL/src/sevp.f90:443,553,618).  The preconditioner
! arguments CL..CT are accepted and unused, as in the reference, whose
! PRECONDITIONING is a copy.

subroutine p_bicgstab_le(al, ab, ac, ar, at, b, x, cl, cb, cc, cr, ct, i2, j2)
  use iso_c_binding
  use sevp_hip
  implicit none
  integer, intent(in) :: i2, j2
  real, intent(in) :: al(i2, j2), ab(i2, j2), ac(i2, j2), ar(i2, j2), at(i2, j2), b(i2, j2)
  real, intent(in) :: cl(i2, j2), cb(i2, j2), cc(i2, j2), cr(i2, j2), ct(i2, j2)
  real(8), intent(inout) :: x(i2 + 2, j2 + 2)
  type(sevp_krylov_opts) :: o
  type(sevp_krylov_stats) :: st
  call sevp_check(sevp_krylov_defaults(o, 0_c_int), 'P_BICGSTAB_LE')
  o%periodic = sevp_krylov_periodic
  o%ref_order = sevp_krylov_ref_order
  call sevp_check(c_sevp_p_bicgstab_le(al, ab, ac, ar, at, b, x, i2, j2, o, st), 'P_BICGSTAB_LE')
end subroutine

subroutine p_csi(al, ab, ac, ar, at, mineig, maxeig, b, x, cl, cb, cc, cr, ct, i2, j2)
  use iso_c_binding
  use sevp_hip
  implicit none
  integer, intent(in) :: i2, j2
  real(8), intent(in) :: mineig, maxeig
  real, intent(in) :: al(i2, j2), ab(i2, j2), ac(i2, j2), ar(i2, j2), at(i2, j2), b(i2, j2)
  real, intent(in) :: cl(i2, j2), cb(i2, j2), cc(i2, j2), cr(i2, j2), ct(i2, j2)
  real(8), intent(inout) :: x(i2 + 2, j2 + 2)
  type(sevp_krylov_opts) :: o
  type(sevp_krylov_stats) :: st
  call sevp_check(sevp_krylov_defaults(o, 1_c_int), 'P_CSI')
  o%periodic = sevp_krylov_periodic
  o%ref_order = sevp_krylov_ref_order
  call sevp_check(c_sevp_p_csi(al, ab, ac, ar, at, mineig, maxeig, b, x, i2, j2, o, st), 'P_CSI')
end subroutine

subroutine lanczos(al, ab, ac, ar, at, mineig, maxeig, i2, j2)
  use iso_c_binding
  use sevp_hip
  implicit none
  integer, intent(in) :: i2, j2
  real, intent(in) :: al(i2, j2), ab(i2, j2), ac(i2, j2), ar(i2, j2), at(i2, j2)
  real(8), intent(out) :: mineig, maxeig
  type(sevp_krylov_opts) :: o
  call sevp_check(sevp_krylov_defaults(o, 0_c_int), 'LANCZOS')
  o%ref_order = sevp_krylov_ref_order
  call sevp_check(c_sevp_lanczos(al, ab, ac, ar, at, i2, j2, 10_c_int, o, mineig, maxeig, &
                                 c_null_ptr, c_null_ptr), 'LANCZOS')
end subroutine
