! dropin_driver.f90 -- test program in the role of the reference's callers.
!
! Reads a case written by tests/test_gpu_fortran.py, then calls the drop-in
! external procedures PRE, REPBIR, PRE1 and REP1 (musoc_solver_amd/lib/
! libsevp_fortran.so) with the reference's calling conventions:
!   - PREP's strip loop (PREP/template/GLOBAL/src/prep_GLOBAL.f90:286-317):
!     PRE on AL(IL,1)... for each strip, the seam strip on gathered copies;
!   - FS_SOLVER's sweep loop (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:
!     1904-1966): REPBIR on X(IL,1) for each strip, the seam strip on XX;
!   - PRE1 and REP1 on the full width (FL_EVP_STP = 0);
!   - FS_SOLVER's cross-checks (fs_GLOBAL.f90:1853,2024-2026): P_BICGSTAB_LE,
!     LANCZOS and P_CSI, with the inner products in the reference's order.
! Writes the influence matrices and the solutions for comparison.
program dropin_driver
  use sevp_hip, only: sevp_krylov_ref_order
  implicit none
  integer :: i2, j2, nbir, nb0, nsweep, ibir, inc, i0, j0, j1, n, il, i, j, it, zp
  integer, allocatable :: ie(:)
  real, allocatable :: al(:,:), ab(:,:), ac(:,:), ar(:,:), at(:,:), s(:,:)
  real, allocatable :: src(:,:), cl(:,:), cb(:,:), cc(:,:), cr(:,:), ct(:,:)
  real(8), allocatable :: rinv(:,:,:), rinv1(:,:,:), x(:,:), xx(:,:), h(:,:), dum0(:,:)
  real(8), allocatable :: dum1(:), dum2(:)
  real(8), allocatable :: rf(:,:,:), rf1(:,:,:), xf(:,:), hf(:,:), df0(:,:), df1(:), df2(:)
  real(8), allocatable :: xk(:,:), xc(:,:)
  real(8) :: emin, emax
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  read (10) i2, j2, nbir, nb0, nsweep
  allocate (ie(nb0))
  read (10) ie
  allocate (al(i2, j2), ab(i2, j2), ac(i2, j2), ar(i2, j2), at(i2, j2), s(i2, j2))
  read (10) al, ab, ac, ar, at, s
  i0 = i2 + 2
  j0 = j2 + 2
  j1 = j2 + 1
  allocate (x(i0, j0))
  read (10) x
  close (10)

  ibir = 2*i2/nbir
  inc = ibir/2
  allocate (rinv(ibir*ibir, nb0, nbir), rinv1(ibir*ibir, nb0 - 1, nbir))
  allocate (h(ibir + 2, j0), xx(ibir + 2, j0), dum0(ibir, nb0 - 1), dum1(ibir), dum2(ibir))
  allocate (src(ibir, j2), cl(ibir, j2), cb(ibir, j2), cc(ibir, j2), cr(ibir, j2), ct(ibir, j2))
  rinv = 0.d0
  rinv1 = 0.d0
  xx = 0.d0

  ! PREP: influence matrices of every strip
  do n = 1, nbir
    il = (n - 1)*inc + 1
    if (n /= nbir) then
      call pre(al(il, 1), ab(il, 1), ac(il, 1), ar(il, 1), at(il, 1), rinv(1, 1, n), &
               rinv1(1, 1, n), h, ie, i2, ibir + 2, ibir, nb0)
    else
      call gather(il)
      call pre(cl, cb, cc, cr, ct, rinv(1, 1, n), rinv1(1, 1, n), h, ie, ibir, ibir + 2, ibir, nb0)
    end if
  end do

  ! MAIN: BIR sweeps
  do it = 1, nsweep
    do n = 1, nbir
      il = (n - 1)*inc + 1
      if (n /= nbir) then
        do j = 1, j2
          do i = 1, ibir
            src(i, j) = s(il + i - 1, j)
          end do
        end do
        call repbir(al(il, 1), ab(il, 1), ac(il, 1), ar(il, 1), at(il, 1), rinv(1, 1, n), &
                    rinv1(1, 1, n), dum0, dum1, dum2, src, h, x(il, 1), ie, i0, i2, ibir + 2, &
                    ibir, nb0)
      else
        do j = 1, j2
          xx(1, j) = x(il, j)
          xx(ibir + 2, j) = x(inc + 2, j)
        end do
        call gather(il)
        call repbir(cl, cb, cc, cr, ct, rinv(1, 1, n), rinv1(1, 1, n), dum0, dum1, dum2, src, h, &
                    xx, ie, ibir + 2, ibir, ibir + 2, ibir, nb0)
        do j = 2, j1
          do i = 1, inc
            x(il + i, j) = xx(i + 1, j)
          end do
          do i = inc + 1, ibir
            x(i - inc + 1, j) = xx(i + 1, j)
          end do
        end do
        do j = 2, j1
          x(1, j) = x(i0 - 1, j)
          x(i0, j) = x(2, j)
        end do
      end if
    end do
  end do

  ! full width, periodic
  allocate (rf(i2, i2, nb0), rf1(i2, i2, nb0 - 1), xf(i0, j0), hf(i0, j0), df0(i2, nb0 - 1))
  allocate (df1(i2), df2(i2))
  call pre1(al, ab, ac, ar, at, rf, rf1, hf, ie, i0, i2, nb0)
  xf = 0.d0
  call rep1(al, ab, ac, ar, at, rf, rf1, df0, df1, df2, s, hf, xf, ie, i0, i2, nb0)

  ! the Krylov cross-check solvers, summing in the reference's order
  sevp_krylov_ref_order = 1
  allocate (xk(i0, j0), xc(i0, j0))
  xk = 0.d0
  call p_bicgstab_le(al, ab, ac, ar, at, s, xk, al, ab, ac, ar, at, i2, j2)
  call lanczos(al, ab, ac, ar, at, emin, emax, i2, j2)
  xc = 0.d0
  call p_csi(al, ab, ac, ar, at, emin, emax, s, xc, al, ab, ac, ar, at, i2, j2)

  open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (11) rinv, rinv1, x, rf, rf1, xf, xk, emin, emax, xc
  close (11)
  zp = 0
  print '(A)', 'dropin_driver done'

contains

  ! seam strip: columns il..il+inc-1, then 1..inc
  subroutine gather(il)
    integer, intent(in) :: il
    integer :: i, j
    do j = 1, j2
      do i = 1, inc
        src(i, j) = s(il + i - 1, j)
        cl(i, j) = al(il + i - 1, j)
        cb(i, j) = ab(il + i - 1, j)
        cc(i, j) = ac(il + i - 1, j)
        cr(i, j) = ar(il + i - 1, j)
        ct(i, j) = at(il + i - 1, j)
      end do
      do i = inc + 1, ibir
        src(i, j) = s(i - inc, j)
        cl(i, j) = al(i - inc, j)
        cb(i, j) = ab(i - inc, j)
        cc(i, j) = ac(i - inc, j)
        cr(i, j) = ar(i - inc, j)
        ct(i, j) = at(i - inc, j)
      end do
    end do
  end subroutine

end program dropin_driver
