! The EVP record as the reference's own Fortran statements write and read it,
! compiled by flang for the EVP file tests:
!   PREP  WRITE(FNO_EVP) AL,AR,AB,AT,AC,RINV,RINV1,IE
!         (PREP/template/GLOBAL/src/prep_GLOBAL.f90:320)
!   MAIN  READ(FN_EVP) AL,AR,AB,AT,AC,OAC,RINV,RINV1,IE
!         (MAIN/template/GLOBAL/src/input_GLOBAL.f90:438-443)
! both on a unit opened CONVERT='BIG_ENDIAN', FORM='unformatted'.
!
! Usage: evp_io MODE RAW EVP I2 J2 NB0 W2 NMAT
!   MODE  write-prep | write-main   RAW -> EVP
!         read-prep  | read-main    EVP -> RAW
! RAW is a native-endian stream file holding AL,AR,AB,AT,AC,OAC,RINV,RINV1,IE
! (OAC always present there); RINV is (W2, NB0, NMAT), RINV1 (W2, NB0-1, NMAT).
! Exit status 3 when the READ fails (the record does not hold the list).
program evp_io
  implicit none
  character(len=1024) :: mode, raw, evp, arg
  integer :: i2, j2, nb0, w2, nmat, ios, k
  real, allocatable :: al(:,:), ar(:,:), ab(:,:), at(:,:), ac(:,:), oac(:,:)
  real(8), allocatable :: rinv(:,:,:), rinv1(:,:,:)
  integer, allocatable :: ie(:)
  integer :: dims(5)

  call get_command_argument(1, mode)
  call get_command_argument(2, raw)
  call get_command_argument(3, evp)
  do k = 1, 5
    call get_command_argument(3 + k, arg)
    read (arg, *) dims(k)
  end do
  i2 = dims(1); j2 = dims(2); nb0 = dims(3); w2 = dims(4); nmat = dims(5)
  allocate (al(i2,j2), ar(i2,j2), ab(i2,j2), at(i2,j2), ac(i2,j2), oac(i2,j2))
  allocate (rinv(w2,nb0,nmat), rinv1(w2,nb0-1,nmat), ie(nb0))

  if (mode(1:5) == 'write') then
    open (20, file=trim(raw), access='stream', form='unformatted', status='old')
    read (20) al, ar, ab, at, ac, oac, rinv, rinv1, ie
    close (20)
    open (10, convert='BIG_ENDIAN', file=trim(evp), form='unformatted', status='replace')
    if (trim(mode) == 'write-prep') then
      write (10) al, ar, ab, at, ac, rinv, rinv1, ie
    else
      write (10) al, ar, ab, at, ac, oac, rinv, rinv1, ie
    end if
    close (10)
  else
    oac = 0.
    open (10, convert='BIG_ENDIAN', file=trim(evp), form='unformatted', status='old')
    if (trim(mode) == 'read-prep') then
      read (10, iostat=ios) al, ar, ab, at, ac, rinv, rinv1, ie
    else
      read (10, iostat=ios) al, ar, ab, at, ac, oac, rinv, rinv1, ie
    end if
    close (10)
    if (ios /= 0) call exit(3)
    open (20, file=trim(raw), access='stream', form='unformatted', status='replace')
    write (20) al, ar, ab, at, ac, oac, rinv, rinv1, ie
    close (20)
  end if
end program evp_io
