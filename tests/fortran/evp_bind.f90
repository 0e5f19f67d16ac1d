! Reads an EVP record through the sevp_hip module's sevp_evp_read (either
! layout) and dumps it as evp_io.f90 does: a native stream file holding
! AL,AR,AB,AT,AC,OAC,RINV,RINV1,IE (OAC zero when the record has none), then
! writes it back in MAIN's layout with sevp_evp_write.
! Usage: evp_bind EVP RAW OUT I2 J2 NB0 FORM NBIR
! Host calls only: no device is used.
program evp_bind
  use iso_c_binding
  use sevp_hip
  implicit none
  character(len=1024) :: evp, raw, outf, arg
  integer :: dims(5), k, w2, nmat, had
  type(sevp_desc), target :: d
  real(c_float), allocatable, target :: al(:,:), ar(:,:), ab(:,:), at(:,:), ac(:,:), oac(:,:)
  real(c_double), allocatable :: rinv(:), rinv1(:)
  integer(c_int), allocatable, target :: ie(:)

  call get_command_argument(1, evp)
  call get_command_argument(2, raw)
  call get_command_argument(3, outf)
  do k = 1, 5
    call get_command_argument(3 + k, arg)
    read (arg, *) dims(k)
  end do
  d%i2 = dims(1); d%j2 = dims(2); d%nb0 = dims(3); d%form = dims(4); d%nbir = dims(5)
  d%nsweep = 0; d%periodic = 0; d%nrhs_max = 1; d%precision = SEVP_PREC_FP64; d%seam_fix = 0
  if (d%form == SEVP_FORM_BIR) then
    w2 = (2 * d%i2 / d%nbir)**2; nmat = d%nbir
  else
    w2 = d%i2 * d%i2; nmat = 1
  end if
  allocate (al(d%i2,d%j2), ar(d%i2,d%j2), ab(d%i2,d%j2), at(d%i2,d%j2), ac(d%i2,d%j2))
  allocate (oac(d%i2,d%j2), rinv(w2*d%nb0*nmat), rinv1(max(w2*(d%nb0-1)*nmat, 1)), ie(d%nb0))
  oac = 0.
  ie = 0
  d%ie = c_loc(ie)
  call sevp_check(sevp_evp_read(trim(evp)//c_null_char, d, al, ab, ac, ar, at, rinv, rinv1, ie, &
                                c_loc(oac), -1_c_int, had), 'sevp_evp_read')
  open (20, file=trim(raw), access='stream', form='unformatted', status='replace')
  write (20) al, ar, ab, at, ac, oac, rinv, rinv1(1:w2*(d%nb0-1)*nmat), ie
  close (20)
  call sevp_check(sevp_evp_write(trim(outf)//c_null_char, d, al, ab, ac, ar, at, rinv, rinv1, &
                                 c_loc(oac), 0_c_long), 'sevp_evp_write')
  print '(A,I0)', 'had_oac=', had
end program evp_bind
