! qs3x_dsflan.f90 -- continued-fraction evaluation of S(q, w) from the
! alpha / beta^2 file the DSF recurrence writes (QS3's DSFLan,
! source_DSFLan/main.f90:1-96).
!
! Reads alpha_beta_values_E0-tmp.d from the working directory (columns:
! ex_i, n, itr, alpha_n, beta2_n, norm1, E0', sztot), evaluates
! S(w) = -Im G(w) / pi on a grid and writes DSF_itrnum<M>_eta00<eta*100>.dat.
! Defaults are DSFLan's: eta = 0.04, w = 0, 0.0001, ..., 30 (300,001 points)
! and its evaluation of the deepest level.  Optional arguments:
!   qs3x_dsflan [eta [wmax [npoints [exact]]]]
! where exact = 1 divides the deepest level's imaginary part by its
! denominator (the exact continued fraction).
program qs3x_dsflan
  use iso_c_binding
  use qs3x_iso_c, only: qs3x_continued_fraction, qs3x_check
  implicit none
  integer, parameter :: maxlev = 100000
  real(8), allocatable :: alpha(:), beta2(:), omega(:), dsf(:)
  real(8) :: norm1, E0, sztot, a, b, eta, wmax
  integer :: n, ex_i0, dummy, itr, ios, npts, om, quirk, nargs
  character(len=64) :: arg, itr_name, eta_name
  eta = 0.04d0
  wmax = 30.0d0
  npts = 300001
  quirk = 1
  nargs = command_argument_count()
  if (nargs >= 1) then
    call get_command_argument(1, arg); read(arg, *) eta
  end if
  if (nargs >= 2) then
    call get_command_argument(2, arg); read(arg, *) wmax
  end if
  if (nargs >= 3) then
    call get_command_argument(3, arg); read(arg, *) npts
  end if
  if (nargs >= 4) then
    call get_command_argument(4, arg); read(arg, *) n
    if (n == 1) quirk = 0
  end if
  allocate(alpha(maxlev), beta2(maxlev))
  open(162, file='alpha_beta_values_E0-tmp.d', status='old')
  n = 0
  do
    read(162, *, iostat=ios) ex_i0, dummy, itr, a, b, norm1, E0, sztot
    if (ios /= 0) exit
    n = n + 1
    alpha(n) = a
    beta2(n) = b
  end do
  close(162)
  if (n < 1) stop "no coefficients in alpha_beta_values_E0-tmp.d"
  allocate(omega(npts), dsf(npts))
  do om = 1, npts
    omega(om) = wmax * dble(om - 1) / dble(max(npts - 1, 1))
  end do
  call qs3x_check(qs3x_continued_fraction(int(n, c_int32_t), alpha, beta2, norm1, E0, eta, &
    int(npts, c_int64_t), omega, int(quirk, c_int32_t), dsf), "qs3x_dsflan")
  write(itr_name, '(I0)') n + 1
  write(eta_name, '(I0)') int(eta * 100)
  open(40, file='DSF_itrnum'//trim(itr_name)//'_eta00'//trim(eta_name)//'.dat', position='append')
  do om = 1, npts
    write(40, '(2e23.15)') omega(om), dsf(om)
  end do
  close(40)
end program qs3x_dsflan
