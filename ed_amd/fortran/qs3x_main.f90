! qs3x_main.f90 -- driver for the MI355X engine with QS3's input contract.
!
! Reads input.dat from stdin (namelists &input_parameters, &input_static,
! &input_dynamic, then &input_lancz or &input_TRLan), builds the (N_down, k)
! sector on the GPU, runs the solver selected by ALG, the static observables
! and the dynamical structure factor, and writes the reference's output files
! into OUTDIR.  Same flow as QS3's driver (main.f90:21-185); the modules it
! uses are the GPU-backed ones of qs3x_modules.f90.
program qs3x_main
  use state_lists, only: get_combination, combination, mk_shift_x_SQ, mk_shift_y_SQ, mk_shift_z_SQ, &
    allocate_lists_omp_SQ
  use input_param, only: read_ip, THS, NOS, NOD, NOV, LX, LY, LZ, shift_x_SQ, shift_y_SQ, shift_z_SQ, &
    ene, psi, ALG, cal_lm, cal_cf, cal_dsf, NOLM, NOCF, FILElm, FILECF, FILEpos, list_r, list_s, &
    qx, qy, qz, kvec_calc, rfield, itr_dsf, wr_wf, write_wf, re_wf, read_wf, OUTDIR
  use eigen_solver, only: my_trlanczos_routines_complex, Full_diag_routines, calcu_FM_energy, &
    i_vec_min, i_vec_max, read_input_TRLan, NOE
  use lanczos, only: read_lanczos_para, lanczos_routines
  use get_expectation_values, only: allocate_expe_mem, calcu_DSF_wavevector, get_lm_2_wave_vector, &
    get_cf_2_wave_vector
  implicit none
  real(8) :: eneFM
  complex(8) :: prdct
  integer :: k, nvec
  integer(8) :: c0, c1, rate

  call system_clock(c0, rate)
  write(*,'("********************************* Start QS^3 (MI355X) *******************************")')
  call read_ip
  call get_combination(NOS, NOD + 1)
  if (NOD > 0) then
    THS = combination(NOS, NOD)
  else
    THS = 1
  end if
  write(*,*) "  THS   = ", THS
  if (THS < 0) stop "THS is bigger than the upper limit of the integer kind (build with -fdefault-integer-8)"
  if (THS == 1) then
    eneFM = 0.0d0
    call calcu_FM_energy(eneFM)
    open(10, file=trim(adjustl(OUTDIR))//'energy.dat', position='append')
    write(10, *) eneFM
    close(10)
    write(*,*) "eneFM=", eneFM
    stop
  end if
  shift_x_SQ = mk_shift_x_SQ(LX, LY, LZ)
  shift_y_SQ = mk_shift_y_SQ(LX, LY, LZ)
  shift_z_SQ = mk_shift_z_SQ(LX, LY, LZ)
  call allocate_lists_omp_SQ
  write(*,*) "  THS(k)   = ", THS

  if (ALG == 1) then
    if (re_wf /= 1) then
      call lanczos_routines(ene, THS, psi)
      if (wr_wf == 1) call write_wf(THS, 1, 1, 1, ene)
    else
      call read_lanczos_para
      call read_wf(THS, 1, 1, 1)
    end if
    nvec = 1
  else if (ALG == 2) then
    if (re_wf /= 1) then
      call my_trlanczos_routines_complex(ene, psi, THS)
      if (wr_wf == 1) call write_wf(THS, i_vec_min, i_vec_max, NOE, ene)
    else
      call read_input_TRLan
      call read_wf(THS, i_vec_min, i_vec_max, NOE)
    end if
    if (NOV >= 2) then
      prdct = dot_product(psi(:, 1), psi(:, 2))
      print *, prdct
    end if
    do k = 1, NOV
      prdct = dot_product(psi(:, k), psi(:, k))
      print *, k, prdct
      psi(1:THS, k) = psi(1:THS, k) / sqrt(abs(prdct))
    end do
    nvec = NOV
  else
    call Full_diag_routines(ene, psi, THS)
    nvec = THS
  end if

  if (cal_lm == 1 .or. cal_cf == 1) call allocate_expe_mem(NOS, nvec, NOLM, NOCF, FILElm, FILECF, FILEpos)
  if (cal_lm == 1) call get_lm_2_wave_vector(psi, NOD, THS, nvec, NOS)
  if (cal_cf == 1) call get_cf_2_wave_vector(psi, NOCF, NOD, THS, nvec, list_r, list_s, NOS)

  if (cal_dsf == 1) then
    kvec_calc(1) = qx
    kvec_calc(2) = qy
    kvec_calc(3) = qz
    call calcu_DSF_wavevector(psi, ene(1:2), kvec_calc, rfield, NOS, NOD, itr_dsf, THS)
  end if
  call system_clock(c1)
  write(*,'("Time:",f10.3," [sec]")') dble(c1 - c0) / dble(rate)
end program qs3x_main
