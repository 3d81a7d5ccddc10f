! qs3x_modules.f90 -- Fortran modules with the reference's module names and
! procedure interfaces, implemented over libqs3x through qs3x_iso_c.
!
! A Fortran driver written against QS3's modules (input_param, state_lists,
! ham2vec, my_l1_blas, lanczos, eigen_solver, get_expectation_values) links
! these instead: the namelists, public variables, argument lists and output
! files are the reference's contract (input_param.f90:3-54; state_lists.f90;
! ham2vec_v3.f90:37-46; lanczos.f90:5-48; eigen_solver.f90:5-49;
! get_expectation_values.f90:4-38,197-279,379-605).  The work itself runs on
! the GPU; arrays handed back (list_s, list_r, ene, psi) are host copies with
! the reference's shapes.  Build with -fdefault-integer-8 for sectors whose
! C(NOS, N_down) exceeds 2^31 (L = 1000), as the reference needs too.

module input_param
  use qs3x_iso_c
  implicit none
  integer :: NOS, NOD, NOV, THS, NOLM, NOCF, NOxxz, ALG
  integer, allocatable :: p_xxz(:,:)
  real(8), allocatable :: Jint(:,:), sJint(:,:)
  character(100) :: FILExxz, FILElm, FILECF, FILEwf, FILEp, FILEp2
  character(100) :: OUTDIR
  integer :: random, check_p, wr_wf, re_wf, cal_lm, cal_cf, cal_dsf
  integer :: spsmsz, itr_dsf
  real(8) :: qx, qy, qz, kvec_calc(1:3), rfield
  character(100) :: FILEpos
  integer :: LX, LY, LZ, KX, KY, KZ
  real(8) :: PI, rkx, rky, rkz
  complex(8), allocatable :: explist(:,:,:)
  integer, allocatable :: list_s(:)
  real(8), allocatable :: list_r(:)
  integer, allocatable :: shift_x_SQ(:), shift_y_SQ(:), shift_z_SQ(:)
  integer :: NOD_new, THS_new
  real(8), allocatable :: list_r_new(:)
  integer, allocatable :: list_s_new(:)
  real(8) :: rkx_new, rky_new, rkz_new
  complex(8), allocatable :: explist_new(:,:,:)
  namelist /input_parameters/ NOS, NOD, LX, LY, LZ, KX, KY, KZ, NOxxz, &
    ALG, cal_lm, cal_cf, cal_dsf, wr_wf, re_wf, FILExxz, FILEwf, OUTDIR
  namelist /input_static/ NOV, NOLM, NOCF, FILElm, FILECF
  namelist /input_dynamic/ spsmsz, itr_dsf, qx, qy, qz, rfield, FILEpos
  real(8), allocatable :: ene(:)
  complex(8), allocatable :: psi(:,:)
contains

  ! Namelists from stdin, wave vector, phase table, bond file (swap to
  ! p1 < p2, sJint = (Jxy/2, Jz/4)), input_xxz.dat; then the GPU model.
  subroutine read_ip
    integer :: i, j, k, tmp
    integer(c_int32_t), allocatable :: pairs(:)
    real(c_double), allocatable :: jxy(:), jz(:)
    write(*,'(" ### Read input_parameters. ")')
    read(*, input_parameters)
    write(*, input_parameters)
    write(*,'(" ### Read input_static. ")')
    read(*, input_static)
    write(*, input_static)
    write(*,'(" ### Read input_dynamic. ")')
    read(*, input_dynamic)
    write(*, input_dynamic)
    PI = acos(-1.0d0)
    rkx = 2.0d0 * PI * dble(KX) / dble(LX)
    rky = 2.0d0 * PI * dble(KY) / dble(LY)
    rkz = 2.0d0 * PI * dble(KZ) / dble(LZ)
    allocate(p_xxz(2, NOxxz), Jint(NOxxz, 2:3), sJint(NOxxz, 2:3), shift_x_SQ(NOS), &
      shift_y_SQ(NOS), shift_z_SQ(NOS), explist(0:LX, 0:LY, 0:LZ))
    do k = 0, LZ
      do j = 0, LY
        do i = 0, LX
          explist(i, j, k) = exp((0.0d0, 1.0d0) * (rkx * dble(i) + rky * dble(j) + rkz * dble(k)))
        end do
      end do
    end do
    open(10, file=trim(adjustl(FILExxz)), status='old')
    do i = 1, NOxxz
      read(10, *) p_xxz(1, i), p_xxz(2, i), Jint(i, 2), Jint(i, 3)
      sJint(i, 2) = Jint(i, 2) * 0.5d0
      sJint(i, 3) = Jint(i, 3) * 0.25d0
      if (p_xxz(1, i) > p_xxz(2, i)) then
        tmp = p_xxz(1, i); p_xxz(1, i) = p_xxz(2, i); p_xxz(2, i) = tmp
      end if
    end do
    close(10)
    open(10, file=trim(adjustl(OUTDIR))//'input_xxz.dat', position='append')
    write(10, '("************************************************************************************")')
    write(10, *) "         i         j          J^{2}_{ij}           J^{3}_{ij}"
    write(10, '("************************************************************************************")')
    do i = 1, NOxxz
      write(10, *) p_xxz(1, i), p_xxz(2, i), Jint(i, 2), Jint(i, 3)
    end do
    write(10, '("************************************************************************************")')
    close(10)
    allocate(pairs(2 * NOxxz), jxy(NOxxz), jz(NOxxz))
    do i = 1, NOxxz
      pairs(2 * i - 1) = int(p_xxz(1, i), c_int32_t)
      pairs(2 * i) = int(p_xxz(2, i), c_int32_t)
      jxy(i) = Jint(i, 2); jz(i) = Jint(i, 3)
    end do
    call qs3x_check(qs3x_model_create(int(NOS, c_int32_t), int(LX, c_int32_t), int(LY, c_int32_t), &
      int(LZ, c_int32_t), int(NOxxz, c_int32_t), pairs, jxy, jz, qs3x_model_h), "read_ip")
  end subroutine read_ip

  ! Wave functions as unformatted sequential records (input_param.f90:119-147).
  subroutine write_wf(dim, i_vec_min, i_vec_max, NOE, ene)
    integer, intent(in) :: dim, i_vec_min, i_vec_max, NOE
    real(8), intent(in) :: ene(1:NOE)
    character(200) :: file_name
    character(500) :: file_name1
    integer :: i_vec, i
    if (wr_wf /= 1) return
    do i_vec = i_vec_min, i_vec_max
      write(file_name, '(A100,A2,I1,A4)') trim(adjustl(FILEwf)), 'wf', i_vec, '.bin'
      open(10, file=trim(adjustl(file_name)), status='unknown', action='write', form='unformatted')
      write(10) psi(1:dim, i_vec)
      close(10)
    end do
    write(file_name1, '(A300,A10)') trim(adjustl(FILEwf)), 'energy.dat'
    open(10, file=trim(adjustl(file_name1)), form='formatted')
    do i = 1, NOE
      write(10, '(2ES24.15)') dble(i), ene(i)
    end do
    close(10)
  end subroutine write_wf

  subroutine read_wf(dim, i_vec_min, i_vec_max, NOE)
    integer, intent(in) :: dim, i_vec_min, i_vec_max, NOE
    character(200) :: file_name
    character(500) :: file_name1
    integer :: i_vec, i
    real(8) :: dummy
    if (re_wf /= 1) return
    allocate(psi(1:dim, i_vec_min:i_vec_max))
    do i_vec = i_vec_min, i_vec_max
      write(file_name, '(A100,A2,I1,A4)') trim(adjustl(FILEwf)), 'wf', i_vec, '.bin'
      open(10, file=trim(adjustl(file_name)), form='unformatted')
      read(10) psi(1:dim, i_vec)
      close(10)
    end do
    write(file_name1, '(A300,A10)') trim(adjustl(FILEwf)), 'energy.dat'
    allocate(ene(1:NOE))
    ene(1:NOE) = 0.0d0
    open(10, file=trim(adjustl(file_name1)), status='old')
    do i = 1, NOE
      read(10, *) dummy, ene(i)
    end do
    close(10)
  end subroutine read_wf

end module input_param


module my_l1_blas
  implicit none
contains
  subroutine my_zaxpy(n, a, x, y)
    integer, intent(in) :: n
    complex(8), intent(in) :: a, x(n)
    complex(8), intent(inout) :: y(n)
    y(1:n) = a * x(1:n) + y(1:n)
  end subroutine my_zaxpy
  subroutine my_zdscal(n, a, x)
    integer, intent(in) :: n
    real(8), intent(in) :: a
    complex(8), intent(inout) :: x(n)
    x(1:n) = a * x(1:n)
  end subroutine my_zdscal
end module my_l1_blas


module state_lists
  use input_param
  implicit none
  integer, allocatable :: combination(:,:)
contains

  subroutine get_combination(NOS, NOD)
    integer, intent(in) :: NOS, NOD
    integer :: i, j
    if (allocated(combination)) deallocate(combination)
    allocate(combination(NOS, NOD))
    do j = 1, NOD
      do i = 1, NOS
        combination(i, j) = f_combination(i, j)
      end do
    end do
  end subroutine get_combination

  function f_combination(n, k) result(c)
    integer, intent(in) :: n, k
    integer :: c, i
    if (n < k) then
      c = 0
      return
    end if
    c = 1
    do i = 1, k
      c = c * (n - k + i)
      c = c / i
    end do
  end function f_combination

  function mk_shift_x_SQ(LX, LY, LZ) result(s)
    integer, intent(in) :: LX, LY, LZ
    integer :: s(LX*LY*LZ), x, y, z
    do z = 1, LZ
      do y = 1, LY
        do x = 1, LX
          s(x + LX*((y-1) + LY*(z-1))) = (mod(x, LX)+1) + LX*(y-1) + LX*LY*(z-1)
        end do
      end do
    end do
  end function mk_shift_x_SQ

  function mk_shift_y_SQ(LX, LY, LZ) result(s)
    integer, intent(in) :: LX, LY, LZ
    integer :: s(LX*LY*LZ), x, y, z
    do z = 1, LZ
      do y = 1, LY
        do x = 1, LX
          s(x + LX*((y-1) + LY*(z-1))) = x + LX*mod(y, LY) + LX*LY*(z-1)
        end do
      end do
    end do
  end function mk_shift_y_SQ

  function mk_shift_z_SQ(LX, LY, LZ) result(s)
    integer, intent(in) :: LX, LY, LZ
    integer :: s(LX*LY*LZ), x, y, z
    do z = 1, LZ
      do y = 1, LY
        do x = 1, LX
          s(x + LX*((y-1) + LY*(z-1))) = x + LX*(y-1) + LX*LY*mod(z, LZ)
        end do
      end do
    end do
  end function mk_shift_z_SQ

  ! Sector (NOD, rkx..rkz) built on the device; THS := dimension; host copies
  ! of list_s (ranks) and list_r (norms).
  subroutine allocate_lists_omp_SQ
    integer(c_int64_t) :: d
    integer(c_int64_t), allocatable :: ls(:)
    write(*,'(" ### Build the sector on the GPU. ")')
    if (c_associated(qs3x_sector_h)) call qs3x_check(qs3x_sector_destroy(qs3x_sector_h), "allocate_lists_omp_SQ")
    call qs3x_check(qs3x_sector_build(qs3x_model_h, int(NOD, c_int32_t), rkx, rky, rkz, qs3x_sector_h, d), &
      "allocate_lists_omp_SQ")
    THS = int(d)
    if (allocated(list_s)) deallocate(list_s)
    if (allocated(list_r)) deallocate(list_r)
    allocate(list_s(THS), list_r(THS), ls(max(THS, 1)))
    call qs3x_check(qs3x_sector_lists(qs3x_sector_h, ls, list_r), "allocate_lists_omp_SQ")
    list_s = int(ls(1:THS))
  end subroutine allocate_lists_omp_SQ

  ! DSF target sector (NOD_new, rk_new) (state_lists.f90:418-486).
  subroutine allocate_lists_omp_SQ_new
    integer(c_int64_t) :: d
    integer(c_int64_t), allocatable :: ls(:)
    if (c_associated(qs3x_new_h)) call qs3x_check(qs3x_sector_destroy(qs3x_new_h), "allocate_lists_omp_SQ_new")
    call qs3x_check(qs3x_sector_build(qs3x_model_h, int(NOD_new, c_int32_t), rkx_new, rky_new, rkz_new, &
      qs3x_new_h, d), "allocate_lists_omp_SQ_new")
    THS_new = int(d)
    if (allocated(list_s_new)) deallocate(list_s_new)
    if (allocated(list_r_new)) deallocate(list_r_new)
    allocate(list_s_new(THS_new), list_r_new(THS_new), ls(max(THS_new, 1)))
    call qs3x_check(qs3x_sector_lists(qs3x_new_h, ls, list_r_new), "allocate_lists_omp_SQ_new")
    list_s_new = int(ls(1:THS_new))
  end subroutine allocate_lists_omp_SQ_new

end module state_lists


module ham2vec
  use input_param
  implicit none
  type(c_ptr), save :: cached_h = c_null_ptr
  integer, save :: cached_lv = -1
contains

  ! v0 = H v1 (ham2vec_v3.f90:37-72).  The sector is the device one whose
  ! dimension and N_down match (main or DSF target); lists from elsewhere are
  ! uploaded once and cached.
  subroutine ham_to_vec_wave_vector(v0, v1, lv, NOD_, list_s_, list_r_, explist_)
    integer, intent(in) :: lv, NOD_
    complex(8), intent(in) :: v1(lv)
    complex(8), intent(out) :: v0(lv)
    integer, intent(in) :: list_s_(1:lv)
    real(8), intent(in) :: list_r_(1:lv)
    complex(8), intent(in) :: explist_(0:LX, 0:LY, 0:LZ)
    type(c_ptr) :: h
    integer(c_int64_t), allocatable :: ls(:)
    if (c_associated(qs3x_sector_h) .and. lv == THS .and. NOD_ == NOD) then
      h = qs3x_sector_h
    else if (c_associated(qs3x_new_h) .and. lv == THS_new .and. NOD_ == NOD_new) then
      h = qs3x_new_h
    else
      if (.not. (c_associated(cached_h) .and. cached_lv == lv)) then
        if (c_associated(cached_h)) call qs3x_check(qs3x_sector_destroy(cached_h), "ham_to_vec_wave_vector")
        allocate(ls(lv))
        ls = int(list_s_, c_int64_t)
        call qs3x_check(qs3x_sector_from_lists(qs3x_model_h, int(NOD_, c_int32_t), int(lv, c_int64_t), ls, &
          list_r_, explist_, cached_h), "ham_to_vec_wave_vector")
        cached_lv = lv
      end if
      h = cached_h
    end if
    call qs3x_check(qs3x_ham_to_vec_wave_vector(h, v0, v1), "ham_to_vec_wave_vector")
  end subroutine ham_to_vec_wave_vector

  ! Dense H (ham2vec_v3.f90:5-35).  The main sector's is assembled on the
  ! device, one H e_j per column; any other sector of dimension lv goes
  ! column by column through ham_to_vec_wave_vector.
  subroutine make_full_hamiltonian(lv, Ham)
    integer, intent(in) :: lv
    complex(8), intent(inout) :: Ham(lv, lv)
    complex(8), allocatable :: e(:)
    integer :: j
    if (c_associated(qs3x_sector_h) .and. lv == THS) then
      call qs3x_check(qs3x_make_full_hamiltonian(qs3x_sector_h, Ham, int(lv, c_int64_t)), "make_full_hamiltonian")
      return
    end if
    allocate(e(lv))
    do j = 1, lv
      e = (0.0d0, 0.0d0)
      e(j) = (1.0d0, 0.0d0)
      call ham_to_vec_wave_vector(Ham(:, j), e, lv, NOD, list_s, list_r, explist)
    end do
  end subroutine make_full_hamiltonian

end module ham2vec


module lanczos
  use my_l1_blas
  use ham2vec
  implicit none
  integer :: ite_lancz, maxitr, dim, minitr, itrint
  real(8) :: lnc_ene_conv
  namelist /input_lancz/ lnc_ene_conv, maxitr, minitr, itrint
contains

  subroutine read_lanczos_para
    read(*, input_lancz)
    write(*, input_lancz)
  end subroutine read_lanczos_para

  ! lanczos_routines (lanczos.f90:14-48): device-resident plain Lanczos.
  subroutine lanczos_routines(ene, THS_, psi)
    real(8), allocatable, intent(out) :: ene(:)
    integer, intent(in) :: THS_
    complex(8), allocatable, intent(out) :: psi(:,:)
    complex(8), allocatable :: xpsi(:,:)
    real(8) :: alpha(2)
    integer(c_int64_t) :: it
    dim = THS_
    allocate(ene(2))
    ene = 99.0d0
    read(*, input_lancz)
    write(*, input_lancz)
    write(*,'("************************************************************************************")')
    write(*,*) "Eigen solver: Lanczos (MI355X)"
    allocate(psi(dim, 1))
    call qs3x_check(qs3x_lanczos(qs3x_sector_h, qs3x_env_precision(), int(maxitr, c_int64_t), &
      int(minitr, c_int64_t), int(itrint, c_int64_t), lnc_ene_conv, c_null_ptr, 1_c_int64_t, &
      ene(1:2), psi(:, 1), it), "lanczos_routines")
    ene(2) = 99.0d0
    ite_lancz = int(it)
    write(*,'("************************************************************************************")')
    write(*,*) "total lanczos step:", ite_lancz
    write(*,'("E_0 = ",es25.15)') ene(1)
    allocate(xpsi(dim, 2))
    call ham_to_vec_wave_vector(xpsi(:, 1), psi(:, 1), dim, NOD, list_s, list_r, explist)
    call ham_to_vec_wave_vector(xpsi(:, 2), xpsi(:, 1), dim, NOD, list_s, list_r, explist)
    alpha(1) = dble(dot_product(psi(:, 1), xpsi(:, 1)))
    alpha(2) = dble(dot_product(psi(:, 1), xpsi(:, 2)))
    write(*,'("quality of eigenvector")')
    write(*,'("                   <GS| H |GS>  = ", es25.15)') alpha(1)
    write(*,'("|1 - (<GS|H|GS>)^2/<GS|H^2|GS>| = ", es25.15)') 1.0d0 - (alpha(1)**2)/alpha(2)
  end subroutine lanczos_routines

end module lanczos


module eigen_solver
  use my_l1_blas
  use ham2vec
  implicit none
  integer :: NOE, NOK, NOM, maxitr, ite_lancz, i_vec_min, i_vec_max
  integer :: dim
  real(8) :: lnc_ene_conv
  namelist /input_TRLan/ NOE, NOK, NOM, maxitr, lnc_ene_conv, i_vec_min, i_vec_max
contains

  ! my_trlanczos_routines_complex (eigen_solver.f90:10-44): device-resident
  ! thick restart; psi(:, 1:NOE) Ritz vectors, ene(1:NOE) Ritz values.
  subroutine my_trlanczos_routines_complex(ene, psi, THS_)
    real(8), allocatable, intent(out) :: ene(:)
    complex(8), allocatable, intent(out) :: psi(:,:)
    integer, intent(in) :: THS_
    real(8) :: stats(8)
    integer :: i
    read(*, input_TRLan)
    write(*, input_TRLan)
    if (NOE > NOK) stop "ne must be eaqual or smaller than NOK!"
    if (NOM < 2*NOK) NOM = 2*NOK
    dim = THS_
    if (dim < NOE .or. dim < NOK .or. dim < NOM) stop 999
    allocate(psi(dim, NOM+1), ene(NOM))
    psi = (0.0d0, 0.0d0)
    ene = 0.0d0
    write(*,*) "Reduced THS=", THS_
    write(*,*) "Eigen solver: thick restart Lanczos (MI355X)"
    call qs3x_check(qs3x_trlanczos(qs3x_sector_h, qs3x_env_precision(), int(NOE, c_int32_t), &
      int(NOK, c_int32_t), int(NOM, c_int32_t), int(maxitr, c_int64_t), lnc_ene_conv, c_null_ptr, &
      1_c_int64_t, ene, psi, stats), "my_trlanczos_routines_complex")
    ite_lancz = int(stats(2) + stats(3))
    write(*,*) "total lanczos step:", ite_lancz
    write(*,'("    i           E(i)")')
    do i = 1, NOE
      write(*,'(i5, 1es25.15)') i, ene(i)
    end do
    open(10, file=trim(adjustl(OUTDIR))//'energy.dat', position='append')
    do i = 1, NOE
      write(10,'(i5, 1es25.15)') i, ene(i)
    end do
    close(10)
  end subroutine my_trlanczos_routines_complex

  subroutine read_input_TRLan
    read(*, input_TRLan)
    write(*, input_TRLan)
  end subroutine read_input_TRLan

  ! Full_diag_routines / fulldia_routines_with_momentum
  ! (eigen_solver.f90:51-67,263-339): dense H of the main sector and all its
  ! eigenpairs on the device; full_eigenvalues.dat and the ground state in
  ! full_eigenvectors.dat as the reference writes them.
  subroutine Full_diag_routines(ene, psi, THS_)
    real(8), allocatable, intent(out) :: ene(:)
    complex(8), allocatable, intent(out) :: psi(:,:)
    integer, intent(in) :: THS_
    integer :: k
    dim = THS_
    allocate(ene(max(dim, 2)), psi(dim, dim))
    ene = 99.0d0
    print *, "make ham"
    print *, "diag starts"
    call qs3x_check(qs3x_full_diag(qs3x_sector_h, ene, psi, int(dim, c_int64_t)), "Full_diag_routines")
    print *, "Eigenvalues by Full Diagonalization"
    open(111, file=trim(adjustl(OUTDIR))//'full_eigenvalues.dat', position='append')
    do k = 1, dim
      write(111, '(es23.15)') ene(k)
    end do
    close(111)
    open(112, file=trim(adjustl(OUTDIR))//'full_eigenvectors.dat', position='append')
    do k = 1, dim
      write(112, *) psi(k, 1)
    end do
    close(112)
  end subroutine Full_diag_routines

  subroutine calcu_FM_energy(ene0)
    real(8), intent(inout) :: ene0
    call qs3x_check(qs3x_fm_energy(qs3x_model_h, ene0), "calcu_FM_energy")
  end subroutine calcu_FM_energy

end module eigen_solver


module get_expectation_values
  use my_l1_blas
  use input_param
  implicit none
  integer, allocatable :: list_expe_sz(:), list_expe_ss(:,:)
  real(8), allocatable :: expe_sz(:,:)
  real(8), allocatable :: expe_szsz(:,:)
  complex(8), allocatable :: expe_spsm(:,:)
  real(8), allocatable :: Spos(:,:)
contains

  ! allocate_expe_mem (get_expectation_values.f90:11-38): the same files.
  subroutine allocate_expe_mem(NOS_, NOV_, NOLM_, NOCF_, FILElm_, FILECF_, FILEpos_)
    integer, intent(in) :: NOS_, NOV_, NOLM_, NOCF_
    character(*), intent(in) :: FILElm_, FILECF_, FILEpos_
    integer :: i
    allocate(list_expe_sz(NOLM_), list_expe_ss(2, NOCF_), expe_sz(NOLM_, NOV_), expe_szsz(NOCF_, NOV_), &
      expe_spsm(NOCF_, NOV_))
    open(10, file=trim(adjustl(FILElm_)), status='old')
    do i = 1, NOLM_
      read(10, *) list_expe_sz(i)
    end do
    close(10)
    open(10, file=trim(adjustl(FILECF_)), status='old')
    do i = 1, NOCF_
      read(10, *) list_expe_ss(2, i), list_expe_ss(1, i)
    end do
    close(10)
    allocate(Spos(1:3, 1:NOS_))
    open(10, file=trim(adjustl(FILEpos_)), status='old')
    do i = 1, NOS_
      read(10, *) Spos(1, i), Spos(2, i), Spos(3, i)
    end do
    close(10)
  end subroutine allocate_expe_mem

  ! get_lm_2_wave_vector (get_expectation_values.f90:197-232): local_mag.dat.
  subroutine get_lm_2_wave_vector(psi_, NOD_, dim, nvec, NOS_)
    integer, intent(in) :: NOD_, nvec, NOS_, dim
    complex(8), intent(in) :: psi_(1:dim, 1:nvec)
    real(8), allocatable :: ez(:,:)
    integer :: j, k
    allocate(ez(NOS_, nvec))
    do k = 1, nvec
      call qs3x_check(qs3x_local_mag(qs3x_sector_h, psi_(:, k), ez(:, k)), "get_lm_2_wave_vector")
    end do
    open(10, file=trim(adjustl(OUTDIR))//'local_mag.dat', position='append')
    do j = 1, NOS_
      write(10, '(i8,10000es23.15)') j, (ez(j, k), k = 1, nvec)
    end do
    close(10)
    if (NOD_ < 0) stop
  end subroutine get_lm_2_wave_vector

  ! get_cf_2_wave_vector (get_expectation_values.f90:234-279): SzSz-corr.dat
  ! and SpSm-corr.dat for the FILECF pair list.
  subroutine get_cf_2_wave_vector(psi_, NOCF_, NOD_, dim, nvec, list_r_, list_s_, NOS_)
    integer, intent(in) :: NOCF_, NOD_, nvec, NOS_, dim
    complex(8), intent(in) :: psi_(1:dim, 1:nvec)
    real(8), intent(in) :: list_r_(1:dim)
    integer, intent(in) :: list_s_(1:dim)
    integer(c_int32_t), allocatable :: pairs(:)
    integer :: j, k
    allocate(pairs(2 * NOCF_))
    do j = 1, NOCF_
      pairs(2*j - 1) = int(list_expe_ss(2, j), c_int32_t)   ! file column 1
      pairs(2*j) = int(list_expe_ss(1, j), c_int32_t)       ! file column 2
    end do
    do k = 1, nvec
      call qs3x_check(qs3x_correlations(qs3x_sector_h, psi_(:, k), int(NOCF_, c_int64_t), pairs, &
        expe_szsz(:, k), expe_spsm(:, k)), "get_cf_2_wave_vector")
    end do
    open(10, file=trim(adjustl(OUTDIR))//'SzSz-corr.dat', position='append')
    do j = 1, NOCF_
      write(10, '(2i8,10000es25.15)') list_expe_ss(2, j), list_expe_ss(1, j), (expe_szsz(j, k), k = 1, nvec)
    end do
    close(10)
    open(10, file=trim(adjustl(OUTDIR))//'SpSm-corr.dat', position='append')
    do j = 1, NOCF_
      write(10, '(2i8,10000es25.15)') list_expe_ss(2, j), list_expe_ss(1, j), (expe_spsm(j, k), k = 1, nvec)
    end do
    close(10)
    if (NOD_ < 0 .or. NOS_ < 0 .or. size(list_r_) < 0 .or. size(list_s_) < 0) stop
  end subroutine get_cf_2_wave_vector

  ! calcu_DSF_wavevector (get_expectation_values.f90:379-605): target sector,
  ! S^a_q |GS>, the recurrence, and the alpha_beta_values_* files.
  subroutine calcu_DSF_wavevector(psi_, ene_, kvec, rfield_, NOS_, NOD_, itr_dsf_, dim)
    use state_lists, only: allocate_lists_omp_SQ_new
    integer, intent(in) :: dim, NOS_, NOD_, itr_dsf_
    complex(8), intent(in) :: psi_(1:dim, 1)
    real(8), intent(in) :: ene_(1:2), rfield_, kvec(1:3)
    complex(8), allocatable :: fn(:)
    real(8), allocatable :: alpha(:), beta2(:), pos(:)
    real(8) :: norm1(1), sztot, eps
    integer :: i, j, k, ex_i, itr
    character(2) :: ex_index
    character(10) :: sztot_index
    eps = 1.0d-8
    if (NOD_ == NOS_/2 .and. spsmsz == 2) then
      print *, "You don't need to calculate S(q,w)^-"
      stop
    end if
    if (NOD_ == 0 .or. NOD_ == 1 .and. spsmsz == 1) then
      print *, "You can not calculate S+(q,w) on the FM state and sztot=NOS/2-1 state!"
      stop
    end if
    rkx_new = rkx + qx
    rky_new = rky + qy
    rkz_new = rkz + qz
    if (spsmsz == 1) then
      NOD_new = NOD_ - 1
    else if (spsmsz == 2) then
      NOD_new = NOD_ + 1
    else
      NOD_new = NOD_
    end if
    call allocate_lists_omp_SQ_new
    allocate(explist_new(0:LX, 0:LY, 0:LZ))
    do k = 0, LZ
      do j = 0, LY
        do i = 0, LX
          explist_new(i, j, k) = exp((0.0d0, 1.0d0) * (rkx_new * dble(i) + rky_new * dble(j) + rkz_new * dble(k)))
        end do
      end do
    end do
    if (abs(ene_(1) - ene_(2)) < eps) then
      print *, "Caution!!"
      print *, "E(2) and E(1) are degenerate"
    end if
    allocate(fn(THS_new), alpha(itr_dsf_), beta2(itr_dsf_), pos(3 * NOS_))
    do i = 1, NOS_
      pos(3*i - 2) = Spos(1, i); pos(3*i - 1) = Spos(2, i); pos(3*i) = Spos(3, i)
    end do
    call qs3x_check(qs3x_apply_sq(qs3x_sector_h, qs3x_new_h, int(spsmsz, c_int32_t), kvec, pos, psi_(:, 1), fn), &
      "calcu_DSF_wavevector")
    call qs3x_check(qs3x_dsf_recurrence(qs3x_new_h, qs3x_env_precision(), 1_c_int32_t, fn, &
      int(itr_dsf_, c_int32_t), rfield_, alpha, beta2, norm1), "calcu_DSF_wavevector")
    ex_i = 1
    itr = 1
    sztot = dble(NOS_) / 2.0d0 - dble(NOD_)
    write(ex_index, '(i1)') ex_i - 1
    open(162, file=trim(adjustl(OUTDIR))//'alpha_beta_values_E'//trim(adjustl(ex_index))//'-tmp.d', &
      position='append')
    do j = 1, min(itr_dsf_, THS_new)
      write(162, '(3I6, 8e23.15)') ex_i, j, itr, alpha(j), beta2(j), norm1(1), ene_(ex_i) - rfield_*sztot, sztot
    end do
    close(162)
    write(sztot_index, '(a,i4.4)') 'sztot', int(sztot*10)
    open(162, file=trim(adjustl(OUTDIR))//'alpha_beta_values_'//trim(adjustl(sztot_index))//'_E'// &
      trim(adjustl(ex_index))//'.d', position='append')
    do j = 1, itr_dsf_
      write(162, '(3I6, 8e23.15)') ex_i, j, itr, alpha(j), beta2(j), norm1(1), ene_(ex_i) - rfield_*sztot, sztot
    end do
    close(162)
  end subroutine calcu_DSF_wavevector

end module get_expectation_values
