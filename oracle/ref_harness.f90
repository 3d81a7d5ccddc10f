! ref_harness.f90 -- C-callable harness around the reference's own modules.
!
! TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh compiles the reference's
! unmodified input_param.f90, my_l1_blas.f90, state_lists.f90, ham2vec_v3.f90
! and get_expectation_values.f90 (source/ of QS-Cube/ED) together with this
! file into oracle/_ref/libqs3ref.so.  The harness only sets the reference's
! public module variables the way read_ip (input_param.f90:58-117) and main
! (main.f90:24-54) do, then calls the reference routines unchanged.  Built
! with -fdefault-integer-8, so every default integer is 64-bit.
module qs3ref_harness
  use iso_c_binding
  use input_param
  use state_lists
  use ham2vec
  use get_expectation_values, only: calcu_Sq_a, mk_list_ij, mk_list_n, Spos
  implicit none
contains

  ! Model set-up mirroring read_ip (input_param.f90:75-103) and main.f90:25,49-51.
  subroutine ref_setup(nos_, lx_, ly_, lz_, nod_, kx_, ky_, kz_, noxxz_, pi_, pj_, jxy_, jz_) &
      bind(C, name="ref_setup")
    integer(c_int64_t), value :: nos_, lx_, ly_, lz_, nod_, kx_, ky_, kz_, noxxz_
    integer(c_int64_t), intent(in) :: pi_(noxxz_), pj_(noxxz_)
    real(c_double), intent(in) :: jxy_(noxxz_), jz_(noxxz_)
    integer :: i, j, k, tmp
    NOS = nos_; LX = lx_; LY = ly_; LZ = lz_; NOD = nod_
    KX = kx_; KY = ky_; KZ = kz_; NOxxz = noxxz_
    PI = acos(-1.0d0)
    rkx = 2.0d0 * PI * dble(KX)/dble(LX)
    rky = 2.0d0 * PI * dble(KY)/dble(LY)
    rkz = 2.0d0 * PI * dble(KZ)/dble(LZ)
    if (allocated(p_xxz)) deallocate(p_xxz, Jint, sJint, shift_x_SQ, shift_y_SQ, shift_z_SQ, explist)
    allocate(p_xxz(2,NOxxz),Jint(NOxxz,2:3),sJint(NOxxz,2:3),shift_x_SQ(NOS),shift_y_SQ(NOS),&
      shift_z_SQ(NOS),explist(0:LX,0:LY,0:LZ))
    forall(i=0:LX,j=0:LY,k=0:LZ) explist(i,j,k) = &
      exp((0.0d0,1.0d0)*(rkx*dble(i)+rky*dble(j)+rkz*dble(k)))
    do i = 1, NOxxz
      p_xxz(1,i) = pi_(i); p_xxz(2,i) = pj_(i)
      Jint(i,2) = jxy_(i); Jint(i,3) = jz_(i)
      sJint(i,2) = Jint(i,2) * 0.5d0
      sJint(i,3) = Jint(i,3) * 0.25d0
      if (p_xxz(1,i) > p_xxz(2,i)) then
        tmp = p_xxz(1,i); p_xxz(1,i) = p_xxz(2,i); p_xxz(2,i) = tmp
      end if
    end do
    if (allocated(combination)) deallocate(combination)
    call get_combination(NOS, NOD+1)
    shift_x_SQ = mk_shift_x_SQ(LX,LY,LZ)
    shift_y_SQ = mk_shift_y_SQ(LX,LY,LZ)
    shift_z_SQ = mk_shift_z_SQ(LX,LY,LZ)
  end subroutine ref_setup

  ! main.f90:28-33 then allocate_lists_omp_SQ (main.f90:54).
  subroutine ref_build_sector(ths_out) bind(C, name="ref_build_sector")
    integer(c_int64_t), intent(out) :: ths_out
    if (allocated(list_s)) deallocate(list_s)
    if (allocated(list_r)) deallocate(list_r)
    if (NOD > 0) then
      THS = combination(NOS, NOD)
    else
      THS = 1
    end if
    call allocate_lists_omp_SQ
    ths_out = THS
  end subroutine ref_build_sector

  subroutine ref_get_lists(ls, lr) bind(C, name="ref_get_lists")
    integer(c_int64_t), intent(out) :: ls(THS)
    real(c_double), intent(out) :: lr(THS)
    ls = list_s(1:THS); lr = list_r(1:THS)
  end subroutine ref_get_lists

  ! One H.v with the module's own sector (ham2vec_v3.f90:37-72).
  subroutine ref_hv(v0, v1) bind(C, name="ref_hv")
    complex(c_double_complex), intent(out) :: v0(THS)
    complex(c_double_complex), intent(in) :: v1(THS)
    call ham_to_vec_wave_vector(v0, v1, THS, NOD, list_s, list_r, explist)
  end subroutine ref_hv

  ! Lists built elsewhere (the reference's own allocate_lists_omp_SQ takes hours
  ! at L = 1000): THS rows of list_s / list_r into the module globals.
  subroutine ref_set_lists(ths_, ls, lr) bind(C, name="ref_set_lists")
    integer(c_int64_t), value :: ths_
    integer(c_int64_t), intent(in) :: ls(ths_)
    real(c_double), intent(in) :: lr(ths_)
    if (allocated(list_s)) deallocate(list_s)
    if (allocated(list_r)) deallocate(list_r)
    THS = ths_
    allocate(list_s(THS), list_r(THS))
    list_s = ls
    list_r = lr
  end subroutine ref_set_lists

  ! Rows i0..i1 (1-based) of ham_to_vec_wave_vector (ham2vec_v3.f90:52-70): the
  ! same loop body over a window of rows, calling the reference's list_fly,
  ! j_flip_ni, representative_SQ and findstate against the full list, with the
  ! rows spread over OpenMP threads as the reference spreads them.  For timing
  ! the reference on a bounded sample of a large sector.
  subroutine ref_hv_rows(i0, i1, v0, v1) bind(C, name="ref_hv_rows")
    integer(c_int64_t), value :: i0, i1
    complex(c_double_complex), intent(inout) :: v0(THS)
    complex(c_double_complex), intent(in) :: v1(THS)
    integer :: i, j, ell(3), s, id
    integer, allocatable :: ni(:), st_list(:)
    logical :: f1, f2
    allocate(ni(NOD), st_list(NOD))
    !$omp parallel do private(f1,f2,s,ni,id,ell,i,j,st_list)
    do i = i0, i1
      v0(i) = (0.0d0, 0.0d0)
      st_list = list_fly(list_s(i), NOD, NOS)
      do j = 1, NOxxz
        f1 = any(st_list == p_xxz(1,j))
        f2 = any(st_list == p_xxz(2,j))
        if (f1 .neqv. f2) then
          v0(i) = v0(i) - sJint(j,3) * v1(i)
          ni = j_flip_ni(p_xxz(1,j), p_xxz(2,j), st_list, NOD)
          call representative_SQ(NOD, s, ell, ni)
          call findstate(s, list_s, id, THS)
          if (id > 0) v0(i) = v0(i) + sJint(j,2) * list_r(id) / list_r(i) &
               * explist(ell(1),ell(2),ell(3)) * v1(id)
        else
          v0(i) = v0(i) + sJint(j,3) * v1(i)
        end if
      end do
    end do
  end subroutine ref_hv_rows

  subroutine ref_explist(ex) bind(C, name="ref_explist")
    complex(c_double_complex), intent(out) :: ex((LX+1)*(LY+1)*(LZ+1))
    ex = reshape(explist, [(LX+1)*(LY+1)*(LZ+1)])
  end subroutine ref_explist

  subroutine ref_representative(nod_, n, t, ell) bind(C, name="ref_representative")
    integer(c_int64_t), value :: nod_
    integer(c_int64_t), intent(inout) :: n(nod_)
    integer(c_int64_t), intent(out) :: t, ell(3)
    integer :: nn(nod_), e(3), tt
    nn = n
    call representative_SQ(nod_, tt, e, nn)
    n = nn; t = tt; ell = e
  end subroutine ref_representative

  subroutine ref_checkstate(s, nod_, ro) bind(C, name="ref_checkstate")
    integer(c_int64_t), value :: s, nod_
    real(c_double), intent(out) :: ro
    call checkstate_SQ(s, ro, nod_, rkx, rky, rkz)
  end subroutine ref_checkstate

  subroutine ref_list_fly(t, nod_, n) bind(C, name="ref_list_fly")
    integer(c_int64_t), value :: t, nod_
    integer(c_int64_t), intent(out) :: n(nod_)
    n = list_fly(t, nod_, NOS)
  end subroutine ref_list_fly

  subroutine ref_j_flip(i, j, nod_, n, nd) bind(C, name="ref_j_flip")
    integer(c_int64_t), value :: i, j, nod_
    integer(c_int64_t), intent(in) :: n(nod_)
    integer(c_int64_t), intent(out) :: nd(nod_)
    nd = j_flip_ni(i, j, n, nod_)
  end subroutine ref_j_flip

  ! calcu_cf_trans_2 (ham2vec_v3.f90:111-174) for the orbit of one pair, built
  ! by mk_list_ij (get_expectation_values.f90:97-131).  Returns noij = 0 when
  ! the pair is not the orbit's first element.
  subroutine ref_cf_pair(r1, r2, v, noij_out, szsz, spsm) bind(C, name="ref_cf_pair")
    integer(c_int64_t), value :: r1, r2
    complex(c_double_complex), intent(in) :: v(THS)
    integer(c_int64_t), intent(out) :: noij_out
    real(c_double), intent(out) :: szsz
    complex(c_double_complex), intent(out) :: spsm
    integer :: list_ij(2,NOS), noij, r(2)
    real(8), allocatable :: ez(:)
    complex(8), allocatable :: es(:)
    allocate(ez(NOS**2), es(NOS**2))
    ez = 0.0d0; es = (0.0d0, 0.0d0)
    r(1) = r1; r(2) = r2
    call mk_list_ij(list_ij, noij, NOS, r)
    noij_out = noij
    szsz = 0.0d0; spsm = (0.0d0, 0.0d0)
    if (noij == 0) return
    call calcu_cf_trans_2(v, THS, NOD, list_s, list_r, explist, NOS, list_ij, noij, ez, es)
    szsz = ez(list_ij(1,1) + (list_ij(2,1)-1)*NOS)
    spsm = es(list_ij(1,1) + (list_ij(2,1)-1)*NOS)
  end subroutine ref_cf_pair

  ! calcu_lm_trans_2 (ham2vec_v3.f90:74-109) for the orbit of site n (mk_list_n).
  subroutine ref_lm_site(n, v, non_out, val) bind(C, name="ref_lm_site")
    integer(c_int64_t), value :: n
    complex(c_double_complex), intent(in) :: v(THS)
    integer(c_int64_t), intent(out) :: non_out
    real(c_double), intent(out) :: val
    integer :: list_n(NOS), non
    real(8) :: ez(NOS)
    ez = 0.0d0
    call mk_list_n(list_n, non, NOS, n)
    non_out = non
    val = 0.0d0
    if (non == 0) return
    call calcu_lm_trans_2(v, THS, NOD, ez, NOS, list_n, non)
    val = ez(list_n(1))
  end subroutine ref_lm_site

  ! DSF target sector (calcu_DSF_wavevector, get_expectation_values.f90:430-454):
  ! NOD_new for spsmsz, rk_new = rk + q, allocate_lists_omp_SQ_new, explist_new.
  subroutine ref_build_new_sector(spsmsz_, qx_, qy_, qz_, ths_new_out) bind(C, name="ref_build_new_sector")
    integer(c_int64_t), value :: spsmsz_
    real(c_double), value :: qx_, qy_, qz_
    integer(c_int64_t), intent(out) :: ths_new_out
    integer :: i, j, k
    spsmsz = spsmsz_; qx = qx_; qy = qy_; qz = qz_
    rkx_new = rkx + qx
    rky_new = rky + qy
    rkz_new = rkz + qz
    if (spsmsz == 1) then
      NOD_new = NOD - 1
    else if (spsmsz == 2) then
      NOD_new = NOD + 1
    else
      NOD_new = NOD
    end if
    if (allocated(list_s_new)) deallocate(list_s_new)
    if (allocated(list_r_new)) deallocate(list_r_new)
    if (allocated(explist_new)) deallocate(explist_new)
    THS_new = combination(NOS, NOD_new)
    call allocate_lists_omp_SQ_new
    allocate(explist_new(0:LX,0:LY,0:LZ))
    forall(i=0:LX,j=0:LY,k=0:LZ) explist_new(i,j,k) = &
      exp((0.0d0,1.0d0)*(rkx_new*dble(i)+rky_new*dble(j)+rkz_new*dble(k)))
    ths_new_out = THS_new
  end subroutine ref_build_new_sector

  subroutine ref_get_new_lists(ls, lr) bind(C, name="ref_get_new_lists")
    integer(c_int64_t), intent(out) :: ls(THS_new)
    real(c_double), intent(out) :: lr(THS_new)
    ls = list_s_new(1:THS_new); lr = list_r_new(1:THS_new)
  end subroutine ref_get_new_lists

  ! calcu_Sq_a (get_expectation_values.f90:281-377) into the target sector.
  subroutine ref_sq(pos, psi, fn) bind(C, name="ref_sq")
    real(c_double), intent(in) :: pos(3, NOS)
    complex(c_double_complex), intent(in) :: psi(THS)
    complex(c_double_complex), intent(out) :: fn(THS_new)
    real(8) :: kv(3)
    if (allocated(Spos)) deallocate(Spos)
    allocate(Spos(3, NOS))
    Spos = pos
    kv(1) = qx; kv(2) = qy; kv(3) = qz
    call calcu_Sq_a(spsmsz, psi, fn, kv, rkx_new, rky_new, rkz_new, THS, THS_new, Spos)
  end subroutine ref_sq

  ! H.v in the target sector (get_expectation_values.f90:551-552).
  subroutine ref_hv_new(v0, v1) bind(C, name="ref_hv_new")
    complex(c_double_complex), intent(out) :: v0(THS_new)
    complex(c_double_complex), intent(in) :: v1(THS_new)
    call ham_to_vec_wave_vector(v0, v1, THS_new, NOD_new, list_s_new, list_r_new, explist_new)
  end subroutine ref_hv_new

end module qs3ref_harness
