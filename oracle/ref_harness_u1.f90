! ref_harness_u1.f90 -- C-callable entry points into the reference's own
! U(1)-only modules (source_only_u1/), for the CPU parity tests of the
! U(1)-only H.v.  Test infrastructure only: built into oracle/_ref/ by
! build_ref.sh and never shipped.
!
! ref_u1_setup mirrors read_ip (source_only_u1/input_param.f90:49-91: the
! p1 < p2 swap, sJint = (Jxy/2, Jz/4)) and main.f90:23-29 (get_combination,
! THS = combination(NOS, NOD)); ref_u1_hv calls ham_to_vec_wave_vector
! (source_only_u1/ham2vec_v3.f90:35-65) unchanged.
module ref_harness_u1
  use iso_c_binding
  use input_param, only: NOS, NOD, NOxxz, p_xxz, Jint, sJint, THS
  use state_lists, only: get_combination, combination
  use ham2vec, only: ham_to_vec_wave_vector
  implicit none
contains
  subroutine ref_u1_setup(nos_, nod_, noxxz_, pi_, pj_, jxy_, jz_, ths_out) bind(C, name="ref_u1_setup")
    integer(c_int64_t), value :: nos_, nod_, noxxz_
    integer(c_int64_t), intent(in) :: pi_(noxxz_), pj_(noxxz_)
    real(c_double), intent(in) :: jxy_(noxxz_), jz_(noxxz_)
    integer(c_int64_t), intent(out) :: ths_out
    integer :: i, tmp
    NOS = nos_; NOD = nod_; NOxxz = noxxz_
    if (allocated(p_xxz)) deallocate(p_xxz, Jint, sJint)
    allocate(p_xxz(2,NOxxz), Jint(NOxxz,2:3), sJint(NOxxz,2:3))
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
    call get_combination(NOS, NOD + 1)
    THS = combination(NOS, NOD)
    ths_out = THS
  end subroutine ref_u1_setup

  subroutine ref_u1_hv(v0, v1) bind(C, name="ref_u1_hv")
    complex(c_double_complex), intent(out) :: v0(THS)
    complex(c_double_complex), intent(in) :: v1(THS)
    call ham_to_vec_wave_vector(v0, v1, THS, NOD)
  end subroutine ref_u1_hv
end module ref_harness_u1
