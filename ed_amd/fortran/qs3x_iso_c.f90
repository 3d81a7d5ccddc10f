! qs3x_iso_c.f90 -- ISO_C_BINDING interfaces to libqs3x (include/qs3x.h).
!
! This is the binding a Fortran driver uses to reach the MI355X engine.  The
! module-level handles hold the model (bonds, lattice) and the sectors the
! shim modules in qs3x_modules.f90 create; complex(8) arrays are passed as the
! interleaved (re, im) doubles the C ABI expects.
module qs3x_iso_c
  use iso_c_binding
  implicit none
  integer(c_int32_t), parameter :: QS3X_FP64 = 0, QS3X_FP32 = 1
  type(c_ptr) :: qs3x_model_h = c_null_ptr   ! model of the current run
  type(c_ptr) :: qs3x_sector_h = c_null_ptr  ! main (N_down, k) sector
  type(c_ptr) :: qs3x_new_h = c_null_ptr     ! DSF target sector

  interface
    function qs3x_last_error() bind(C, name="qs3x_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function qs3x_model_create(NOS, LX, LY, LZ, NOxxz, p_xxz, jxy, jz, out) &
        bind(C, name="qs3x_model_create") result(rc)
      import :: c_int32_t, c_double, c_ptr, c_int
      integer(c_int32_t), value :: NOS, LX, LY, LZ, NOxxz
      integer(c_int32_t), intent(in) :: p_xxz(*)
      real(c_double), intent(in) :: jxy(*), jz(*)
      type(c_ptr), intent(out) :: out
      integer(c_int) :: rc
    end function
    function qs3x_fm_energy(m, e) bind(C, name="qs3x_fm_energy") result(rc)
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: m
      real(c_double), intent(out) :: e
      integer(c_int) :: rc
    end function
    function qs3x_sector_build(m, nod, rkx, rky, rkz, out, dim) bind(C, name="qs3x_sector_build") result(rc)
      import :: c_ptr, c_int32_t, c_double, c_int64_t, c_int
      type(c_ptr), value :: m
      integer(c_int32_t), value :: nod
      real(c_double), value :: rkx, rky, rkz
      type(c_ptr), intent(out) :: out
      integer(c_int64_t), intent(out) :: dim
      integer(c_int) :: rc
    end function
    function qs3x_sector_from_lists(m, nod, dim, list_s, list_r, explist, out) &
        bind(C, name="qs3x_sector_from_lists") result(rc)
      import :: c_ptr, c_int32_t, c_int64_t, c_double, c_int
      type(c_ptr), value :: m
      integer(c_int32_t), value :: nod
      integer(c_int64_t), value :: dim
      integer(c_int64_t), intent(in) :: list_s(*)
      real(c_double), intent(in) :: list_r(*)
      complex(c_double_complex), intent(in) :: explist(*)
      type(c_ptr), intent(out) :: out
      integer(c_int) :: rc
    end function
    function qs3x_sector_destroy(s) bind(C, name="qs3x_sector_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: s
      integer(c_int) :: rc
    end function
    function qs3x_sector_lists(s, list_s, list_r) bind(C, name="qs3x_sector_lists") result(rc)
      import :: c_ptr, c_int64_t, c_double, c_int
      type(c_ptr), value :: s
      integer(c_int64_t), intent(out) :: list_s(*)
      real(c_double), intent(out) :: list_r(*)
      integer(c_int) :: rc
    end function
    function qs3x_ham_to_vec_wave_vector(s, v0, v1) bind(C, name="qs3x_ham_to_vec_wave_vector") result(rc)
      import :: c_ptr, c_double_complex, c_int
      type(c_ptr), value :: s
      complex(c_double_complex), intent(out) :: v0(*)
      complex(c_double_complex), intent(in) :: v1(*)
      integer(c_int) :: rc
    end function
    function qs3x_make_full_hamiltonian(s, ham, ld) bind(C, name="qs3x_make_full_hamiltonian") result(rc)
      import :: c_ptr, c_double_complex, c_int64_t, c_int
      type(c_ptr), value :: s
      complex(c_double_complex), intent(out) :: ham(*)
      integer(c_int64_t), value :: ld
      integer(c_int) :: rc
    end function
    function qs3x_full_diag(s, ene, psi, ld) bind(C, name="qs3x_full_diag") result(rc)
      import :: c_ptr, c_double, c_double_complex, c_int64_t, c_int
      type(c_ptr), value :: s
      real(c_double), intent(out) :: ene(*)
      complex(c_double_complex), intent(out) :: psi(*)
      integer(c_int64_t), value :: ld
      integer(c_int) :: rc
    end function
    function qs3x_lanczos(s, prec, maxitr, minitr, itrint, conv, start, seed, ene, psi, iters) &
        bind(C, name="qs3x_lanczos") result(rc)
      import :: c_ptr, c_int32_t, c_int64_t, c_double, c_double_complex, c_int
      type(c_ptr), value :: s
      integer(c_int32_t), value :: prec
      integer(c_int64_t), value :: maxitr, minitr, itrint, seed
      real(c_double), value :: conv
      type(c_ptr), value :: start
      real(c_double), intent(out) :: ene(2)
      complex(c_double_complex), intent(out) :: psi(*)
      integer(c_int64_t), intent(out) :: iters
      integer(c_int) :: rc
    end function
    function qs3x_trlanczos(s, prec, NOE, NOK, NOM, maxitr, conv, start, seed, ene, psi, stats) &
        bind(C, name="qs3x_trlanczos") result(rc)
      import :: c_ptr, c_int32_t, c_int64_t, c_double, c_double_complex, c_int
      type(c_ptr), value :: s
      integer(c_int32_t), value :: prec, NOE, NOK, NOM
      integer(c_int64_t), value :: maxitr, seed
      real(c_double), value :: conv
      type(c_ptr), value :: start
      real(c_double), intent(out) :: ene(*)
      complex(c_double_complex), intent(out) :: psi(*)
      real(c_double), intent(out) :: stats(8)
      integer(c_int) :: rc
    end function
    function qs3x_correlations(s, psi, npairs, pairs, szsz, spsm) bind(C, name="qs3x_correlations") result(rc)
      import :: c_ptr, c_int64_t, c_int32_t, c_double, c_double_complex, c_int
      type(c_ptr), value :: s
      complex(c_double_complex), intent(in) :: psi(*)
      integer(c_int64_t), value :: npairs
      integer(c_int32_t), intent(in) :: pairs(*)
      real(c_double), intent(out) :: szsz(*)
      complex(c_double_complex), intent(out) :: spsm(*)
      integer(c_int) :: rc
    end function
    function qs3x_local_mag(s, psi, expe_sz) bind(C, name="qs3x_local_mag") result(rc)
      import :: c_ptr, c_double, c_double_complex, c_int
      type(c_ptr), value :: s
      complex(c_double_complex), intent(in) :: psi(*)
      real(c_double), intent(out) :: expe_sz(*)
      integer(c_int) :: rc
    end function
    function qs3x_apply_sq(gs, tgt, spsmsz, q, pos, psi, f) bind(C, name="qs3x_apply_sq") result(rc)
      import :: c_ptr, c_int32_t, c_double, c_double_complex, c_int
      type(c_ptr), value :: gs, tgt
      integer(c_int32_t), value :: spsmsz
      real(c_double), intent(in) :: q(3), pos(*)
      complex(c_double_complex), intent(in) :: psi(*)
      complex(c_double_complex), intent(out) :: f(*)
      integer(c_int) :: rc
    end function
    function qs3x_dsf_recurrence(tgt, prec, nvec, f, itr_dsf, rfield, alpha, beta2, norm1) &
        bind(C, name="qs3x_dsf_recurrence") result(rc)
      import :: c_ptr, c_int32_t, c_double, c_double_complex, c_int
      type(c_ptr), value :: tgt
      integer(c_int32_t), value :: prec, nvec, itr_dsf
      complex(c_double_complex), intent(in) :: f(*)
      real(c_double), value :: rfield
      real(c_double), intent(out) :: alpha(*), beta2(*), norm1(*)
      integer(c_int) :: rc
    end function
    function qs3x_continued_fraction(nlev, alpha, beta2, norm1, E0, eta, nomega, omega, quirk, dsf) &
        bind(C, name="qs3x_continued_fraction") result(rc)
      import :: c_int32_t, c_int64_t, c_double, c_int
      integer(c_int32_t), value :: nlev, quirk
      real(c_double), intent(in) :: alpha(*), beta2(*), omega(*)
      real(c_double), value :: norm1, E0, eta
      integer(c_int64_t), value :: nomega
      real(c_double), intent(out) :: dsf(*)
      integer(c_int) :: rc
    end function
  end interface

contains

  ! Non-zero status -> stop with the library's message (the reference stops on
  ! every error condition).
  subroutine qs3x_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(qs3x_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
      if (msg(n+1) == c_null_char) exit
      n = n + 1
    end do
    write(*,'(a,": ",1024a)') where, msg(1:n)
    error stop 1
  end subroutine qs3x_check

  ! Vector precision of the solvers: QS3X_PRECISION=fp32 selects the
  ! FP32-mixed path, anything else FP64.
  integer(c_int32_t) function qs3x_env_precision() result(p)
    character(len=16) :: v
    integer :: l, st
    p = QS3X_FP64
    call get_environment_variable("QS3X_PRECISION", v, l, st)
    if (st == 0 .and. l >= 4) then
      if (v(1:4) == "fp32" .or. v(1:4) == "FP32") p = QS3X_FP32
    end if
  end function qs3x_env_precision

end module qs3x_iso_c
