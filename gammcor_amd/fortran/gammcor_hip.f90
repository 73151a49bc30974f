! Fortran interfaces of the gammcor_hip C ABI (include/gammcor_hip.h), for
! Fortran callers: integer(c_int) status returns, 0 = success.
module gammcor_hip
  use iso_c_binding
  implicit none
  private
  public :: GMC_FP64, GMC_FP32, GMC_MIXED, GMC_MIXED2
  public :: gmc_init, gmc_finalize, gmc_twono1, gmc_ab_cas, gmc_erpasymm1, gmc_aceneerpa, gmc_aceinteg, &
            gmc_aceinteg_fofo, gmc_acecorr, gmc_ac0cas, gmc_ac0cas_fofo, gmc_forget_resident, gmc_check, gmc_prec_from_env

  integer(c_int), parameter :: GMC_FP64 = 0, GMC_FP32 = 1, GMC_MIXED = 2, GMC_MIXED2 = 3

  interface
    integer(c_int) function gmc_init() bind(C, name="gmc_init")
      import :: c_int
    end function

    integer(c_int) function gmc_finalize() bind(C, name="gmc_finalize")
      import :: c_int
    end function

    type(c_ptr) function gmc_last_error() bind(C, name="gmc_last_error")
      import :: c_ptr
    end function

    integer(c_int) function gmc_forget_resident() bind(C, name="gmc_forget_resident")
      import :: c_int
    end function

    integer(c_int) function gmc_twono1(tno, ure, nbasis, ninte2) bind(C, name="gmc_twono1")
      import :: c_int, c_int64_t, c_double
      real(c_double) :: tno(*)
      real(c_double), intent(in) :: ure(*)
      integer(c_int), value :: nbasis
      integer(c_int64_t), value :: ninte2
    end function

    integer(c_int) function gmc_ab_cas(abplus, abmin, etot, ure, occ, xone, twono, ninte2, indn, ndimx, nbasis, &
                                       ninact, nact, igem, rdm2act, acalpha, cicoef) bind(C, name="gmc_ab_cas")
      import :: c_int, c_int64_t, c_double
      real(c_double) :: abplus(*), abmin(*), etot, cicoef(*)
      real(c_double), intent(in) :: ure(*), occ(*), xone(*), twono(*), rdm2act(*)
      integer(c_int64_t), value :: ninte2
      integer(c_int), intent(in) :: indn(*), igem(*)
      integer(c_int), value :: ndimx, nbasis, ninact, nact
      real(c_double), value :: acalpha
    end function

    integer(c_int) function gmc_erpasymm1(eigvecr, eig, aplsqrt, abmin, ndimx, nneg) bind(C, name="gmc_erpasymm1")
      import :: c_int, c_double
      real(c_double) :: eigvecr(*), eig(*), aplsqrt(*), abmin(*)
      integer(c_int), value :: ndimx
      integer(c_int) :: nneg
    end function

    integer(c_int) function gmc_aceneerpa(ecorr, eigvecr, eig, twono, ninte2, occ, igem, cicoef, indn, nbasis, &
                                          ndimx) bind(C, name="gmc_aceneerpa")
      import :: c_int, c_int64_t, c_double
      real(c_double) :: ecorr
      real(c_double), intent(in) :: eigvecr(*), eig(*), twono(*), occ(*), cicoef(*)
      integer(c_int64_t), value :: ninte2
      integer(c_int), intent(in) :: igem(*), indn(*)
      integer(c_int), value :: nbasis, ndimx
    end function

    integer(c_int) function gmc_aceinteg(prec, ecorr, egone, twono, ninte2, ure, occ, xone, igem, rdm2act, indn, &
                                         ndimx, nbasis, ninact, nact, acalpha, abplus, abmin, eigvecr, eig) &
                                         bind(C, name="gmc_aceinteg")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: prec
      real(c_double) :: ecorr, egone
      real(c_double), intent(in) :: twono(*), ure(*), occ(*), xone(*), rdm2act(*)
      integer(c_int64_t), value :: ninte2
      integer(c_int), intent(in) :: igem(*), indn(*)
      integer(c_int), value :: ndimx, nbasis, ninact, nact
      real(c_double), value :: acalpha
      real(c_double) :: abplus(*), abmin(*), eigvecr(*), eig(*)
    end function

    integer(c_int) function gmc_ac0cas(ecorr, etot, twono, ninte2, ure, occ, xone, igem, rdm2act, indn, ndimx, &
                                       nbasis, ninact, nact) bind(C, name="gmc_ac0cas")
      import :: c_int, c_int64_t, c_double
      real(c_double) :: ecorr, etot
      real(c_double), intent(in) :: twono(*), ure(*), occ(*), xone(*), rdm2act(*)
      integer(c_int64_t), value :: ninte2
      integer(c_int), intent(in) :: igem(*), indn(*)
      integer(c_int), value :: ndimx, nbasis, ninact, nact
    end function

    integer(c_int) function gmc_ac0cas_fofo(ecorr, etot, jints, kints, ure, occ, xone, igem, rdm2act, indn, ndimx, &
                                            nbasis, ninact, nact) bind(C, name="gmc_ac0cas_fofo")
      import :: c_int, c_double
      real(c_double) :: ecorr, etot
      real(c_double), intent(in) :: jints(*), kints(*), ure(*), occ(*), xone(*), rdm2act(*)
      integer(c_int), intent(in) :: igem(*), indn(*)
      integer(c_int), value :: ndimx, nbasis, ninact, nact
    end function

    integer(c_int) function gmc_aceinteg_fofo(prec, ecorr, egone, jints, kints, ure, occ, xone, igem, rdm2act, &
                                              indn, ndimx, nbasis, ninact, nact, acalpha, abplus, abmin, eigvecr, &
                                              eig) bind(C, name="gmc_aceinteg_fofo")
      import :: c_int, c_double
      integer(c_int), value :: prec
      real(c_double) :: ecorr, egone
      real(c_double), intent(in) :: jints(*), kints(*), ure(*), occ(*), xone(*), rdm2act(*)
      integer(c_int), intent(in) :: igem(*), indn(*)
      integer(c_int), value :: ndimx, nbasis, ninact, nact
      real(c_double), value :: acalpha
      real(c_double) :: abplus(*), abmin(*), eigvecr(*), eig(*)
    end function

    integer(c_int) function gmc_acecorr(prec, alpha_batch, nbasis, ninact, nact, twono, ninte2, occ, igem, xone, &
                                        ure, rdm2act, ndimx, indn, nalpha, alphas, weights, ecorr, walpha, ecasscf, &
                                        nneg) bind(C, name="gmc_acecorr")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: prec, alpha_batch, nbasis, ninact, nact, ndimx, nalpha
      real(c_double), intent(in) :: twono(*), occ(*), xone(*), ure(*), rdm2act(*), alphas(*), weights(*)
      integer(c_int64_t), value :: ninte2
      integer(c_int), intent(in) :: igem(*), indn(*)
      real(c_double) :: ecorr, walpha(*), ecasscf
      integer(c_int) :: nneg(*)
    end function
  end interface

contains

  ! A non-zero status ends the run with the library's message, as the
  ! reference ends on its own errors (Stop, misc.f:567).
  subroutine gmc_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(gmc_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write (6, '(3A,I0,2A)') ' gammcor_hip: ', where, ' failed with status ', rc, ': ', transfer(msg(1:n), repeat(' ', n))
    stop 1
  end subroutine

  ! Arithmetic mode from GAMMCOR_HIP_PREC (fp64 | mixed | mixed2 | fp32);
  ! plain FP64 when unset.
  integer(c_int) function gmc_prec_from_env() result(prec)
    character(len=16) :: v
    integer :: l, st
    prec = GMC_FP64
    call get_environment_variable("GAMMCOR_HIP_PREC", v, l, st)
    if (st /= 0) return
    select case (trim(v))
    case ("mixed", "MIXED")
      prec = GMC_MIXED
    case ("mixed2", "MIXED2")
      prec = GMC_MIXED2
    case ("fp32", "FP32")
      prec = GMC_FP32
    end select
  end function

end module gammcor_hip
