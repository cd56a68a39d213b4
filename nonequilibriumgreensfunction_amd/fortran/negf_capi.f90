! iso_c_binding interfaces of the C ABI in include/negf_capi.h.
!
! Complex arrays are COMPLEX(c_double_complex), real arrays REAL(c_double); all
! are passed by reference as assumed-size dummies, so module arrays need no
! TARGET attribute.  Every function returns the library's int status.
module negf_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: NEGF_OK = 0, NEGF_ERR_ARG = -1, NEGF_ERR_HIP = -2, &
      NEGF_ERR_NO_DEVICE = -3, NEGF_ERR_MEMORY = -4, NEGF_ERR_STATE = -5
  integer(c_int), parameter :: NEGF_STATUS_SR_NOT_CONVERGED = 1048576, NEGF_STATUS_SR_NAN = 2097152
  integer(c_int), parameter :: NEGF_PREC_FP64 = 0, NEGF_PREC_MIXED = 1, NEGF_PREC_REFINED = 2, &
                                NEGF_PREC_LU32 = 3
  integer(c_int), parameter :: NEGF_FLAG_RGF_THREE_PASS = 1, NEGF_FLAG_SR_SEPARATE_LEADS = 2, &
      NEGF_FLAG_SINGLE_LANE = 4, NEGF_FLAG_CACHE_CONTACTS = 8, NEGF_FLAG_GRAPH = 16, &
      NEGF_FLAG_LU32_FP32_MFMA = 32, NEGF_FLAG_SR_NO_MIRROR = 64

  type, bind(C) :: negf_config
    integer(c_int) :: npx, npy, npz, norb
    real(c_double) :: eta
    real(c_double) :: sr_tol
    integer(c_int) :: sr_max_iter
    integer(c_int) :: sr_extra_eta
    integer(c_int) :: precision
    integer(c_int) :: device
    real(c_double) :: hbm_budget_gb
    integer(c_int) :: max_batch
    integer(c_int) :: flags
    integer(c_int) :: lanes
    real(c_double) :: gamma_trunc
  end type negf_config

  interface
    function negf_init(ctx, cfg) result(rc) bind(C, name="negf_init")
      import :: c_ptr, c_int, negf_config
      type(c_ptr), intent(out) :: ctx
      type(negf_config), intent(in) :: cfg
      integer(c_int) :: rc
    end function

    function negf_finalize(ctx) result(rc) bind(C, name="negf_finalize")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function negf_set_hamiltonian(ctx, Htotd, HopX, gM) result(rc) bind(C, name="negf_set_hamiltonian")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: Htotd(*), HopX(*), gM(*)
      integer(c_int) :: rc
    end function

    function negf_energy_sweep_f(ctx, energy, eCtr, deltaE, mu, muT, mask, T12, Modec1, Modec2, &
        DOSc1, DOSc2, EDensity, HDensity, RhoN, RhoP, SigX, SigY, SigZ, ContactCurrent, Jx, Jy, Jz, &
        ERJx, ERJy, ERJz, ERRhon, ORJx, ORJy, ORJz, ORRhon) &
        result(rc) bind(C, name="negf_energy_sweep_f")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: energy, deltaE, muT
      integer(c_int), value :: eCtr, mask
      real(c_double), intent(in) :: mu(2)
      ! optional: an absent (or unallocated allocatable) actual argument passes NULL,
      ! which the library treats as an output not wanted
      real(c_double), optional, intent(inout) :: T12(*), Modec1(*), Modec2(*), DOSc1(*), DOSc2(*), &
          EDensity(*), HDensity(*), RhoN(*), RhoP(*), SigX(*), SigY(*), SigZ(*), ContactCurrent(*), &
          Jx(*), Jy(*), Jz(*), ERJx(*), ERJy(*), ERJz(*), ERRhon(*), ORJx(*), ORJy(*), ORJz(*), &
          ORRhon(*)
      integer(c_int) :: rc
    end function

    function negf_submit(ctx, count, energies, eCtr, deltaE) result(rc) bind(C, name="negf_submit")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: count
      real(c_double), intent(in) :: energies(*), deltaE(*)
      integer(c_int), intent(in) :: eCtr(*)
      integer(c_int) :: rc
    end function

    function negf_flush_f(ctx, mu, muT, mask, T12, Modec1, Modec2, DOSc1, DOSc2, EDensity, &
        HDensity, RhoN, RhoP, SigX, SigY, SigZ, ContactCurrent, Jx, Jy, Jz, ERJx, ERJy, ERJz, &
        ERRhon, ORJx, ORJy, ORJz, ORRhon) result(rc) bind(C, name="negf_flush_f")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: muT
      integer(c_int), value :: mask
      real(c_double), intent(in) :: mu(2)
      ! optional: an absent (or unallocated allocatable) actual argument passes NULL,
      ! which the library treats as an output not wanted
      real(c_double), optional, intent(inout) :: T12(*), Modec1(*), Modec2(*), DOSc1(*), DOSc2(*), &
          EDensity(*), HDensity(*), RhoN(*), RhoP(*), SigX(*), SigY(*), SigZ(*), ContactCurrent(*), &
          Jx(*), Jy(*), Jz(*), ERJx(*), ERJy(*), ERJz(*), ERRhon(*), ORJx(*), ORJy(*), ORJz(*), &
          ORRhon(*)
      integer(c_int) :: rc
    end function

    function negf_contact_self_energy(ctx, energy, SigmaL, SigmaR, GammaL, GammaR, dos, iters) &
        result(rc) bind(C, name="negf_contact_self_energy")
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: energy
      complex(c_double_complex), intent(out) :: SigmaL(*), SigmaR(*), GammaL(*), GammaR(*)
      real(c_double), intent(out) :: dos(2)
      integer(c_int), intent(out) :: iters(2)
      integer(c_int) :: rc
    end function

    function negf_sancho_rubio(ctx, alpha, beta, Sigma, iters) result(rc) bind(C, name="negf_sancho_rubio")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: alpha(*), beta(*)
      complex(c_double_complex), intent(out) :: Sigma(*)
      integer(c_int), intent(out) :: iters
      integer(c_int) :: rc
    end function

    function negf_retarded(ctx, InvGRd, GRd, GRu, GRl, GRLef, GRRgt) result(rc) bind(C, name="negf_retarded")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: InvGRd(*)
      complex(c_double_complex), intent(out) :: GRd(*), GRu(*), GRl(*), GRLef(*), GRRgt(*)
      integer(c_int) :: rc
    end function

    function negf_gnp(ctx, f1, f2, deltaE, GRLef, GRRgt, GammaL, GammaR, Gnd, Gnu, Gnl, Gpd, Gpu, &
        Gpl, RhoN, RhoP) result(rc) bind(C, name="negf_gnp")
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: f1, f2, deltaE
      complex(c_double_complex), intent(in) :: GRLef(*), GRRgt(*), GammaL(*), GammaR(*)
      complex(c_double_complex), intent(out) :: Gnd(*), Gnu(*), Gnl(*), Gpd(*), Gpu(*), Gpl(*)
      real(c_double), intent(inout) :: RhoN(*), RhoP(*)
      integer(c_int) :: rc
    end function

    function negf_observables_f(ctx, eCtr, f1, f2, deltaE, Gnd, Gpd, Gnu, Gnl, Gpu, Gpl, GRRgt, &
        GammaL, GammaR, mask, T12, Modec1, Modec2, DOSc1, DOSc2, EDensity, HDensity, RhoN, RhoP, &
        SigX, SigY, SigZ, ContactCurrent, Jx, Jy, Jz) result(rc) bind(C, name="negf_observables_f")
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: eCtr, mask
      real(c_double), value :: f1, f2, deltaE
      complex(c_double_complex), intent(in) :: Gnd(*), Gpd(*), Gnu(*), Gnl(*), Gpu(*), Gpl(*), &
          GRRgt(*), GammaL(*), GammaR(*)
      real(c_double), intent(inout) :: T12(*), Modec1(*), Modec2(*), DOSc1(*), DOSc2(*), &
          EDensity(*), HDensity(*), RhoN(*), RhoP(*), SigX(*), SigY(*), SigZ(*), ContactCurrent(*), &
          Jx(*), Jy(*), Jz(*)
      integer(c_int) :: rc
    end function

    function negf_invert_matrix(n, G, invG, info) result(rc) bind(C, name="negf_invert_matrix")
      import :: c_int, c_double_complex
      integer(c_int), value :: n
      complex(c_double_complex), intent(in) :: G(*)
      complex(c_double_complex), intent(out) :: invG(*)
      integer(c_int), intent(out) :: info
      integer(c_int) :: rc
    end function

    ! The lead mirror the paired decimation uses (axis 0: none, 1: y, 2: z); paired = 1
    ! when one sector's decimation serves both.  No counterpart in the reference.
    function negf_contact_mirror(ctx, axis, perm, phase, paired) result(rc) bind(C, name="negf_contact_mirror")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: axis, perm(*), phase(*), paired
      integer(c_int) :: rc
    end function

    function negf_find_mirror(npy, npz, norb, Hend, HopX, axis, perm, phase, paired) result(rc) &
        bind(C, name="negf_find_mirror")
      import :: c_int, c_double_complex
      integer(c_int), value :: npy, npz, norb
      complex(c_double_complex), intent(in) :: Hend(*), HopX(*)
      integer(c_int), intent(out) :: axis, perm(*), phase(*), paired
      integer(c_int) :: rc
    end function

    function negf_last_error() result(p) bind(C, name="negf_last_error")
      import :: c_ptr
      type(c_ptr) :: p
    end function
  end interface

contains

  ! The library's last error message as a Fortran string.
  function negf_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = negf_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function negf_error_message

end module negf_capi
