! Drop-in replacements for the reference's hot-path subroutines, with the same
! names and (argument-less, module-state) interfaces:
!   Energy_Sweep          (Energy_Sweep.F90:1-46)
!   Contact_Self_Energy   (Contact_Self_Energy.F90:1-106, semi-infinite contacts)
!   Sancho_Rubio          (Contact_Self_Energy.F90:110-223)
!   Get_Retarded_Greens   (Recursive_Greens.F90:2-107)
!   Calculate_Gnp         (Calculate_Gnp.F90:1-136)
!   Calculate_Observables (Calculate_Observables.F90:1-169, non-BdG branches)
!   Invert_Matrix         (MKL_Routines_Cluster.F90:2-33)
! Each forwards to the gfx950 library through the C ABI (negf_capi).  Built
! standalone they USE negf_state; built inside the reference with
! -DNEGF_USE_REFERENCE_MODULE they USE the reference's own SIP_Module (which
! re-exports NEGF_Module), replacing Energy_Sweep.F90, Contact_Self_Energy.F90,
! Recursive_Greens.F90, Calculate_Gnp.F90, Calculate_Observables.F90 and the
! Invert_Matrix of MKL_Routines_Cluster.F90 in that build.
!
! Energy_Sweep runs the fused device path (module arrays InvGRd, GR*, Gn*/Gp*,
! Sigma*, Gamma* are not formed; the accumulated outputs are).  Setting
! negf_shim_staged = .true. runs the per-stage path instead, which fills every
! module array exactly as the reference does (and the bond currents Jx/Jy/Jz).
! Energy_Sweep_Submit + NEGF_Flush is the batched form of the energy loop.
#ifdef NEGF_USE_REFERENCE_MODULE
#define NEGF_STATE_MODULE SIP_Module
#else
#define NEGF_STATE_MODULE negf_state
#endif

module negf_shim_ctl
  use iso_c_binding
  use negf_capi
  implicit none
  type(c_ptr), save :: negf_shim_ctx = c_null_ptr
  integer, save :: negf_shim_h_ite = -huge(1)     ! iteration whose Hamiltonian is on the device
  ! solver options; the defaults reproduce the reference (Contact_Self_Energy.F90:141,154)
  real(c_double), save :: negf_shim_sr_tol = 1d-8
  integer(c_int), save :: negf_shim_sr_max_iter = 20000
  integer(c_int), save :: negf_shim_sr_extra_eta = 1
  integer(c_int), save :: negf_shim_precision = NEGF_PREC_FP64
  integer(c_int), save :: negf_shim_device = 0
  integer(c_int), save :: negf_shim_flags = 0      ! NEGF_FLAG_* algorithm selection
  real(c_double), save :: negf_shim_gamma_trunc = 0d0  ! negf_config%gamma_trunc (0: whole Gamma)
  logical, save :: negf_shim_staged = .false.
contains
  subroutine negf_shim_fail(where, rc)
    character(*), intent(in) :: where
    integer(c_int), intent(in) :: rc
    write(*, *) "NEGF device error in ", where, ": status ", rc, " ", trim(negf_error_message())
    stop 1
  end subroutine negf_shim_fail
end module negf_shim_ctl

! Binds the device context to the current dimensions and uploads Htotd / HopX /
! gMx,gMy,gMz once per Poisson iteration (after Build_Hamiltonian, NEGF_Main.F90:15).
subroutine NEGF_Prepare_Device
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  type(negf_config) :: cfg
  complex(8), allocatable :: gm(:, :, :)
  integer(c_int) :: rc
  if (.not. c_associated(negf_shim_ctx)) then
    cfg%npx = Npx; cfg%npy = Npy; cfg%npz = Npz; cfg%norb = Norb
    cfg%eta = eta_sigma
    cfg%sr_tol = negf_shim_sr_tol
    cfg%sr_max_iter = negf_shim_sr_max_iter
    cfg%sr_extra_eta = negf_shim_sr_extra_eta
    cfg%precision = negf_shim_precision
    cfg%device = negf_shim_device
    cfg%hbm_budget_gb = 0d0
    cfg%max_batch = 0
    cfg%flags = negf_shim_flags
    cfg%lanes = 0
    cfg%gamma_trunc = negf_shim_gamma_trunc
    rc = negf_init(negf_shim_ctx, cfg)
    if (rc /= 0) call negf_shim_fail("negf_init", rc)
    negf_shim_h_ite = -huge(1)
  end if
  if (negf_shim_h_ite /= ite) then
    allocate(gm(Norb, Norb, 3))
    gm = (0d0, 0d0)
    gm(1:Nh, 1:Nh, 1) = gMx
    gm(1:Nh, 1:Nh, 2) = gMy
    gm(1:Nh, 1:Nh, 3) = gMz
    rc = negf_set_hamiltonian(negf_shim_ctx, Htotd, HopX, gm)
    if (rc /= 0) call negf_shim_fail("negf_set_hamiltonian", rc)
    negf_shim_h_ite = ite
  end if
end subroutine NEGF_Prepare_Device

! Forces a re-upload of the Hamiltonian (call after changing Htotd/HopX within an iteration).
subroutine NEGF_Reload_Hamiltonian
  use negf_shim_ctl
  implicit none
  negf_shim_h_ite = -huge(1)
  call NEGF_Prepare_Device
end subroutine NEGF_Reload_Hamiltonian

subroutine NEGF_Finalize_Device
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc
  if (c_associated(negf_shim_ctx)) rc = negf_finalize(negf_shim_ctx)
  negf_shim_ctx = c_null_ptr
end subroutine NEGF_Finalize_Device

! Step occupations and energy step of Energy_Sweep.F90:7-23.
subroutine negf_shim_occupations
  use NEGF_STATE_MODULE
  implicit none
  if (eCtr /= (eStpnum - 1)) deltaE = EAxis(eCtr + 2) - EAxis(eCtr + 1)
  if (energy <= muT - mu(1)) then
    f1 = 1d0
  else
    f1 = 0d0
  end if
  if (energy <= muT - mu(2)) then
    f2 = 1d0
  else
    f2 = 0d0
  end if
  f1v = 1d0 - f1
  f2v = 1d0 - f2
end subroutine negf_shim_occupations

! Accumulator mask: every per-energy output, densities, contact and bond currents on
! the observables iteration (Energy_Sweep.F90:44); otherwise only what
! Contact_Self_Energy and Calculate_Gnp accumulate every iteration (DOSc, RhoN, RhoP,
! and the resolved densities ERRhon / ORRhon).  er / orb: erCD / orCD = 1 with the
! resolved arrays allocated (Calculate_Gnp.F90:125-130, Calculate_Observables.F90:111-168).
integer(c_int) function negf_shim_mask(obs, spin, jxyz, er, orb)
  use iso_c_binding
  implicit none
  logical, intent(in) :: obs, spin, jxyz, er, orb
  negf_shim_mask = ibset(ibset(ibset(ibset(0, 3), 4), 7), 8)
  if (er) negf_shim_mask = ibset(negf_shim_mask, 19)   ! ERRhon
  if (orb) negf_shim_mask = ibset(negf_shim_mask, 23)  ! ORRhon
  if (obs) then
    negf_shim_mask = ior(negf_shim_mask, int(b'1000001100111', c_int))  ! T12, Modec1/2, E/HDensity, ContactCurrent
    if (spin) negf_shim_mask = ior(negf_shim_mask, int(b'0111000000000', c_int))
    if (jxyz) negf_shim_mask = ior(negf_shim_mask, ishft(7_c_int, 13))  ! Jx, Jy, Jz
    if (er) negf_shim_mask = ior(negf_shim_mask, ishft(7_c_int, 16))    ! ERJx, ERJy, ERJz
    if (orb) negf_shim_mask = ior(negf_shim_mask, ishft(7_c_int, 20))   ! ORJx, ORJy, ORJz
  end if
end function negf_shim_mask

! Which outputs this sweep feeds: the observables gate (Energy_Sweep.F90:44), spin
! (calcSigExp, Read_In_Allocate.F90:93), bond currents and the resolved arrays
! (erCD / orCD), each only where its arrays are allocated.
subroutine negf_shim_outputs(obs, spin, jxyz, er, orb)
  use NEGF_STATE_MODULE
  implicit none
  logical, intent(out) :: obs, spin, jxyz, er, orb
  obs = (ite == iteMax .or. poissonIsConverged == 1)
  spin = (calcSigExp == 1)
  if (spin) spin = allocated(SigXExp)
  jxyz = allocated(Jx)
  er = (erCD == 1)
  if (er) er = allocated(ERJx)
  orb = (orCD == 1)
  if (orb) orb = allocated(ORJx)
end subroutine negf_shim_outputs

! The fused sweep / flush with the module's arrays: an unallocated one arrives as an
! absent optional argument (NULL), and the mask keeps the library off the arrays
! this iteration does not feed.
subroutine negf_shim_sweep_call(mask, rc)
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int), intent(in) :: mask
  integer(c_int), intent(out) :: rc
  rc = negf_energy_sweep_f(negf_shim_ctx, energy, eCtr, deltaE, mu, muT, mask, T12, Modec1, Modec2, &
      DOSc1, DOSc2, EDensity, HDensity, RhoN, RhoP, SigXExp, SigYExp, SigZExp, ContactCurrent, &
      Jx, Jy, Jz, ERJx, ERJy, ERJz, ERRhon, ORJx, ORJy, ORJz, ORRhon)
end subroutine negf_shim_sweep_call

subroutine negf_shim_flush_call(mask, rc)
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int), intent(in) :: mask
  integer(c_int), intent(out) :: rc
  rc = negf_flush_f(negf_shim_ctx, mu, muT, mask, T12, Modec1, Modec2, DOSc1, DOSc2, EDensity, &
      HDensity, RhoN, RhoP, SigXExp, SigYExp, SigZExp, ContactCurrent, Jx, Jy, Jz, ERJx, ERJy, ERJz, &
      ERRhon, ORJx, ORJy, ORJz, ORRhon)
end subroutine negf_shim_flush_call

subroutine Energy_Sweep
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc, mask
  integer(c_int), external :: negf_shim_mask
  logical :: obs, spin, jxyz, er, orb
  if (mpiRank == mpiRoot) write(*, *) eCtr
  call negf_shim_occupations
  if (contactType /= 0 .or. isBdGHam /= 0) then
    write(*, *) "NEGF device path: only semi-infinite normal contacts (contactType 0, no BdG)"
    stop 1
  end if
  call NEGF_Prepare_Device
  if (negf_shim_staged) then
    call negf_shim_staged_sweep
    return
  end if
  call negf_shim_outputs(obs, spin, jxyz, er, orb)
  mask = negf_shim_mask(obs, spin, jxyz, er, orb)
  call negf_shim_sweep_call(mask, rc)
  if (rc < 0) call negf_shim_fail("Energy_Sweep", rc)
  if (rc > 0) then
    write(*, *) "NEGF: energy status ", rc, " (SR not converged / NaN / singular block), Energy = ", energy
    errorFlag = rc
  end if
end subroutine Energy_Sweep

! Per-stage sweep: the reference's own sequence of calls (Energy_Sweep.F90:25-46).
subroutine negf_shim_staged_sweep
  use NEGF_STATE_MODULE
  implicit none
  integer :: x, i
  GRd = c0; GRu = c0; GRl = c0; InvGRd = c0; GRLef = c0; GRRgt = c0
  SigmaL = c0; SigmaR = c0; GammaL = c0; GammaR = c0
  call Contact_Self_Energy
  do x = 1, Npx
    InvGRd(:, :, x) = -Htotd(:, :, x)
    do i = 1, NNz
      InvGRd(i, i, x) = InvGRd(i, i, x) + cmplx(energy, eta_sigma, 8)
    end do
  end do
  InvGRd(:, :, 1) = InvGRd(:, :, 1) - SigmaL
  InvGRd(:, :, Npx) = InvGRd(:, :, Npx) - SigmaR
  call Get_Retarded_Greens
  call Calculate_Gnp
  if (ite == iteMax .or. poissonIsConverged == 1) call Calculate_Observables
end subroutine negf_shim_staged_sweep

! Batched form: queue this energy (use in place of CALL Energy_Sweep in the loop
! of NEGF_Main.F90:21-27), then CALL NEGF_Flush before MPI_Gather_Iteration.
subroutine Energy_Sweep_Submit
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc
  real(c_double) :: e(1), d(1)
  integer(c_int) :: k(1)
  call negf_shim_occupations
  call NEGF_Prepare_Device
  e(1) = energy; d(1) = deltaE; k(1) = eCtr
  rc = negf_submit(negf_shim_ctx, 1, e, k, d)
  if (rc /= 0) call negf_shim_fail("negf_submit", rc)
end subroutine Energy_Sweep_Submit

subroutine NEGF_Flush
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc, mask
  integer(c_int), external :: negf_shim_mask
  logical :: obs, spin, jxyz, er, orb
  if (.not. c_associated(negf_shim_ctx)) return
  call negf_shim_outputs(obs, spin, jxyz, er, orb)
  mask = negf_shim_mask(obs, spin, jxyz, er, orb)
  call negf_shim_flush_call(mask, rc)
  if (rc < 0) call negf_shim_fail("NEGF_Flush", rc)
  if (rc > 0) errorFlag = rc
end subroutine NEGF_Flush

subroutine Contact_Self_Energy
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc, iters(2)
  real(c_double) :: dos(2)
  call NEGF_Prepare_Device
  rc = negf_contact_self_energy(negf_shim_ctx, energy, SigmaL, SigmaR, GammaL, GammaR, dos, iters)
  if (rc < 0) call negf_shim_fail("Contact_Self_Energy", rc)
  if (rc > 0) then
    write(*, *) "Error in SR algorithm, Sigma not converging, Energy = ", energy
    errorFlag = 1
  end if
  DOSc1(eCtr + 1) = DOSc1(eCtr + 1) + dos(1)
  DOSc2(eCtr + 1) = DOSc2(eCtr + 1) + dos(2)
end subroutine Contact_Self_Energy

subroutine Sancho_Rubio(NNin, alpha, beta, Sigma, errFlag)
  use negf_shim_ctl
  implicit none
  integer(4) :: NNin, errFlag
  complex(8) :: alpha(NNin, NNin), beta(NNin, NNin), Sigma(NNin, NNin)
  integer(c_int) :: rc, iters
  call NEGF_Prepare_Device
  rc = negf_sancho_rubio(negf_shim_ctx, alpha, beta, Sigma, iters)
  if (rc < 0) call negf_shim_fail("Sancho_Rubio", rc)
  errFlag = 0
  if (rc > 0) errFlag = 1
end subroutine Sancho_Rubio

subroutine Get_Retarded_Greens
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc
  call NEGF_Prepare_Device
  rc = negf_retarded(negf_shim_ctx, InvGRd, GRd, GRu, GRl, GRLef, GRRgt)
  if (rc < 0) call negf_shim_fail("Get_Retarded_Greens", rc)
  if (rc > 0) then
    write(*, *) "ERROR: Inverse of GR is singular... Dumping data"
    write(*, *) "INFO = ", rc
    write(*, *) "Energy = ", energy
  end if
end subroutine Get_Retarded_Greens

subroutine Calculate_Gnp
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc
  call NEGF_Prepare_Device
  rc = negf_gnp(negf_shim_ctx, f1, f2, deltaE, GRLef, GRRgt, GammaL, GammaR, Gnd, Gnu, Gnl, Gpd, &
      Gpu, Gpl, RhoN, RhoP)
  if (rc /= 0) call negf_shim_fail("Calculate_Gnp", rc)
end subroutine Calculate_Gnp

subroutine Calculate_Observables
  use NEGF_STATE_MODULE
  use negf_shim_ctl
  implicit none
  integer(c_int) :: rc, mask
  logical :: spin
  real(c_double) :: dmy(1)
  call NEGF_Prepare_Device
  if (isBdGHam /= 0) then
    write(*, *) "NEGF device path: the BdG branch of Calculate_Observables is not provided"
    stop 1
  end if
  spin = (calcSigExp == 1)
  if (spin) spin = allocated(SigXExp)
  ! T12, Modec1/2, E/HDensity, ContactCurrent, Jx, Jy, Jz (+ spin)
  mask = ior(int(b'1111000001100111', c_int), merge(int(b'0111000000000', c_int), 0_c_int, spin))
  dmy = 0d0
  if (spin) then
    rc = negf_observables_f(negf_shim_ctx, eCtr, f1, f2, deltaE, Gnd, Gpd, Gnu, Gnl, Gpu, Gpl, GRRgt, &
        GammaL, GammaR, mask, T12, Modec1, Modec2, dmy, dmy, EDensity, HDensity, dmy, dmy, &
        SigXExp, SigYExp, SigZExp, ContactCurrent, Jx, Jy, Jz)
  else
    rc = negf_observables_f(negf_shim_ctx, eCtr, f1, f2, deltaE, Gnd, Gpd, Gnu, Gnl, Gpu, Gpl, GRRgt, &
        GammaL, GammaR, mask, T12, Modec1, Modec2, dmy, dmy, EDensity, HDensity, dmy, dmy, &
        dmy, dmy, dmy, ContactCurrent, Jx, Jy, Jz)
  end if
  if (rc /= 0) call negf_shim_fail("Calculate_Observables", rc)
end subroutine Calculate_Observables

subroutine Invert_Matrix(invG, G, N, errFlag)
  use negf_capi
  use negf_shim_ctl, only: negf_shim_fail
  implicit none
  integer(4) :: N, errFlag
  complex(8) :: invG(N, N), G(N, N)
  integer(c_int) :: rc, info
  rc = negf_invert_matrix(N, G, invG, info)
  if (rc /= 0) call negf_shim_fail("Invert_Matrix", rc)
  errFlag = info
end subroutine Invert_Matrix
