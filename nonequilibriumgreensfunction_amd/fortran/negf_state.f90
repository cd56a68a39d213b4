! Module state read and written by the per-energy solve, with the names, kinds and
! shapes the reference's NEGF_Module / SIP_Module give them (NEGF_Module.F90:27-105,
! Read_In_Allocate.F90:104-149; poissonIsConverged from SIP_Module).  A standalone
! host (and the tests) USE this module; inside the reference build the shims USE
! the reference's own SIP_Module instead (negf_shims.F90, NEGF_USE_REFERENCE_MODULE).
! eta_sigma is a variable here (a PARAMETER in the reference, Constants.F90:7).
module negf_state
  implicit none
  complex(8), parameter :: c0 = (0d0, 0d0), c1 = (1d0, 0d0), ci = (0d0, 1d0)
  real(8), parameter :: pi = 3.14159265358979323846d0

  real(8) :: eta_sigma = 1d-4

  integer :: Npx, Npy, Npz, Norb, Nh, NN, NNz, Ntot
  integer :: eStpnum, eCtr, ite = 1, iteMax = 1
  integer :: poissonIsConverged = 0
  integer(4) :: calcSigExp = 1, erCD = 0, orCD = 0, isBdGHam = 0
  integer :: contactType = 0
  integer :: errorFlag = 0
  integer(4) :: mpiRank = 0, mpiRoot = 0

  real(8) :: energy, deltaE, f1, f2, f1v, f2v
  real(8), dimension(2) :: mu
  real(8) :: muT
  real(8), dimension(:), allocatable :: EAxis

  complex(8), dimension(:,:,:), allocatable :: Htotd
  complex(8), dimension(:,:), allocatable :: HopX
  complex(8), dimension(:,:), allocatable :: gMx, gMy, gMz

  complex(8), dimension(:,:,:), allocatable :: InvGRd, GRd, GRu, GRl, Gnd, Gpd, Gnu, Gpu, Gnl, Gpl
  complex(8), dimension(:,:), allocatable :: GRLef, GRRgt
  complex(8), dimension(:,:), allocatable :: SigmaL, SigmaR, GammaL, GammaR

  real(8), dimension(:,:,:), allocatable :: RhoN, RhoP, SigXExp, SigYExp, SigZExp, Jx, Jy, Jz
  ! energy- and orbital-resolved outputs, allocated when erCD / orCD = 1
  ! (NEGF_Module.F90:100, Read_In_Allocate.F90:228-239)
  real(8), dimension(:,:,:,:), allocatable :: ERJx, ERJy, ERJz, ERRhon, ORJx, ORJy, ORJz, ORRhon
  real(8), dimension(:), allocatable :: T12, Modec1, Modec2, DOSc1, DOSc2, EDensity, HDensity
  real(8), dimension(:), allocatable :: ContactCurrent

contains

  ! Read_In_Allocate's allocations for the hot path (Read_In_Allocate.F90:104-149),
  ! zero-initialised; dimensions must be set first.
  subroutine negf_state_allocate()
    NN = Npy * Norb
    NNz = Npz * NN
    Ntot = Npx * NNz
    Nh = Norb
    if (allocated(Htotd)) call negf_state_deallocate()
    allocate(Htotd(NNz, NNz, Npx), HopX(NNz, NNz), InvGRd(NNz, NNz, Npx))
    allocate(GRd(NNz, NNz, Npx), GRu(NNz, NNz, Npx - 1), GRl(NNz, NNz, Npx - 1))
    allocate(GRLef(Ntot, NNz), GRRgt(Ntot, NNz))
    allocate(Gnd(NNz, NNz, Npx), Gpd(NNz, NNz, Npx), Gnu(NNz, NNz, Npx - 1), Gpu(NNz, NNz, Npx - 1), &
             Gnl(NNz, NNz, Npx - 1), Gpl(NNz, NNz, Npx - 1))
    allocate(RhoN(Npx, Npy, Npz), RhoP(Npx, Npy, Npz), SigXExp(Npx, Npy, Npz), &
             SigYExp(Npx, Npy, Npz), SigZExp(Npx, Npy, Npz), Jx(Npx, Npy, Npz), Jy(Npx, Npy, Npz), &
             Jz(Npx, Npy, Npz))
    allocate(GammaL(NNz, NNz), GammaR(NNz, NNz), SigmaL(NNz, NNz), SigmaR(NNz, NNz))
    allocate(T12(eStpnum), Modec1(eStpnum), Modec2(eStpnum), DOSc1(eStpnum), DOSc2(eStpnum), &
             EDensity(eStpnum), HDensity(eStpnum), EAxis(eStpnum), ContactCurrent(2))
    allocate(gMx(Nh, Nh), gMy(Nh, Nh), gMz(Nh, Nh))
    Htotd = c0; HopX = c0; InvGRd = c0; GRd = c0; GRu = c0; GRl = c0; GRLef = c0; GRRgt = c0
    Gnd = c0; Gpd = c0; Gnu = c0; Gpu = c0; Gnl = c0; Gpl = c0
    GammaL = c0; GammaR = c0; SigmaL = c0; SigmaR = c0
    RhoN = 0d0; RhoP = 0d0; SigXExp = 0d0; SigYExp = 0d0; SigZExp = 0d0; Jx = 0d0; Jy = 0d0; Jz = 0d0
    T12 = 0d0; Modec1 = 0d0; Modec2 = 0d0; DOSc1 = 0d0; DOSc2 = 0d0; EDensity = 0d0; HDensity = 0d0
    EAxis = 0d0; ContactCurrent = 0d0
    gMx = c0; gMy = c0; gMz = c0
    if (erCD == 1) then
      allocate(ERJx(Npx, Npy, Npz, eStpnum), ERJy(Npx, Npy, Npz, eStpnum), &
               ERJz(Npx, Npy, Npz, eStpnum), ERRhon(Npx, Npy, Npz, eStpnum))
      ERJx = 0d0; ERJy = 0d0; ERJz = 0d0; ERRhon = 0d0
    end if
    if (orCD == 1) then
      allocate(ORJx(Npx, Npy, Npz, Norb * Norb), ORJy(Npx, Npy, Npz, Norb * Norb), &
               ORJz(Npx, Npy, Npz, Norb * Norb), ORRhon(Npx, Npy, Npz, Norb))
      ORJx = 0d0; ORJy = 0d0; ORJz = 0d0; ORRhon = 0d0
    end if
  end subroutine negf_state_allocate

  subroutine negf_state_deallocate()
    deallocate(Htotd, HopX, InvGRd, GRd, GRu, GRl, GRLef, GRRgt, Gnd, Gpd, Gnu, Gpu, Gnl, Gpl)
    deallocate(RhoN, RhoP, SigXExp, SigYExp, SigZExp, Jx, Jy, Jz, GammaL, GammaR, SigmaL, SigmaR)
    deallocate(T12, Modec1, Modec2, DOSc1, DOSc2, EDensity, HDensity, EAxis, ContactCurrent)
    deallocate(gMx, gMy, gMz)
    if (allocated(ERJx)) deallocate(ERJx, ERJy, ERJz, ERRhon)
    if (allocated(ORJx)) deallocate(ORJx, ORJy, ORJz, ORRhon)
  end subroutine negf_state_deallocate

end module negf_state
