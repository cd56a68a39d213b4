! Test host: a Fortran program that drives the reference-named shims the way
! NEGF_Main does (NEGF_Main.F90:21-38): fill module state, loop over energies
! calling Energy_Sweep (or Energy_Sweep_Submit + NEGF_Flush), then write the
! accumulated outputs.  Usage: negf_fdriver <input.bin> <output.bin>
! Input (stream, native endian): int32 Npx Npy Npz Norb eStpnum staged submit
! sr_max_iter; float64 eta sr_tol mu1 mu2 muT; float64 EAxis(eStpnum); complex128
! Htotd(n,n,Npx), HopX(n,n), gMx, gMy, gMz (Norb,Norb).
! Output: float64 T12 Modec1 Modec2 DOSc1 DOSc2 EDensity HDensity (eStpnum each),
! RhoN RhoP SigXExp SigYExp SigZExp Jx Jy Jz (Npx*Npy*Npz each), ContactCurrent(2);
! int32 errorFlag.
program negf_fdriver
  use negf_state
  use negf_shim_ctl
  implicit none
  character(len=1024) :: fin, fout
  integer(4) :: hdr(8)
  real(8) :: par(5)
  integer :: k

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) par
  Npx = hdr(1); Npy = hdr(2); Npz = hdr(3); Norb = hdr(4); eStpnum = hdr(5)
  eta_sigma = par(1)
  mu(1) = par(3); mu(2) = par(4); muT = par(5)
  call negf_state_allocate()
  read(10) EAxis
  read(10) Htotd
  read(10) HopX
  read(10) gMx
  read(10) gMy
  read(10) gMz
  close(10)
  negf_shim_sr_tol = par(2)
  negf_shim_sr_max_iter = hdr(8)
  negf_shim_staged = (hdr(6) == 1)
  calcSigExp = 1
  ite = 1; iteMax = 1
  deltaE = 0d0
  if (eStpnum > 1) deltaE = EAxis(2) - EAxis(1)
  do k = 0, eStpnum - 1
    eCtr = k
    energy = EAxis(eCtr + 1)
    if (hdr(7) == 1) then
      call Energy_Sweep_Submit
    else
      call Energy_Sweep
    end if
  end do
  if (hdr(7) == 1) call NEGF_Flush
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) T12, Modec1, Modec2, DOSc1, DOSc2, EDensity, HDensity
  write(11) RhoN, RhoP, SigXExp, SigYExp, SigZExp, Jx, Jy, Jz, ContactCurrent
  write(11) int(errorFlag, 4)
  close(11)
  call NEGF_Finalize_Device
end program negf_fdriver
