! Test-only driver around the reference's own Get_Retarded_Greens and
! Calculate_Gnp, compiled unmodified from Source/ (Recursive_Greens.F90,
! Calculate_Gnp.F90, MKL_Routines_Cluster.F90 and the modules they USE).
! It fills the NEGF_Module arrays the two subroutines read, calls them, and
! copies their outputs out through a C ABI so the tests can compare the CPU
! oracle with the reference itself.  Nothing here is part of the product.
module negf_ref_driver
  use iso_c_binding
  use NEGF_Module
  implicit none
contains

  subroutine reset_arrays()
    if (allocated(InvGRd)) deallocate(InvGRd)
    if (allocated(HopX)) deallocate(HopX)
    if (allocated(GRd)) deallocate(GRd)
    if (allocated(GRu)) deallocate(GRu)
    if (allocated(GRl)) deallocate(GRl)
    if (allocated(GRLef)) deallocate(GRLef)
    if (allocated(GRRgt)) deallocate(GRRgt)
    if (allocated(GammaL)) deallocate(GammaL)
    if (allocated(GammaR)) deallocate(GammaR)
    if (allocated(Gnd)) deallocate(Gnd)
    if (allocated(Gnu)) deallocate(Gnu)
    if (allocated(Gnl)) deallocate(Gnl)
    if (allocated(Gpd)) deallocate(Gpd)
    if (allocated(Gpu)) deallocate(Gpu)
    if (allocated(Gpl)) deallocate(Gpl)
    if (allocated(RhoN)) deallocate(RhoN)
    if (allocated(RhoP)) deallocate(RhoP)
  end subroutine reset_arrays

  ! Get_Retarded_Greens on a given InvGRd / HopX (shapes as Read_In_Allocate.F90:104-113).
  function ref_retarded(n, L, inv_in, hop_in, grd_out, gru_out, grl_out, lef_out, rgt_out) &
      result(info) bind(C, name="ref_retarded")
    integer(c_int), value :: n, L
    complex(c_double_complex), intent(in) :: inv_in(n, n, L), hop_in(n, n)
    complex(c_double_complex), intent(out) :: grd_out(n, n, L), gru_out(n, n, L - 1), &
        grl_out(n, n, L - 1), lef_out(n * L, n), rgt_out(n * L, n)
    integer(c_int) :: info
    call reset_arrays()
    NNz = n; Npx = L; Ntot = n * L
    energy = 0d0; errorFlag = 0
    allocate(InvGRd(n, n, L), HopX(n, n), GRd(n, n, L), GRu(n, n, L - 1), GRl(n, n, L - 1))
    allocate(GRLef(Ntot, n), GRRgt(Ntot, n))
    InvGRd = inv_in; HopX = hop_in
    GRd = c0; GRu = c0; GRl = c0; GRLef = c0; GRRgt = c0
    call Get_Retarded_Greens
    grd_out = GRd; gru_out = GRu; grl_out = GRl; lef_out = GRLef; rgt_out = GRRgt
    info = errorFlag
  end function ref_retarded

  ! Calculate_Gnp on given columns and broadenings (erCD = orCD = 0).
  subroutine ref_gnp(n, L, py, pz, norb_in, f1_in, f2_in, de, lef_in, rgt_in, gl_in, gr_in, &
      gnd_o, gnu_o, gnl_o, gpd_o, gpu_o, gpl_o, rhon_o, rhop_o) bind(C, name="ref_gnp")
    integer(c_int), value :: n, L, py, pz, norb_in
    real(c_double), value :: f1_in, f2_in, de
    complex(c_double_complex), intent(in) :: lef_in(n * L, n), rgt_in(n * L, n), gl_in(n, n), &
        gr_in(n, n)
    complex(c_double_complex), intent(out) :: gnd_o(n, n, L), gnu_o(n, n, L - 1), &
        gnl_o(n, n, L - 1), gpd_o(n, n, L), gpu_o(n, n, L - 1), gpl_o(n, n, L - 1)
    real(c_double), intent(out) :: rhon_o(L, py, pz), rhop_o(L, py, pz)
    call reset_arrays()
    Npx = L; Npy = py; Npz = pz; Norb = norb_in; Nh = norb_in
    NN = Npy * Norb; NNz = Npz * NN; Ntot = Npx * NNz
    f1 = f1_in; f2 = f2_in; f1v = 1d0 - f1; f2v = 1d0 - f2
    deltaE = de; erCD = 0; orCD = 0; eCtr = 0
    allocate(GRLef(Ntot, NNz), GRRgt(Ntot, NNz), GammaL(NNz, NNz), GammaR(NNz, NNz))
    allocate(Gnd(NNz, NNz, Npx), Gpd(NNz, NNz, Npx), Gnu(NNz, NNz, Npx - 1), &
             Gpu(NNz, NNz, Npx - 1), Gnl(NNz, NNz, Npx - 1), Gpl(NNz, NNz, Npx - 1))
    allocate(RhoN(Npx, Npy, Npz), RhoP(Npx, Npy, Npz))
    GRLef = lef_in; GRRgt = rgt_in; GammaL = gl_in; GammaR = gr_in
    RhoN = 0d0; RhoP = 0d0
    call Calculate_Gnp
    gnd_o = Gnd; gnu_o = Gnu; gnl_o = Gnl; gpd_o = Gpd; gpu_o = Gpu; gpl_o = Gpl
    rhon_o = RhoN; rhop_o = RhoP
  end subroutine ref_gnp

  ! Invert_Matrix (MKL_Routines_Cluster.F90:2-33) as the reference calls it.
  function ref_invert(n, g_in, inv_out) result(info) bind(C, name="ref_invert")
    integer(c_int), value :: n
    complex(c_double_complex), intent(in) :: g_in(n, n)
    complex(c_double_complex), intent(out) :: inv_out(n, n)
    integer(c_int) :: info
    integer(4) :: nn4, flag
    nn4 = n; flag = 0
    call Invert_Matrix(inv_out, g_in, nn4, flag)
    info = flag
  end function ref_invert

end module negf_ref_driver
