/*
 * negf_capi.h -- C ABI of the MI355X (gfx950) per-energy NEGF solve.
 *
 * Drop-in boundary for the Energy_Sweep path of the reference
 * (brian-dellabetta/nonEquilibriumGreensFunction).  The reference's hot-path
 * subroutines take no arguments and read/write NEGF_Module globals
 * (NEGF_Module.F90:51-100); these entry points take the same arrays as plain
 * pointers, so a Fortran caller binds them with iso_c_binding and passes its
 * module arrays (see the shims in fortran/negf_shims.F90).
 *
 * Conventions
 *  - Complex arrays are COMPLEX(8): interleaved (re, im) doubles, column-major,
 *    passed as `double*`.  Htotd(i,j,x) is element ((x*n)+j)*n+i (0-based),
 *    n = NNz = Npy*Npz*Norb, with basis index orb + Norb*(y + Npy*z).
 *  - Spatial arrays (RhoN, SigXExp, Jx, ...) are REAL(8) (Npx,Npy,Npz), Fortran order.
 *  - Per-energy arrays (T12, Modec1, ...) are REAL(8)(eStpnum), indexed by eCtr (0-based).
 *  - Every call returns an int status: 0 = ok; > 0 = LAPACK-style INFO of the first
 *    singular block (or the NEGF_STATUS_* bits below); < 0 = API/HIP error, with a
 *    message available from negf_last_error().
 *  - Single host thread per context.  Device work is asynchronous inside
 *    negf_submit(); negf_flush() must return before the host reads accumulators
 *    (the reference reads them at MPI_Gather_Iteration, NEGF_Main.F90:30).
 *  - There is no host (CPU) fallback: without a usable gfx950 device every entry
 *    point fails with NEGF_ERR_NO_DEVICE.
 */
#ifndef NEGF_CAPI_H
#define NEGF_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define NEGF_OK 0
#define NEGF_ERR_ARG (-1)
#define NEGF_ERR_HIP (-2)
#define NEGF_ERR_NO_DEVICE (-3)
#define NEGF_ERR_MEMORY (-4)
#define NEGF_ERR_STATE (-5)
/* positive status bits (per energy), combined with a LAPACK INFO in the low 20 bits */
#define NEGF_STATUS_SR_NOT_CONVERGED (1 << 20) /* Contact_Self_Energy.F90:202-205 */
#define NEGF_STATUS_SR_NAN (1 << 21)           /* Contact_Self_Energy.F90:182-191 */

/* precision forms */
#define NEGF_PREC_FP64 0  /* plain complex128 on FP64 MFMA */
#define NEGF_PREC_MIXED 1 /* FP64 contacts, inverses and T12 chain; complex64 MFMA for the
                             first-column chain, whose site blocks feed the densities */
#define NEGF_PREC_REFINED 2 /* NEGF_PREC_MIXED with every RGF block inverse taken in complex64
                               (Gauss-Jordan on FP32 MFMA) and refined in FP64 by Newton steps
                               X <- X + X (I - M X) until max|I - M X| <= 1e-12 (at most two);
                               a block whose complex64 inverse is singular or whose residual does
                               not contract is inverted again in FP64 (escalation) */
#define NEGF_PREC_LU32 3 /* block LU in complex64: every RGF block inverse (Gauss-Jordan on
                            FP32 MFMA) and both column chains in complex64; FP64 contacts.
                            The leading 64 columns of both block columns' factor products
                            (G_xL W_R, G_x1 W_L with Gamma = W W') are solved to FP64
                            accuracy by iterative refinement of the block-tridiagonal system
                            (FP64 residuals, corrections through the complex64 block LU, two
                            sweeps); every output sums those refined columns with the other
                            columns' complex64 values.  T12's part beyond W_R's leading
                            columns comes from a complex64 first-row chain and a Gaussian
                            sketch (32 columns) of both factors' tails.  An energy whose
                            complex64 inverse is singular, whose refinement does not contract
                            to an estimated error of 1e-10 of the solution, or whose T12 parts
                            beyond the refined columns exceed 1e-3 (right tail) or 1e-9 (both
                            tails) of max(1, T12) is solved again in the plain FP64 form
                            (escalation).  Needs a site-block-diagonal HopX (otherwise every
                            energy takes the plain FP64 form) and the two-pass RGF. */

typedef struct negf_ctx negf_ctx;

/* Dimensions and solver options.  Replaces the globals Npx, Npy, Npz, Norb
 * (NEGF_Module.F90:27), eta_sigma (Constants.F90:7) and the Sancho-Rubio
 * constants hard-coded at Contact_Self_Energy.F90:141,154. */
typedef struct {
  int npx, npy, npz, norb;
  double eta;          /* eta_sigma */
  double sr_tol;       /* SR stop: max|change| <= sr_tol */
  int sr_max_iter;     /* SR iteration cap, reset per call */
  int sr_extra_eta;    /* 1: SR seed inverts alpha + i*eta*I as Contact_Self_Energy.F90:141 does */
  int precision;       /* NEGF_PREC_* */
  int device;          /* HIP device ordinal */
  double hbm_budget_gb;/* device memory the context may use for in-flight energies; 0 = auto */
  int max_batch;       /* cap on energies in flight; 0 = auto */
  int flags;           /* NEGF_FLAG_* (0 = default algorithms) */
  int lanes;           /* concurrent device lanes of negf_flush, 1..4 (0 = the tuning
                          table's value for this block size, else 4) */
  double gamma_trunc;  /* error-controlled truncation of the contact broadening in the
                          fused solve (no counterpart in the reference, which always uses
                          the whole Gamma): the factor Gamma = W W' of each lead stops once
                          the trace it leaves out is at most gamma_trunc * tr(Gamma), and
                          the column chains, site blocks and traces run on those columns
                          only.  0 = keep every column (the whole Gamma up to 1e-15 of its
                          largest pivot).  The outputs move by about gamma_trunc relative
                          (measured: T12 ~0.02x, densities ~0.2x); see the precision
                          forms' documentation for the values used with them. */
} negf_config;

/* Algorithm selection (results agree to rounding either way).
 * Default fused RGF: one forward pass of left-connected inverses gL_x, then both
 * block columns by block back-substitution (L inverses per energy, 2L resident
 * blocks).  THREE_PASS: the reference's structure with right-connected inverses
 * gR_x for the first column (2L inverses, L resident blocks; for tight memory).
 * Default Sancho-Rubio: when the two end slices are bitwise equal (clean leads of
 * the same material) one decimation serves both leads, whose (t, tt) sequences
 * are each other's swapped; SR_SEPARATE_LEADS runs one decimation per lead. */
#define NEGF_FLAG_RGF_THREE_PASS 1
#define NEGF_FLAG_SR_SEPARATE_LEADS 2
/* negf_flush runs on the calling thread and one HIP stream instead of two
 * concurrent lanes (for tools that cannot follow kernels dispatched from two
 * host threads, e.g. per-dispatch hardware counter collection). */
#define NEGF_FLAG_SINGLE_LANE 4
/* Keep each energy's contact self-energies (Sigma_L, Sigma_R, with the
 * Sancho-Rubio iteration counts) on the device and reuse them in later sweeps of
 * the same energies while HopX and the two end slices are bitwise unchanged --
 * e.g. across Poisson iterations that only move the interior potential
 * (NEGF_Main.F90:7-30).  Reused self-energies are the same bits, so results are
 * unchanged; the Sancho-Rubio cost is paid once per energy.  Up to 15% of the
 * free HBM holds the cache. */
#define NEGF_FLAG_CACHE_CONTACTS 8
/* Capture each batch's launches after the contacts (assembly, block inverses,
 * column chains, site and bond blocks, traces) into a HIP graph once per lane,
 * context, Hamiltonian upload and batch size, and replay it for later batches of
 * the same shape (plain and mixed forms; the refined form decides per block on
 * the host).  Results are bitwise those of the launches issued one by one.
 * Batches of more than about 20000 launches (e.g. blocks of 2304 over 200 slices)
 * run without a graph. */
#define NEGF_FLAG_GRAPH 16
/* NEGF_PREC_LU32 only: take the complex64 column-chain products and the block
 * inverses' trailing updates on the FP32 matrix cores (3M form) instead of the default
 * bf16 matrix cores with every FP32 operand split three ways (four real products of
 * six split products each, FP32 accuracy; 1.3 times the FP32 kernel's throughput for
 * the chains at block 2304). */
#define NEGF_FLAG_LU32_FP32_MFMA 32
/* Paired decimation only (equal end slices): by default, when the end slice and HopX
 * both keep a mirror of the cross-section (negf_find_mirror), every Sancho-Rubio
 * matrix is block diagonal in the mirror's two sectors, and the decimation runs on
 * the n/2 x n/2 sector blocks (a quarter of the arithmetic; an eighth when a second,
 * anticommuting mirror makes the two sectors equivalent); its stop test
 * (max|change| <= sr_tol) and Sigma are taken in the site basis as
 * Contact_Self_Energy.F90:178-209 does.  SR_NO_MIRROR keeps the n x n decimation. */
#define NEGF_FLAG_SR_NO_MIRROR 64

/* Energy-integrated and per-energy outputs, in the reference's layouts.  Every
 * pointer may be NULL (that output is skipped).  All accumulate (+=), as the
 * reference does.  The bond currents and the energy-/orbital-resolved arrays
 * (Calculate_Observables.F90:111-168, Calculate_Gnp.F90:125-130; the reference
 * fills the resolved ones when erCD / orCD = 1) are filled by the fused entry
 * points (negf_energy_sweep, negf_flush, the server); negf_observables takes its
 * bond currents as separate arguments.  Currents accumulate in REAL(8) (the
 * reference passes each term through a REAL(4), Calculate_Observables.F90:7).
 * Orbital pair (ii, jj), 0-based, of the OR current arrays is index ii*Norb + jj
 * ((ii-1)*Norb + jj 1-based in the reference). */
typedef struct {
  double *T12, *Modec1, *Modec2, *DOSc1, *DOSc2, *EDensity, *HDensity; /* [eStpnum] at eCtr */
  double *RhoN, *RhoP;                                                  /* (Npx,Npy,Npz) */
  double *SigXExp, *SigYExp, *SigZExp;                                  /* (Npx,Npy,Npz) */
  double *ContactCurrent;                                               /* [2] */
  double *Jx, *Jy, *Jz;                                                 /* (Npx,Npy,Npz) */
  double *ERJx, *ERJy, *ERJz, *ERRhon;  /* (Npx,Npy,Npz,eStpnum), at eCtr; unweighted by dE/2pi */
  double *ORJx, *ORJy, *ORJz;           /* (Npx,Npy,Npz,Norb*Norb) */
  double *ORRhon;                       /* (Npx,Npy,Npz,Norb) */
} negf_accum;

const char* negf_last_error(void);
int negf_device_count(void);

/* Read_In_Allocate (Read_In_Allocate.F90:104-149) analogue: validates the
 * config and binds a device.  No device memory is reserved yet. */
int negf_init(negf_ctx** ctx, const negf_config* cfg);
int negf_finalize(negf_ctx* ctx);

/* Upload Htotd (n,n,Npx), HopX (n,n) and the spin matrices gMx,gMy,gMz (three
 * column-major Norb x Norb blocks, Read_In_Allocate.F90:186-196).  Call after
 * Build_Hamiltonian / MPI_Broadcast (NEGF_Main.F90:15-19), i.e. once per
 * Poisson iteration. */
int negf_set_hamiltonian(negf_ctx* ctx, const double* Htotd, const double* HopX,
                         const double* gM);

/* Device-side Hamiltonian from a compact description: every slice is the
 * template H0 (n,n) -- the slice block with its y/z hoppings, Peierls phases and
 * Zeeman terms (Build_Hamiltonian.F90:59-100, :186-207, :348-360), the same in all
 * slices -- plus the real site potentials onsite (Npx,Npy,Npz), Fortran order,
 * on every orbital of the site (Insert_Impurity / Anderson disorder,
 * Build_Hamiltonian.F90:154-165, :278-296):
 *     Htotd(a,a,x) = H0(a,a) + onsite(x, site(a)),  Htotd(a,b,x) = H0(a,b) otherwise.
 * The slices are assembled on the device as they are used; n*n + Npx*Npy*Npz
 * values stay resident instead of n*n*Npx.  Results are identical to
 * negf_set_hamiltonian with the Htotd so defined. */
int negf_set_hamiltonian_compact(negf_ctx* ctx, const double* H0, const double* HopX,
                                 const double* gM, const double* onsite);

/* Replace the site potentials of a compact Hamiltonian (the part a Poisson
 * iteration changes: the electrostatic potential on the diagonal,
 * Build_Hamiltonian.F90:104-112) without re-uploading the template. */
int negf_set_onsite(negf_ctx* ctx, const double* onsite);
/* Contact-cache state (NEGF_FLAG_CACHE_CONTACTS): stored energies, and energies
 * served from the cache since it was last invalidated. */
int negf_contact_cache_stats(negf_ctx* ctx, int* entries, long long* hits);
/* NEGF_PREC_REFINED, no counterpart in the reference (whose Invert_Matrix is one
 * FP64 zgetrf + zgetri, MKL_Routines_Cluster.F90:19,23): block inverses taken since
 * negf_init, FP64 Newton steps applied to them, and blocks escalated to the FP64
 * inverse.  NEGF_PREC_LU32: energies solved, FP64 refinement sweeps (summed over
 * energies) and energies escalated to the plain FP64 form. */
int negf_refine_stats(negf_ctx* ctx, long long* blocks, long long* steps, long long* escalated);
/* No counterpart in the reference (whose Sancho_Rubio, Contact_Self_Energy.F90:141-209,
 * works on the whole n x n blocks): the mirror P of the lead cross-section that
 * the end slice Hend and HopX (n x n, column-major complex, n = Npy*Npz*Norb, site
 * y + Npy*z, orbital fastest) both keep to 1e-14 of their largest entry, tried as
 * z -> Npz-1-z first, then y -> Npy-1-y, on an even extent only:
 *     P e(site, o) = i^phase[o] e(mirror site, perm[o]),
 * perm an involution of the orbitals and i^phase[o] i^phase[perm[o]] = 1 (P is a
 * Hermitian unitary involution).  axis = 0 (none), 1 (y) or 2 (z); perm and phase
 * (Norb entries each, may be NULL) are filled when axis > 0.  paired (may be NULL)
 * = 1 when a mirror on the other axis also holds and anticommutes with P: it maps
 * one sector of P onto the other, and one n/2 x n/2 decimation serves both.
 * Norb <= 4.  Host only: needs no device. */
int negf_find_mirror(int npy, int npz, int norb, const double* Hend, const double* HopX, int* axis, int* perm,
                     int* phase, int* paired);
/* The mirror the context's paired decimation uses for the current Hamiltonian
 * (axis 0: none -- no mirror kept, end slices unequal, SR_SEPARATE_LEADS or
 * SR_NO_MIRROR set).  Same outputs as negf_find_mirror. */
int negf_contact_mirror(negf_ctx* ctx, int* axis, int* perm, int* phase, int* paired);
/* NEGF_FLAG_GRAPH, no counterpart in the reference: graphs captured and graph
 * replays (batches launched from a graph) since negf_init. */
int negf_graph_stats(negf_ctx* ctx, long long* built, long long* replays);

/* Energy_Sweep (Energy_Sweep.F90:1-46) for one energy, fused on the device:
 * contacts, RGF, Gn/Gp site blocks and observables.  mu[2], muT give the step
 * occupations of Energy_Sweep.F90:11-23.  Accumulates into acc at eCtr. */
int negf_energy_sweep(negf_ctx* ctx, double energy, int eCtr, double deltaE, const double* mu,
                      double muT, negf_accum* acc);

/* Batched form of the energy loop (NEGF_Main.F90:21-27): queue energies, then
 * flush.  The device runs queued energies in batches sized to the HBM budget.
 * status (optional, [count of queued energies]) receives per-energy status in
 * submission order; sr_iters (optional, 2 per energy) the SR iteration counts. */
int negf_submit(negf_ctx* ctx, int count, const double* energies, const int* eCtr,
                const double* deltaE);
int negf_flush(negf_ctx* ctx, const double* mu, double muT, negf_accum* acc, int* status,
               int* sr_iters);

/* --- Stage entry points (each mirrors one reference subroutine) ---------- */

/* Contact_Self_Energy, semi-infinite branch (Contact_Self_Energy.F90:68-101):
 * SigmaL/R and GammaL/R (n,n); dos[2] = this energy's DOSc1, DOSc2 increments;
 * iters[2] = SR iteration counts (optional). */
int negf_contact_self_energy(negf_ctx* ctx, double energy, double* SigmaL, double* SigmaR,
                             double* GammaL, double* GammaR, double* dos, int* iters);

/* Sancho_Rubio (Contact_Self_Energy.F90:110-223) for one lead given alpha, beta. */
int negf_sancho_rubio(negf_ctx* ctx, const double* alpha, const double* beta, double* Sigma,
                      int* iters);

/* Get_Retarded_Greens (Recursive_Greens.F90:2-107) on a given InvGRd (n,n,Npx)
 * and the context's HopX: GRd (n,n,Npx), GRu, GRl (n,n,Npx-1), GRLef, GRRgt
 * (Npx*n, n).  Any output may be NULL. */
int negf_retarded(negf_ctx* ctx, const double* InvGRd, double* GRd, double* GRu, double* GRl,
                  double* GRLef, double* GRRgt);

/* Calculate_Gnp (Calculate_Gnp.F90:1-136): Gnd/Gpd (n,n,Npx), Gnu/Gpu/Gnl/Gpl
 * (n,n,Npx-1) overwritten (any may be NULL); RhoN/RhoP accumulate. */
int negf_gnp(negf_ctx* ctx, double f1, double f2, double deltaE, const double* GRLef,
             const double* GRRgt, const double* GammaL, const double* GammaR, double* Gnd,
             double* Gnu, double* Gnl, double* Gpd, double* Gpu, double* Gpl, double* RhoN,
             double* RhoP);

/* Calculate_Observables, non-BdG branches (Calculate_Observables.F90:9-60,85-169)
 * with spin enabled: per-energy values go to acc->T12 .. HDensity at eCtr, spatial
 * arrays and ContactCurrent accumulate; Jx/Jy/Jz (optional) accumulate in REAL(8). */
int negf_observables(negf_ctx* ctx, int eCtr, double f1, double f2, double deltaE,
                     const double* Gnd, const double* Gpd, const double* Gnu, const double* Gnl,
                     const double* Gpu, const double* Gpl, const double* GRRgt,
                     const double* GammaL, const double* GammaR, negf_accum* acc, double* Jx,
                     double* Jy, double* Jz);

/* --- Fortran-friendly forms ------------------------------------------------
 * The same operations with the accumulators passed as separate arrays, so a
 * Fortran caller can pass its (non-TARGET) module arrays by reference.  Bit k of
 * `mask` marks the k-th array argument present, in negf_accum's field order
 * (0 T12 ... 12 ContactCurrent, 13-15 Jx Jy Jz, 16-19 ERJx ERJy ERJz ERRhon,
 * 20-23 ORJx ORJy ORJz ORRhon); negf_observables_f takes fields 0-15. */
int negf_energy_sweep_f(negf_ctx* ctx, double energy, int eCtr, double deltaE, const double* mu,
                        double muT, int mask, double* T12, double* Modec1, double* Modec2,
                        double* DOSc1, double* DOSc2, double* EDensity, double* HDensity,
                        double* RhoN, double* RhoP, double* SigXExp, double* SigYExp,
                        double* SigZExp, double* ContactCurrent, double* Jx, double* Jy,
                        double* Jz, double* ERJx, double* ERJy, double* ERJz, double* ERRhon,
                        double* ORJx, double* ORJy, double* ORJz, double* ORRhon);
int negf_flush_f(negf_ctx* ctx, const double* mu, double muT, int mask, double* T12,
                 double* Modec1, double* Modec2, double* DOSc1, double* DOSc2, double* EDensity,
                 double* HDensity, double* RhoN, double* RhoP, double* SigXExp, double* SigYExp,
                 double* SigZExp, double* ContactCurrent, double* Jx, double* Jy, double* Jz,
                 double* ERJx, double* ERJy, double* ERJz, double* ERRhon, double* ORJx,
                 double* ORJy, double* ORJz, double* ORRhon);
int negf_observables_f(negf_ctx* ctx, int eCtr, double f1, double f2, double deltaE,
                       const double* Gnd, const double* Gpd, const double* Gnu, const double* Gnl,
                       const double* Gpu, const double* Gpl, const double* GRRgt,
                       const double* GammaL, const double* GammaR, int mask, double* T12,
                       double* Modec1, double* Modec2, double* DOSc1, double* DOSc2,
                       double* EDensity, double* HDensity, double* RhoN, double* RhoP,
                       double* SigXExp, double* SigYExp, double* SigZExp, double* ContactCurrent,
                       double* Jx, double* Jy, double* Jz);

/* --- L1 primitives --------------------------------------------------------- */

/* Invert_Matrix (MKL_Routines_Cluster.F90:2-33): invG = G^-1 for one n x n
 * host matrix; *info = first zero pivot (1-based) or 0. */
int negf_invert_matrix(int n, const double* G, double* invG, int* info);

/* zgemm (as called at Recursive_Greens.F90:18-95): C = alpha op(A) op(B) + beta C
 * for square n x n host matrices, op in 'N','T','C'. */
int negf_zgemm(char transa, char transb, int n, const double* alpha, const double* A, int lda,
               const double* B, int ldb, const double* beta, double* C, int ldc);

/* Parity checks of the sweep's kernels in each form, no counterpart in the
 * reference (the reference's single form is MKL zgemm / zgetrf+zgetri,
 * Recursive_Greens.F90:18-95, MKL_Routines_Cluster.F90:19,23).  `batch` n x n
 * column-major matrices, contiguous; complex64 != 0: COMPLEX(4) pairs of floats,
 * else COMPLEX(8); m3 != 0: the 3M complex product form (3 real products, normwise
 * accurate), else the 4-product form; m3 == 2 (complex64 only): the products on the
 * bf16 matrix cores with every FP32 operand split three ways (FP32 accuracy, each
 * component as accurate as in the 4-product form), the form NEGF_PREC_LU32 takes.
 *   negf_gemm_batch:   C_b = A_b B_b
 *   negf_invert_batch: invG_b = G_b^-1 (the sweep's Gauss-Jordan inverse; m3 selects
 *                      the form of its trailing updates); info[b] = first zero pivot
 *                      (1-based) or 0. */
int negf_gemm_batch(int complex64, int m3, int n, int batch, const void* A, const void* B, void* C);
int negf_invert_batch(int complex64, int m3, int n, int batch, const void* G, void* invG, int* info);
/*   negf_invert_batch_local: invG_b = G_b^-1 by block Gauss-Jordan with pivoting
 *                      inside the diagonal blocks only (128 columns in complex64, 64
 *                      in complex128), the inverse NEGF_PREC_LU32 takes for its RGF
 *                      blocks and its decimation.  No counterpart in the reference
 *                      (Invert_Matrix pivots over all rows): there are no row
 *                      interchanges between blocks, so it is accurate only when every
 *                      diagonal block met during the elimination is well conditioned,
 *                      and is meant for callers that check their results.  info[b] =
 *                      the first zero pivot inside a diagonal block (1-based) or 0;
 *                      the inverse is then undefined.  group (1..8; complex64 only,
 *                      else taken as 1): diagonal blocks eliminated against their own
 *                      columns before one rank-(128 group) update of the others --
 *                      the same arithmetic, longer products. */
int negf_invert_batch_local(int complex64, int m3, int group, int n, int batch, const void* G, void* invG,
                            int* info);

/* Allocation-free form of negf_invert_batch.  No counterpart in the reference; for
 * callers that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, hosts that forbid hipMalloc.  G, invG (batch contiguous n x n
 * matrices each) and info (batch ints) are DEVICE pointers; G is overwritten (the
 * elimination runs in place) and must not alias invG; the work is queued on
 * `stream` (a hipStream_t, NULL for the default stream) and the call returns
 * without synchronising.  The Gauss-Jordan workspace (per matrix a 128 x n panel
 * and 2n+97 ints) lives in a static 24 MiB device array, so batch * n is limited to
 * about 12000 for complex128 and 23000 for complex64 (larger requests fail with
 * NEGF_ERR_MEMORY); the call makes no hipMalloc/hipFree.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller. */
int negf_invert_batch_noalloc(int complex64, int m3, int n, int batch, void* G, void* invG,
                              int* info, void* stream);

/* --- Serving and batching of whole transport solves (no single counterpart in
 * the reference: it runs one wire per process, NEGF_Main.F90:21-38) ----------
 * A server owns one GPU.  Each request is one wire -- its own dimensions,
 * Hamiltonian (Htotd, HopX, gM) and energy list -- solved exactly as
 * negf_submit/negf_flush would solve it; several requests are in flight at once:
 *  - admission: a request is admitted when its resident footprint (Htotd and
 *    the coupling matrices on the device) fits the server's HBM budget next to
 *    the admitted ones; otherwise it waits (first come, first served);
 *  - continuous batching: the energies of admitted requests are cut into chunks
 *    that two device lanes take in arrival order, so requests that arrive while
 *    others run share the GPU as soon as memory allows;
 *  - each lane owns a fixed working pool, re-carved per chunk for that request's
 *    block size and length; a request whose single energy does not fit a pool is
 *    rejected at submission (NEGF_ERR_MEMORY);
 *  - results are accumulated per request in energy order, so a request's
 *    outputs are identical alone and in any mix.
 * The request's host arrays (Htotd, HopX, gM, energies) must stay valid, and its
 * outputs untouched, until negf_server_wait() returns for it. */
typedef struct negf_server negf_server;

typedef struct {
  double eta, sr_tol;
  int sr_max_iter, sr_extra_eta, precision, device, flags;
  double hbm_budget_gb; /* device memory for resident Hamiltonians + lane pools; 0 = 85% of free */
  double lane_pool_gb;  /* working pool per lane; 0 = half the budget split over the lanes */
  int lanes;            /* worker lanes (HIP streams), 1..4; 0 = 4 */
  double gamma_trunc;   /* as negf_config.gamma_trunc, for every request */
} negf_server_config;

typedef struct {
  int npx, npy, npz, norb;
  const double* Htotd; /* (n, n, npx) complex128 */
  const double* HopX;  /* (n, n) complex128 */
  const double* gM;    /* gMx, gMy, gMz (norb x norb each) complex128, or NULL */
  int n_energies;
  const double* energies;
  double deltaE;
  double mu[2], muT;
  negf_accum acc; /* outputs; per-energy arrays indexed 0 .. n_energies-1 */
  /* compact Hamiltonian (as negf_set_hamiltonian_compact), used when Htotd is NULL */
  const double* H0;     /* (n, n) complex128 slice template */
  const double* onsite; /* (npx, npy, npz) float64 site potentials */
} negf_request;

int negf_server_create(negf_server** srv, const negf_server_config* cfg);
/* Queue a request; *id identifies it.  Returns immediately. */
int negf_server_submit(negf_server* srv, const negf_request* req, long long* id);
/* 1 in *done when the request's outputs are complete. */
int negf_server_poll(negf_server* srv, long long id, int* done);
/* Block until the request completes; *status = its first nonzero per-energy status
 * (or a negative error code), latency_ms = submission to completion. */
int negf_server_wait(negf_server* srv, long long id, int* status, double* latency_ms);
/* Waits for every request, then releases the GPU resources. */
int negf_server_destroy(negf_server* srv);

/* --- Instrumentation (no counterpart in the reference) ---------------------
 * Time the MFMA GEMM launches (the dominant kernels) and the Gauss-Jordan panel
 * launches with HIP events on the stream each is launched on, between begin and
 * end: one launch in sample_every of each kind, in launch order (an event pair
 * per launch costs host time and stream markers, so a uniform sample keeps the
 * timed run close to an untimed one; 1 = every launch).  end() synchronises the
 * device and returns, per kind, the sampled launches' algorithmic work, their
 * summed duration in ms and their count; each argument is an array of 12:
 * [0..7] GEMMs indexed by 4 * (3M complex form) + 2 * (FP32) + (Gauss-Jordan
 * trailing update) -- [0] FP64 product, [1] FP64 update, [2] FP32 product,
 * [3] FP32 update, [4..7] the same in the 3M form -- with work = 8 m n k flops per
 * complex product; [8] FP64 and [9] FP32 Gauss-Jordan panels, with work = the
 * bytes the panel's columns and pivot rows move (read and written once);
 * [10] complex64 product and [11] update on the bf16 matrix cores (split form),
 * work = 8 m n k flops. */
int negf_profile_begin(int sample_every);
int negf_profile_end(double* flops, double* ms, long long* launches);

/* Instrumentation, no counterpart in the reference: the batched complex product
 * kernel of the sweep alone, C_b = A_b B_b for `batch` n x n matrices (complex64 if
 * complex64 != 0, else complex128; 3M form if m3 != 0), `reps` timed launches after
 * one untimed one; the operands are a constant byte pattern.  For hardware-counter
 * (rocprofv3 --pmc) runs of that kernel at the shape the sweep launches it with.
 * *ms_per_launch gets the mean launch time. */
int negf_bench_gemm(int device, int complex64, int n, int batch, int reps, int m3,
                    double* ms_per_launch);
/* The same with the workgroup tile fixed (0: 64x64, 1: 32x32), for the tile autotuner
 * (nonequilibriumgreensfunction_amd/autotune.py). */
int negf_bench_gemm_tile(int device, int complex64, int n, int batch, int reps, int m3, int tile,
                         double* ms_per_launch);
/* Tuning table of the complex product kernel (no counterpart in the reference): for
 * complex64 (complex64 != 0) or complex128 products, `count` rules (n_max[i],
 * tile[i]) with increasing n_max; a product whose larger dimension n is at most
 * n_max[i] (first such rule) runs with tile[i] (0: 64x64, 1: 32x32 workgroup
 * tiles), larger ones with 64x64.  Call before solves start.  The Python layer
 * loads the table the autotuner wrote. */
int negf_set_gemm_tiles(int complex64, int count, const int* n_max, const int* tile);
/* Lanes a context gets when negf_config.lanes is 0: the first rule with block size
 * n <= n_max[i] gives lanes[i] (1..4); no rule: 4.  Set by the autotuner's table. */
int negf_set_lane_table(int count, const int* n_max, const int* lanes);

#ifdef __cplusplus
}
#endif
#endif /* NEGF_CAPI_H */
