// CPU oracle for the per-energy NEGF solve -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, in C++ complex128 with MKL BLAS/LAPACK (the same library
// family the reference links, Implement/Makefile:3), the operation sequence of
// the reference's Energy_Sweep path.  It is the checker the HIP library is
// compared against: only tests/, the smoke check and bench.py's CPU-baseline
// leg load it.  The product library never links or calls it.
//
// Deliberate differences from the reference (each a documented defect there):
//  * the Sancho-Rubio iteration counter is reset per call (the reference's is an
//    implicitly SAVE'd INTEGER, Contact_Self_Energy.F90:119, so its cap is
//    cumulative per process);
//  * ZGETRI gets a correctly sized workspace (MKL_Routines_Cluster.F90:12,23 pass
//    LWORK=N with a 512-entry buffer, undefined for N>512);
//  * eta, the SR tolerance and the SR iteration cap are parameters
//    (Constants.F90:7, Contact_Self_Energy.F90:154 hard-code 1e-4, 1e-8, 20000);
//  * bond currents accumulate in REAL(8) (the reference truncates through a
//    REAL(4) temporary, Calculate_Observables.F90:7).
//
// Complex arrays are interleaved (re, im) doubles, Fortran column-major, exactly
// like COMPLEX(8) module arrays: Htotd(i,j,x) sits at ((x*n)+j)*n+i (0-based).
#include <complex>
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>

typedef std::complex<double> cd;

extern "C" {
void zgemm_(const char*, const char*, const int*, const int*, const int*, const cd*,
            const cd*, const int*, const cd*, const int*, const cd*, cd*, const int*);
void zgetrf_(const int*, const int*, cd*, const int*, int*, int*);
void zgetri_(const int*, cd*, const int*, const int*, cd*, const int*, int*);
void MKL_Set_Num_Threads(int);
}

namespace {

const cd c0(0.0, 0.0), c1(1.0, 0.0), ci(0.0, 1.0);
const double kPi = 3.14159265358979323846;
// Physical constants exactly as Constants.F90:4-6 declares them.
const double kQ = 1.602e-19;
const double kH = 6.626e-34;
const double kHbar = kH / (2.0 * kPi);

typedef std::vector<cd> Mat;

// C = alpha*op(A)*op(B) + beta*C, square n, leading dims given.
void gemm(char ta, char tb, int n, cd alpha, const cd* A, int lda, const cd* B, int ldb,
          cd beta, cd* C, int ldc) {
  zgemm_(&ta, &tb, &n, &n, &n, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
}

// Invert_Matrix (MKL_Routines_Cluster.F90:2-33): ZGETRF then ZGETRI; returns INFO.
int invert(int n, const cd* G, cd* invG) {
  std::memcpy(invG, G, sizeof(cd) * (size_t)n * n);
  std::vector<int> ipiv(n);
  int info = 0;
  zgetrf_(&n, &n, invG, &n, ipiv.data(), &info);
  if (info != 0) return info;
  int lwork = -1;
  cd wq;
  zgetri_(&n, invG, &n, ipiv.data(), &wq, &lwork, &info);
  lwork = std::max(n, (int)wq.real());
  std::vector<cd> work(lwork);
  zgetri_(&n, invG, &n, ipiv.data(), work.data(), &lwork, &info);
  return info;
}

inline cd& at(cd* M, int n, int i, int j) { return M[(size_t)j * n + i]; }
inline const cd& at(const cd* M, int n, int i, int j) { return M[(size_t)j * n + i]; }

}  // namespace

extern "C" {

// Status bits shared with the HIP library's header.
enum {
  OR_OK = 0,
  OR_SR_NOT_CONVERGED = 1 << 20,
  OR_SR_NAN = 1 << 21,
};

void oracle_set_threads(int n) { MKL_Set_Num_Threads(n); }

int oracle_invert_matrix(int n, const double* G, double* invG) {
  return invert(n, (const cd*)G, (cd*)invG);
}

int oracle_zgemm(char ta, char tb, int n, double ar, double ai, const double* A, const double* B,
                 double br, double bi, double* C) {
  gemm(ta, tb, n, cd(ar, ai), (const cd*)A, n, (const cd*)B, n, cd(br, bi), (cd*)C, n);
  return 0;
}

// Sancho_Rubio (Contact_Self_Energy.F90:110-223).  Returns 0, a LAPACK INFO of a
// failed inverse (>0), or OR_SR_* bits.  *iters receives the loop count.
int oracle_sancho_rubio(int n, const double* alpha_, const double* beta_, double* Sigma_,
                        double eta, double tol, int max_iter, int extra_eta, int* iters) {
  const cd* alpha = (const cd*)alpha_;
  const cd* beta = (const cd*)beta_;
  cd* Sigma = (cd*)Sigma_;
  size_t nn = (size_t)n * n;
  Mat betad(nn), tmp(nn), t(nn), tt(nn), CapT(nn), Toldt(nn, c0), cmat1(nn), cmat2(nn),
      change(nn, c1), gfinal(nn), work(nn);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) at(betad.data(), n, i, j) = std::conj(at(beta, n, j, i));
  // :141  ginit = inv(alpha + i*eta*I)   (alpha already holds E+i*eta)
  std::memcpy(work.data(), alpha, sizeof(cd) * nn);
  if (extra_eta)
    for (int i = 0; i < n; ++i) at(work.data(), n, i, i) += cd(0.0, eta);
  int status = 0;
  int info = invert(n, work.data(), tmp.data());
  if (info) status = info;
  // :147, :149
  gemm('N', 'N', n, -c1, tmp.data(), n, betad.data(), n, c0, t.data(), n);
  gemm('N', 'N', n, -c1, tmp.data(), n, beta, n, c0, tt.data(), n);
  CapT = t;
  for (int i = 0; i < n; ++i) at(Toldt.data(), n, i, i) = c1;
  auto maxabs = [&](const Mat& M) {
    double m = 0.0;
    for (const cd& z : M) m = std::max(m, std::abs(z));
    return m;
  };
  int it = 0;
  double mc = maxabs(change);
  while (it < max_iter && mc > tol) {  // :154
    ++it;
    gemm('N', 'N', n, c1, Toldt.data(), n, tt.data(), n, c0, cmat1.data(), n);  // :157
    Toldt = cmat1;
    std::fill(cmat1.begin(), cmat1.end(), c0);  // :160-162
    for (int i = 0; i < n; ++i) at(cmat1.data(), n, i, i) = c1;
    gemm('N', 'N', n, -c1, t.data(), n, tt.data(), n, c1, cmat1.data(), n);
    gemm('N', 'N', n, -c1, tt.data(), n, t.data(), n, c1, cmat1.data(), n);
    info = invert(n, cmat1.data(), tmp.data());  // :163
    if (info) {
      *iters = it;
      return info;  // the reference STOPs here (:164-167)
    }
    gemm('N', 'N', n, c1, tmp.data(), n, t.data(), n, c0, cmat1.data(), n);  // :170-172
    gemm('N', 'N', n, c1, cmat1.data(), n, t.data(), n, c0, cmat2.data(), n);
    t = cmat2;
    gemm('N', 'N', n, c1, tmp.data(), n, tt.data(), n, c0, cmat1.data(), n);  // :174-176
    gemm('N', 'N', n, c1, cmat1.data(), n, tt.data(), n, c0, cmat2.data(), n);
    tt = cmat2;
    gemm('N', 'N', n, c1, Toldt.data(), n, t.data(), n, c0, change.data(), n);  // :178
    for (size_t k = 0; k < nn; ++k) CapT[k] += change[k];                   // :180
    for (size_t k = 0; k < nn; ++k)                                          // :182-191
      if (std::isnan(std::abs(change[k]))) {
        *iters = it;
        return OR_SR_NAN;
      }
    mc = maxabs(change);
  }
  *iters = it;
  if (mc > tol) status |= OR_SR_NOT_CONVERGED;
  // :195-197  g = inv(alpha + beta*CapT)
  gemm('N', 'N', n, c1, beta, n, CapT.data(), n, c0, cmat1.data(), n);
  for (size_t k = 0; k < nn; ++k) cmat1[k] += alpha[k];
  info = invert(n, cmat1.data(), gfinal.data());
  if (info) status = info;
  // :208-209  Sigma = beta*g*beta'
  gemm('N', 'N', n, c1, beta, n, gfinal.data(), n, c0, cmat1.data(), n);
  gemm('N', 'C', n, c1, cmat1.data(), n, beta, n, c0, Sigma, n);
  return status;
}

// Contact_Self_Energy, semi-infinite branch (Contact_Self_Energy.F90:68-101).
// H1, HL: the end slices; outputs Sigma/Gamma for both leads; dos[2] receives
// the DOSc1/DOSc2 increments; iters[2] the SR iteration counts.
int oracle_contact_self_energy(int n, const double* H1_, const double* HL_, const double* HopX_,
                               double energy, double eta, double tol, int max_iter, int extra_eta,
                               double* SigmaL, double* SigmaR, double* GammaL_, double* GammaR_,
                               double* dos, int* iters) {
  const cd* H1 = (const cd*)H1_;
  const cd* HL = (const cd*)HL_;
  const cd* HopX = (const cd*)HopX_;
  size_t nn = (size_t)n * n;
  Mat alpha(nn), beta(nn);
  cd ce(energy, eta);
  int st = 0;
  // left lead: alpha = (E+i eta)I - H(:,:,1), beta = -HopX'   (:71-76)
  for (size_t k = 0; k < nn; ++k) alpha[k] = -H1[k];
  for (int i = 0; i < n; ++i) at(alpha.data(), n, i, i) += ce;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) at(beta.data(), n, i, j) = -std::conj(at(HopX, n, j, i));
  int sL = oracle_sancho_rubio(n, (double*)alpha.data(), (double*)beta.data(), SigmaL, eta, tol,
                               max_iter, extra_eta, &iters[0]);
  // right lead: alpha = (E+i eta)I - H(:,:,Npx), beta = -HopX   (:83-88)
  for (size_t k = 0; k < nn; ++k) alpha[k] = -HL[k];
  for (int i = 0; i < n; ++i) at(alpha.data(), n, i, i) += ce;
  for (size_t k = 0; k < nn; ++k) beta[k] = -HopX[k];
  int sR = oracle_sancho_rubio(n, (double*)alpha.data(), (double*)beta.data(), SigmaR, eta, tol,
                               max_iter, extra_eta, &iters[1]);
  st = sL ? sL : sR;
  const cd* SL = (const cd*)SigmaL;
  const cd* SR = (const cd*)SigmaR;
  double d1 = 0.0, d2 = 0.0;
  for (int i = 0; i < n; ++i) {  // :94-97
    d1 -= at(SL, n, i, i).imag() / kPi;
    d2 -= at(SR, n, i, i).imag() / kPi;
  }
  dos[0] = d1;
  dos[1] = d2;
  cd* GL = (cd*)GammaL_;
  cd* GR = (cd*)GammaR_;
  for (int j = 0; j < n; ++j)  // :100-101
    for (int i = 0; i < n; ++i) {
      at(GL, n, i, j) = ci * (at(SL, n, i, j) - std::conj(at(SL, n, j, i)));
      at(GR, n, i, j) = ci * (at(SR, n, i, j) - std::conj(at(SR, n, j, i)));
    }
  return st;
}

// Energy_Sweep.F90:31-38: InvGRd(:,:,x) = (E+i eta)I - Htotd(:,:,x), minus the
// self-energies on the end slices.
void oracle_assemble_invg(int n, int L, const double* Htotd_, const double* SigmaL_,
                          const double* SigmaR_, double energy, double eta, double* InvGRd_) {
  const cd* H = (const cd*)Htotd_;
  cd* A = (cd*)InvGRd_;
  size_t nn = (size_t)n * n;
  for (size_t k = 0; k < nn * L; ++k) A[k] = -H[k];
  for (int x = 0; x < L; ++x)
    for (int i = 0; i < n; ++i) A[x * nn + (size_t)i * n + i] += cd(energy, eta);
  for (size_t k = 0; k < nn; ++k) {
    A[k] -= ((const cd*)SigmaL_)[k];
    A[(L - 1) * nn + k] -= ((const cd*)SigmaR_)[k];
  }
}

// Get_Retarded_Greens (Recursive_Greens.F90:2-107).  GRLef/GRRgt are (L*n) x n
// column-major (ld = L*n).  Returns the first nonzero INFO.
int oracle_retarded(int n, int L, const double* InvGRd_, const double* HopX_, double* GRd_,
                    double* GRu_, double* GRl_, double* GRLef_, double* GRRgt_) {
  const cd* InvG = (const cd*)InvGRd_;
  size_t nn = (size_t)n * n;
  int Nt = n * L;
  Mat mH(nn);
  for (size_t k = 0; k < nn; ++k) mH[k] = -((const cd*)HopX_)[k];  // "-HopX"
  cd* GRd = (cd*)GRd_;
  cd* GRu = (cd*)GRu_;
  cd* GRl = (cd*)GRl_;
  cd* GRLef = (cd*)GRLef_;
  cd* GRRgt = (cd*)GRRgt_;
  Mat gLR(nn * L), cmat1(nn), cmat2(nn), gRR(nn * L), colA(nn), colB(nn);
  int status = 0;
  for (int x = 0; x < L; ++x) {  // :16-34
    if (x > 0) {
      gemm('C', 'N', n, c1, mH.data(), n, &gLR[(x - 1) * nn], n, c0, cmat1.data(), n);
      gemm('N', 'N', n, c1, cmat1.data(), n, mH.data(), n, c0, cmat2.data(), n);
      for (size_t k = 0; k < nn; ++k) cmat1[k] = InvG[x * nn + k] - cmat2[k];
    } else {
      std::memcpy(cmat1.data(), InvG, sizeof(cd) * nn);
    }
    int info = invert(n, cmat1.data(), &gLR[x * nn]);
    if (info && !status) status = info;
  }
  std::memcpy(&GRd[(L - 1) * nn], &gLR[(L - 1) * nn], sizeof(cd) * nn);  // :37
  for (int x = L - 2; x >= 0; --x) {                                       // :38-51
    gemm('N', 'C', n, -c1, &GRd[(x + 1) * nn], n, mH.data(), n, c0, cmat1.data(), n);
    gemm('N', 'N', n, c1, cmat1.data(), n, &gLR[x * nn], n, c0, &GRl[x * nn], n);
    gemm('N', 'N', n, -c1, &gLR[x * nn], n, mH.data(), n, c0, cmat1.data(), n);
    gemm('N', 'N', n, c1, cmat1.data(), n, &GRd[(x + 1) * nn], n, c0, &GRu[x * nn], n);
    gemm('N', 'N', n, -c1, mH.data(), n, &GRl[x * nn], n, c0, cmat1.data(), n);
    for (int i = 0; i < n; ++i) at(cmat1.data(), n, i, i) += c1;
    gemm('N', 'N', n, c1, &gLR[x * nn], n, cmat1.data(), n, c0, &GRd[x * nn], n);
  }
  // :53-63 last block column (rows of GRRgt with ld = Nt)
  std::fill(GRRgt, GRRgt + (size_t)Nt * n, c0);
  auto put = [&](cd* col, int x, const cd* blk) {
    for (int j = 0; j < n; ++j)
      std::memcpy(&col[(size_t)j * Nt + (size_t)x * n], &blk[(size_t)j * n], sizeof(cd) * n);
  };
  auto get = [&](const cd* col, int x, cd* blk) {
    for (int j = 0; j < n; ++j)
      std::memcpy(&blk[(size_t)j * n], &col[(size_t)j * Nt + (size_t)x * n], sizeof(cd) * n);
  };
  put(GRRgt, L - 1, &GRd[(L - 1) * nn]);
  if (L >= 2) put(GRRgt, L - 2, &GRu[(L - 2) * nn]);
  for (int ii = 1; ii <= L - 2; ++ii) {
    int jj = L - 1 - ii;  // 1-based block jj; 0-based block jj-1
    gemm('N', 'N', n, -c1, &gLR[(jj - 1) * nn], n, mH.data(), n, c0, cmat1.data(), n);
    get(GRRgt, jj, colA.data());
    gemm('N', 'N', n, c1, cmat1.data(), n, colA.data(), n, c0, colB.data(), n);
    put(GRRgt, jj - 1, colB.data());
  }
  // :67-86 right-connected
  for (int x = L - 1; x >= 0; --x) {
    if (x < L - 1) {
      gemm('N', 'N', n, c1, mH.data(), n, &gRR[(x + 1) * nn], n, c0, cmat1.data(), n);
      gemm('N', 'C', n, c1, cmat1.data(), n, mH.data(), n, c0, cmat2.data(), n);
      for (size_t k = 0; k < nn; ++k) cmat1[k] = InvG[x * nn + k] - cmat2[k];
    } else {
      std::memcpy(cmat1.data(), &InvG[x * nn], sizeof(cd) * nn);
    }
    int info = invert(n, cmat1.data(), &gRR[x * nn]);
    if (info && !status) status = info;
  }
  // :88-96 first block column
  std::fill(GRLef, GRLef + (size_t)Nt * n, c0);
  put(GRLef, 0, &GRd[0]);
  if (L >= 2) put(GRLef, 1, &GRl[0]);
  for (int ii = 1; ii <= L - 1; ++ii) {
    gemm('N', 'C', n, -c1, &gRR[ii * nn], n, mH.data(), n, c0, cmat1.data(), n);
    get(GRLef, ii - 1, colA.data());
    gemm('N', 'N', n, c1, cmat1.data(), n, colA.data(), n, c0, colB.data(), n);
    put(GRLef, ii, colB.data());
  }
  return status;
}

// Calculate_Gnp (Calculate_Gnp.F90:1-136).  The loop nests are evaluated with
// ZGEMM (same sums, BLAS summation order).  RhoN/RhoP are (L, npy, npz)
// Fortran arrays and are accumulated (+=).
void oracle_gnp(int n, int L, int npy, int npz, int norb, double f1, double f2, double deltaE,
                const double* GRLef_, const double* GRRgt_, const double* GammaL_,
                const double* GammaR_, double* Gnd_, double* Gnu_, double* Gnl_, double* Gpd_,
                double* Gpu_, double* Gpl_, double* RhoN, double* RhoP) {
  size_t nn = (size_t)n * n;
  int Nt = n * L;
  const cd* cols[2] = {(const cd*)GRLef_, (const cd*)GRRgt_};
  const cd* gam[2] = {(const cd*)GammaL_, (const cd*)GammaR_};
  cd* Gd[2] = {(cd*)Gnd_, (cd*)Gpd_};
  cd* Gu[2] = {(cd*)Gnu_, (cd*)Gpu_};
  cd* Gl[2] = {(cd*)Gnl_, (cd*)Gpl_};
  double fac[2][2] = {{f1, f2}, {1.0 - f1, 1.0 - f2}};  // Energy_Sweep.F90:22-23
  Mat cmat1((size_t)Nt * n);
  for (int np = 0; np < 2; ++np) {
    std::fill(Gd[np], Gd[np] + nn * L, c0);
    if (L > 1) {
      std::fill(Gu[np], Gu[np] + nn * (L - 1), c0);
      std::fill(Gl[np], Gl[np] + nn * (L - 1), c0);
    }
    for (int lead = 0; lead < 2; ++lead) {
      const cd* G = cols[lead];
      // cmat1 = f * G * Gamma   (Ntot x n)
      cd alpha(fac[np][lead], 0.0);
      char N = 'N';
      zgemm_(&N, &N, &Nt, &n, &n, &alpha, G, &Nt, gam[lead], &n, &c0, cmat1.data(), &Nt);
      for (int x = 0; x < L; ++x) {
        // Gd_x += cmat1_x * G_x'
        char C = 'C';
        zgemm_(&N, &C, &n, &n, &n, &c1, &cmat1[(size_t)x * n], &Nt, &G[(size_t)x * n], &Nt, &c1,
               &Gd[np][x * nn], &n);
        if (x < L - 1) {
          zgemm_(&N, &C, &n, &n, &n, &c1, &cmat1[(size_t)x * n], &Nt, &G[(size_t)(x + 1) * n],
                 &Nt, &c1, &Gu[np][x * nn], &n);
          zgemm_(&N, &C, &n, &n, &n, &c1, &cmat1[(size_t)(x + 1) * n], &Nt, &G[(size_t)x * n],
                 &Nt, &c1, &Gl[np][x * nn], &n);
        }
      }
    }
  }
  const cd* Gnd = (const cd*)Gnd_;
  const cd* Gpd = (const cd*)Gpd_;
  int NN = npy * norb;
  for (int x = 0; x < L; ++x)  // :118-124
    for (int y = 0; y < npy; ++y)
      for (int z = 0; z < npz; ++z)
        for (int o = 0; o < norb; ++o) {
          int idx = o + y * norb + z * NN;
          size_t r = (size_t)x + (size_t)L * (y + (size_t)npy * z);
          RhoN[r] += Gnd[x * nn + (size_t)idx * n + idx].real() * (deltaE / (2.0 * kPi));
          RhoP[r] += Gpd[x * nn + (size_t)idx * n + idx].real() * (deltaE / (2.0 * kPi));
        }
}

// Calculate_Observables, non-BdG branches (Calculate_Observables.F90:9-60,85-169)
// with the spin block enabled.  gM holds gMx, gMy, gMz (norb x norb each,
// column-major).  perE[7] receives this energy's T12, Modec1, Modec2, EDensity,
// HDensity (accumulated +=); spatial arrays and ContactCurrent accumulate.
void oracle_observables(int n, int L, int npy, int npz, int norb, double f1, double f2,
                        double deltaE, const double* Htotd_, const double* HopX_,
                        const double* Gnd_, const double* Gpd_, const double* Gnu_,
                        const double* Gnl_, const double* Gpu_, const double* Gpl_,
                        const double* GRRgt_, const double* GammaL_, const double* GammaR_,
                        const double* gM_, double* perE, double* SigX, double* SigY, double* SigZ,
                        double* Jx, double* Jy, double* Jz, double* ContactCurrent) {
  size_t nn = (size_t)n * n;
  int Nt = n * L;
  int NN = npy * norb;
  const cd* Gnd = (const cd*)Gnd_;
  const cd* Gpd = (const cd*)Gpd_;
  const cd* gM = (const cd*)gM_;
  const double w = deltaE / (2.0 * kPi);
  double* Sig[3] = {SigX, SigY, SigZ};
  for (int x = 0; x < L; ++x)  // :10-30
    for (int y = 0; y < npy; ++y)
      for (int z = 0; z < npz; ++z) {
        int idx = y * norb + z * NN;
        size_t r = (size_t)x + (size_t)L * (y + (size_t)npy * z);
        for (int ii = 0; ii < norb; ++ii)
          for (int jj = 0; jj < norb; ++jj) {
            cd g = Gnd[x * nn + (size_t)(idx + jj) * n + idx + ii];
            for (int a = 0; a < 3; ++a)
              Sig[a][r] += (w * g * gM[a * norb * norb + ii * norb + jj]).real();
          }
      }
  double ed = 0.0, hd = 0.0;  // :34-43
  for (int x = 0; x < L; ++x)
    for (int y = 0; y < npy; ++y)
      for (int z = 0; z < npz; ++z)
        for (int ii = 0; ii < norb; ++ii) {
          int idx = ii + y * norb + z * NN;
          ed += Gnd[x * nn + (size_t)idx * n + idx].real() / (2.0 * kPi);
          hd += Gpd[x * nn + (size_t)idx * n + idx].real() / (2.0 * kPi);
        }
  const cd* GL = (const cd*)GammaL_;
  const cd* GR = (const cd*)GammaR_;
  double m1 = 0.0, m2 = 0.0;  // :48-53
  for (int ii = 0; ii < n; ++ii)
    for (int kk = 0; kk < n; ++kk) {
      m1 += (at(GL, n, ii, kk) * (Gnd[(size_t)ii * n + kk] + Gpd[(size_t)ii * n + kk])).real();
      m2 += (at(GR, n, ii, kk) *
             (Gnd[(L - 1) * nn + (size_t)ii * n + kk] + Gpd[(L - 1) * nn + (size_t)ii * n + kk]))
                .real();
    }
  // :86-100  T12 = Re tr(GammaL*G1L*GammaR*G1L')  with G1L = GRRgt(1:n,:)
  const cd* Rg = (const cd*)GRRgt_;
  Mat cm1(nn), cm2(nn), g1L(nn);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) g1L[(size_t)j * n + i] = Rg[(size_t)j * Nt + i];
  gemm('N', 'N', n, c1, GL, n, g1L.data(), n, c0, cm1.data(), n);
  gemm('N', 'C', n, c1, GR, n, g1L.data(), n, c0, cm2.data(), n);
  double T = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) T += (cm1[(size_t)j * n + i] * cm2[(size_t)i * n + j]).real();
  perE[0] += T;
  perE[1] += m1;
  perE[2] += m2;
  perE[3] += ed;
  perE[4] += hd;
  ContactCurrent[0] += kQ * deltaE / kH * (T * (f2 - f1));  // :102-103
  ContactCurrent[1] += kQ * deltaE / kH * (T * (f1 - f2));
  // :111-168 bond currents
  const cd* H = (const cd*)Htotd_;
  const cd* HopX = (const cd*)HopX_;
  const cd* Gnu = (const cd*)Gnu_;
  const cd* Gnl = (const cd*)Gnl_;
  const cd* Gpu = (const cd*)Gpu_;
  const cd* Gpl = (const cd*)Gpl_;
  const cd pref = ci * kQ * kQ / kHbar;
  auto E = [&](const cd* M, int x, int i, int j) { return M[x * nn + (size_t)j * n + i]; };
  for (int x = 0; x < L; ++x)
    for (int y = 0; y < npy; ++y)
      for (int z = 0; z < npz; ++z) {
        int idx = y * norb + z * NN;
        size_t r = (size_t)x + (size_t)L * (y + (size_t)npy * z);
        for (int ii = 0; ii < norb; ++ii)
          for (int jj = 0; jj < norb; ++jj) {
            if (y < npy - 1) {
              int a = idx + ii, b = idx + norb + jj;
              double d = (pref * (E(H, x, a, b) * E(Gnd, x, b, a) - E(Gnd, x, a, b) * E(H, x, b, a) -
                                  E(H, x, a, b) * E(Gpd, x, b, a) + E(Gpd, x, a, b) * E(H, x, b, a)))
                             .real();
              Jy[r] += d * w;
            }
            if (z < npz - 1) {
              int a = idx + ii, b = idx + NN + jj;
              double d = (pref * (E(H, x, a, b) * E(Gnd, x, b, a) - E(Gnd, x, a, b) * E(H, x, b, a) -
                                  E(H, x, a, b) * E(Gpd, x, b, a) + E(Gpd, x, a, b) * E(H, x, b, a)))
                             .real();
              Jz[r] += d * w;
            }
            if (x < L - 1) {
              int a = idx + ii, b = idx + jj;
              cd hx = HopX[(size_t)b * n + a];
              double d = (pref * (hx * E(Gnl, x, b, a) - E(Gnu, x, a, b) * std::conj(hx) -
                                  hx * E(Gpl, x, b, a) + E(Gpu, x, a, b) * std::conj(hx)))
                             .real();
              Jx[r] += d * w;
            }
          }
      }
}

// Energy_Sweep (Energy_Sweep.F90:1-46) for one energy: occupations, contacts,
// InvGRd, RGF, Gn/Gp and observables.  All spatial arrays are (L,npy,npz)
// Fortran order and accumulate.  perE[7] = T12, Modec1, Modec2, EDensity,
// HDensity, DOSc1, DOSc2 (accumulated).  iters[2] = SR iteration counts.
int oracle_energy_sweep(int npx, int npy, int npz, int norb, const double* Htotd,
                        const double* HopX, const double* gM, double energy, double deltaE,
                        double mu1, double mu2, double muT, double eta, double sr_tol,
                        int sr_max_iter, int sr_extra_eta, double* perE, double* RhoN,
                        double* RhoP, double* SigX, double* SigY, double* SigZ, double* Jx,
                        double* Jy, double* Jz, double* ContactCurrent, int* iters) {
  int n = npy * npz * norb;
  int L = npx;
  size_t nn = (size_t)n * n;
  double f1 = (energy <= muT - mu1) ? 1.0 : 0.0;  // :11-23
  double f2 = (energy <= muT - mu2) ? 1.0 : 0.0;
  Mat SL(nn), SR(nn), GL(nn), GR(nn);
  double dos[2];
  int st = oracle_contact_self_energy(n, Htotd, Htotd + 2 * nn * (L - 1), HopX, energy, eta, sr_tol,
                                      sr_max_iter, sr_extra_eta, (double*)SL.data(),
                                      (double*)SR.data(), (double*)GL.data(), (double*)GR.data(),
                                      dos, iters);
  perE[5] += dos[0];
  perE[6] += dos[1];
  size_t Lm = (size_t)std::max(L - 1, 1);
  Mat InvG(nn * L), GRd(nn * L), GRu(nn * Lm), GRl(nn * Lm), GRLef(nn * L), GRRgt(nn * L);
  oracle_assemble_invg(n, L, Htotd, (double*)SL.data(), (double*)SR.data(), energy, eta,
                       (double*)InvG.data());
  int s2 = oracle_retarded(n, L, (double*)InvG.data(), HopX, (double*)GRd.data(),
                           (double*)GRu.data(), (double*)GRl.data(), (double*)GRLef.data(),
                           (double*)GRRgt.data());
  if (!st) st = s2;
  InvG.clear();
  InvG.shrink_to_fit();
  GRd.clear();
  GRd.shrink_to_fit();
  Mat Gnd(nn * L), Gpd(nn * L), Gnu(nn * Lm), Gnl(nn * Lm), Gpu(nn * Lm), Gpl(nn * Lm);
  oracle_gnp(n, L, npy, npz, norb, f1, f2, deltaE, (double*)GRLef.data(), (double*)GRRgt.data(),
             (double*)GL.data(), (double*)GR.data(), (double*)Gnd.data(), (double*)Gnu.data(),
             (double*)Gnl.data(), (double*)Gpd.data(), (double*)Gpu.data(), (double*)Gpl.data(),
             RhoN, RhoP);
  oracle_observables(n, L, npy, npz, norb, f1, f2, deltaE, Htotd, HopX, (double*)Gnd.data(),
                     (double*)Gpd.data(), (double*)Gnu.data(), (double*)Gnl.data(),
                     (double*)Gpu.data(), (double*)Gpl.data(), (double*)GRRgt.data(),
                     (double*)GL.data(), (double*)GR.data(), gM, perE, SigX, SigY, SigZ, Jx, Jy,
                     Jz, ContactCurrent);
  return st;
}

}  // extern "C"
