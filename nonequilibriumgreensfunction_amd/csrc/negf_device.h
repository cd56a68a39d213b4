// Shared device-side definitions for the gfx950 NEGF kernels.
//
// Complex numbers are interleaved (re, im) pairs -- the COMPLEX(8) / COMPLEX(4)
// memory layout of the Fortran host -- and every matrix is column-major.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace negf {

// vector types: double2 / float2 hold one complex number.
template <typename T> struct cplx;
template <> struct cplx<double> { typedef double2 type; };
template <> struct cplx<float> { typedef float2 type; };
template <typename T> using cx = typename cplx<T>::type;

template <typename T> __host__ __device__ inline cx<T> mk(T r, T i) { cx<T> z; z.x = r; z.y = i; return z; }
template <typename V> __host__ __device__ inline V cadd(V a, V b) { V z; z.x = a.x + b.x; z.y = a.y + b.y; return z; }
template <typename V> __host__ __device__ inline V csub(V a, V b) { V z; z.x = a.x - b.x; z.y = a.y - b.y; return z; }
template <typename V> __host__ __device__ inline V cmul(V a, V b) {
  V z; z.x = a.x * b.x - a.y * b.y; z.y = a.x * b.y + a.y * b.x; return z;
}
// a * conj(b)
template <typename V> __host__ __device__ inline V cmulc(V a, V b) {
  V z; z.x = a.x * b.x + a.y * b.y; z.y = a.y * b.x - a.x * b.y; return z;
}
template <typename V> __host__ __device__ inline V cconj(V a) { V z; z.x = a.x; z.y = -a.y; return z; }
template <typename V> __host__ __device__ inline V cneg(V a) { V z; z.x = -a.x; z.y = -a.y; return z; }
// 1/a = conj(a)/|a|^2 with one division: a is first scaled by the power of two
// 2^-e that brings max(|re|,|im|) into [1/2, 1) (exact), so |a|^2 can neither
// overflow nor underflow, and the result is scaled back by 2^-e.  Zero gives
// (inf, 0); inf or NaN parts give NaN.
template <typename V> __device__ inline V cinv(V a) {
  typedef decltype(a.x) R;
  const R s = fmax(fabs(a.x), fabs(a.y));
  V z;
  if (s == (R)0) { z.x = (R)1 / s; z.y = (R)0; return z; }
  if (!(s - s == (R)0)) { z.x = z.y = s - s; return z; }  // inf or NaN
  int e;
  (void)frexp(s, &e);
  const R ar = ldexp(a.x, -e), ai = ldexp(a.y, -e);
  const R rd = (R)1 / (ar * ar + ai * ai);
  z.x = ldexp(ar * rd, -e);
  z.y = ldexp(-ai * rd, -e);
  return z;
}

// Matrix operation selectors (BLAS op flags).
enum Op { OP_N = 0, OP_T = 1, OP_C = 2 };
// GEMM epilogues: standard alpha*AB + beta*C, or the Gauss-Jordan update.
enum Epi { EPI_STD = 0, EPI_GJ = 1 };

// A strided batch of complex GEMMs, C_b = alpha*op(A_b)*op(B_b) + beta*C_b, with
// X_b = X + b*sX (elements).  blockIdx.z selects b, through `active` if given.
// In the Gauss-Jordan form (EPI_GJ) the product is C_b = keep(C_b) + op(A_b)op(B_b)
// where keep() zeroes the rows r with pv_lo <= pivstep_b[r] < pv_hi and the
// columns in [skip_lo, skip_hi) are left untouched.
struct GemmBatch {
  const void* A;
  const void* B;
  void* C;
  long long sA, sB, sC;
  int m, n, k;
  int lda, ldb, ldc;
  const int* active;
  const int* pivstep;
  long long sPiv;
  int pv_lo, pv_hi;
  int skip_lo, skip_hi;
  int m3;  // 1: 3-product complex MFMA form (normwise-accurate, 3/4 of the work);
           // 0: 4-product form (each of Re, Im accurate relative to its own terms)
  // Optional per-matrix column count (plain products with beta 0 or 1): matrix b's
  // B is zero past column ncol[b * sNcol], so whole tiles past it are not
  // computed -- C is cleared there (beta 0) or left as it is (beta 1).
  const int* ncol = nullptr;
  long long sNcol = 0;
  float work = 1.0f;  // fraction of the m x n x k products really computed (timing only)
  // complex64 only, 64x64 tiles: the product on the bf16 matrix cores with each FP32
  // operand split three ways (zgemm_x3.hip; FP32 accuracy, four real products)
  int x3 = 0;
};

// Pivoted Cholesky factors Gam = W W' of both leads' broadening for a batch of
// energies (negf_kernels.hip): lead l's matrices at gam[l] + b*sG (overwritten
// with what the factor leaves out), factors at w[l] + b*sW (complex elements),
// state at st[l] + b*sSt (pchol_state_bytes(n) bytes).  Left in the state: the
// number of nonzero columns (int at st + pchol_rank_off(n)) and, at + 4, the
// truncation rank for tau: the fewest leading columns whose product leaves out a
// part of trace at most tau * tr(Gam) (the rank itself when tau <= 0).
struct PcholArgs {
  double2* gam[2];
  double2* w[2];
  char* st[2];
  long long sG, sW, sSt;
  int n;
  double tau;
};
// State: d[n] doubles, used[n] ints, then (8-byte aligned) three pivot keys, the
// stop threshold, tr(Gam), the rank, and two slots of per-workgroup partial sums
// of the remaining diagonal (at most 16 workgroups of 256 rows: n <= 4095).
constexpr int kPcholMaxBlocks = 16;
__host__ __device__ inline long long pchol_key_off(int n) { return ((12LL * n) + 7) & ~7LL; }
__host__ __device__ inline long long pchol_rank_off(int n) { return pchol_key_off(n) + 40; }
__host__ __device__ inline long long pchol_state_bytes(int n) {
  return pchol_key_off(n) + 48 + 2 * kPcholMaxBlocks * 8;
}
hipError_t launch_pchol(hipStream_t s, const PcholArgs& a, int batch);

// Spin matrices gMx, gMy, gMz (norb <= 4) as kernel arguments.
struct SpinMats {
  double re[3][4][4];  // [matrix][row][col]
  double im[3][4][4];
};

// A mirror of a lead's cross-section, P e(s,o) = i^ph[o] e(R s, perm[o]), with R the
// site reflection y -> Npy-1-y (axis 1) or z -> Npz-1-z (axis 2) of the slice's
// sites (site = y + Npy z, orbital fastest), perm an involution of the orbitals and
// i^ph[o] i^ph[perm[o]] = 1, so P is a Hermitian unitary involution.  With Npy
// (Npz) even no site lies on the mirror, and the +1 and -1 eigenspaces of P (the
// sectors) have the orthonormal bases v(a)+- = (e(r_a) +- i^ph e(m_a)) / sqrt 2:
// a = r norb + o runs over the representative half of the sites (y < Npy/2, r = y +
// Npy/2 z; or z < Npz/2, r = y + Npy z) and every orbital, r_a is that site and
// orbital, m_a = (R site, perm[o]).  A matrix M that commutes with P is block
// diagonal there: M+- (a, b) = M(r_a, r_b) +- i^ph_b M(r_a, m_b).
constexpr int kSymMaxOrb = 4;
struct SymMap {
  int axis;  // 0: no mirror, 1: y, 2: z
  int npy, npz, norb;
  int perm[kSymMaxOrb];
  int ph[kSymMaxOrb];  // power of i, 0..3
};
__host__ __device__ inline void sym_index(const SymMap& m, int a, int& ir, int& im, int& ph) {
  const int o = a % m.norb, r = a / m.norb;
  int y, z, ym, zm;
  if (m.axis == 1) {
    const int hy = m.npy / 2;
    y = r % hy; z = r / hy; ym = m.npy - 1 - y; zm = z;
  } else {
    y = r % m.npy; z = r / m.npy; ym = y; zm = m.npz - 1 - z;
  }
  ir = (y + m.npy * z) * m.norb + o;
  im = (ym + m.npy * zm) * m.norb + m.perm[o];
  ph = m.ph[o];
}
// i^k z (exact: a swap of parts and sign changes)
template <typename V> __host__ __device__ inline V rot_i(V z, int k) {
  V r;
  switch (k & 3) {
    case 0: r = z; break;
    case 1: r.x = -z.y; r.y = z.x; break;
    case 2: r.x = -z.x; r.y = -z.y; break;
    default: r.x = z.y; r.y = -z.x; break;
  }
  return r;
}

// Per-matrix workspace of the batched Gauss-Jordan inverse (gj_inverse.hip).
struct GjWork {
  void* prow;       // kGjNBO x n complex, row-major (row m at prow + m*n)
  int* pivstep;     // n: step at which each row became a pivot (n = not yet)
  int* pivrow;      // n: pivot row of each step
  int* info;        // 1: first zero pivot (1-based) or 0
  long long sProw;  // batch stride of prow (complex elements)
  long long sInt;   // batch stride of pivstep, pivrow and info (ints)
  int m3 = 0;       // complex product form of the trailing updates: 0 4-product, 1 3M,
                    // 2 (complex64) the bf16 three-way split products (GemmBatch::x3)
  void* tmp = nullptr;  // gj_invert_local: n x 128 (complex64) or n x 64 (complex128) per matrix,
  long long sTmp = 0;   //   n x 128 (1 + group) complex64 with group > 1
  int group = 1;        // gj_invert_local, complex64: diagonal blocks per delayed trailing update
};

template <typename T>
hipError_t gemm_batched(hipStream_t s, int opa, int opb, const GemmBatch& g, int batch,
                        cx<T> alpha, cx<T> beta);
// Complex64 products on the bf16 matrix cores (zgemm_x3.hip): op pairs as
// gemm_batched, epi EPI_GJ with (OP_N, OP_T) as gemm_gj_update.
hipError_t gemm_x3(hipStream_t s, int opa, int opb, int epi, const GemmBatch& g, int batch, float2 alpha,
                   float2 beta);
template <typename T>
hipError_t gj_invert(hipStream_t s, cx<T>* S, long long sS, int ld, cx<T>* out, long long sOut,
                     int ldo, int n, int batch, const GjWork& w, const int* active);
// In-place inverse by block Gauss-Jordan with pivoting inside diagonal blocks only
// (gj_inverse.hip; 128-column blocks in complex64, 64 in complex128: the diagonal
// block's inverse holds it in registers).  Workspace per matrix: GjWork::prow holds
// n x kb then kb x kb elements (kb = min(n, LocalBlock<T>::kb); stride sProw),
// GjWork::tmp n x kb (stride sTmp), pivstep n ints; info is not cleared.
template <typename T> struct LocalBlock { static constexpr int kb = sizeof(T) == sizeof(float) ? 128 : 64; };
template <typename T>
hipError_t gj_invert_local(hipStream_t s, cx<T>* S, long long sS, int ld, int n, int batch, const GjWork& w,
                           const int* active);
constexpr int kGjNBO = 128;  // Gauss-Jordan outer block width (columns per trailing update)
// Gauss-Jordan int workspace per matrix: pivstep[n], pivrow[n], info, then the
// panel's pivot candidates (8 per 256-row chunk, n <= 3072).
constexpr int kGjCandInts = 96;
inline long long gj_ints(int n) { return 2LL * n + 1 + kGjCandInts; }

// Opt-in timing of the MFMA GEMM launches (the dominant kernels) and of the
// Gauss-Jordan panel launches: when enabled, one launch in `every` of each kind is
// bracketed by HIP events on its own stream and its algorithmic work is recorded
// with its kind.  Kinds 0..7 = 4 * (3M form) + 2 * (FP32) + (Gauss-Jordan trailing
// update), work = 8 m n k flops per complex product (products with a transposed or
// adjoint operand are not sampled, so each kind is one kernel instance); kinds 8, 9 = the FP64 / FP32
// Gauss-Jordan panel, work = the bytes its columns move (read and written once);
// kinds 10, 11 = the complex64 bf16-split product and trailing update (8 m n k).
constexpr int kTimerKinds = 12;
constexpr int kTimerPanel64 = 8;
constexpr int kTimerX3 = 10;
struct GemmTimer {
  bool on = false;
  int every = 1;         // time one launch in `every` of each kind (uniform sample)
  unsigned long long seen[kTimerKinds] = {};
  size_t used = 0;
  hipEvent_t* ev = nullptr;  // 2 per launch
  double* flops = nullptr;
  int* kind = nullptr;
  size_t cap = 0;
};
GemmTimer& gemm_timer();
// Reserves an event pair for a sampled launch of `kind` with algorithmic `work`;
// false when this launch is not sampled (or timing is off).
bool timer_slot(double work, int kind, hipEvent_t* e0, hipEvent_t* e1);
// While on, this host thread's launches are not sampled (it is capturing a graph).
void timer_capture(bool on);
// GEMM workgroup tile per element type and block size (0: 64x64, 1: 32x32), from
// the tuning table set by negf_set_gemm_tiles; force_gemm_tile(t >= 0) overrides it
// for the calling thread (-1 restores the table).
int gemm_tile(int fp32, int n);
int set_gemm_tiles(int fp32, int count, const int* n_max, const int* tile);
void force_gemm_tile(int tile);

}  // namespace negf

extern "C" void negf_set_error(const char* fmt, ...);
