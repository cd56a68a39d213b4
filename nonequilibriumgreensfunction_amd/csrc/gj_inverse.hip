// Batched dense complex inverse: blocked Gauss-Jordan elimination with partial
// (row) pivoting, for gfx950.
//
// Replaces Invert_Matrix (MKL_Routines_Cluster.F90:2-33, ZGETRF + ZGETRI).
// Gauss-Jordan forms the inverse in n^3 complex multiply-adds, the same count
// as getrf + getri, but every step is a rank-k update of the whole matrix,
// which maps onto the MFMA GEMM kernel.
//
// Pivoting is implicit: rows are never swapped.  Column k is eliminated with
// pivot row p_k (largest |Re|+|Im| among rows not yet used, as izamax), and
// pivstep[p_k] = k, pivrow[k] = p_k are recorded.  In place, column k then holds
// column p_k of the accumulated row transform S, and after all n steps
//     inv(A)[a][b] = S[pivrow[a]][pivstep[b]],
// which the final unpermute kernel applies while copying to the output.
//
// Blocking.  Columns are processed in outer blocks of kNBO and inner panels of
// kNBI.  A panel kernel (one 512-thread workgroup per matrix, panel rows held in
// registers) factors kNBI columns and copies its pivot rows' values in the other
// columns of the outer block into Prow; a rank-kNBI update (vector kernel) then
// brings the rest of the outer block up to date.  After the whole outer block, a gather kernel copies
// the block's pivot rows in every column outside it into Prow, and a rank-kNBO
// update (the GEMM kernel's Gauss-Jordan epilogue) is applied to those columns:
//     A[i][j] <- [p(i) not in block] A[i][j] + sum_m S[i][m] Prow[m][j].
// Two launches per panel: under concurrent streams each dependent launch costs
// tens of microseconds of dispatch latency, more than the panel's own work.
#include "negf_device.h"

namespace negf {

template <typename T>
hipError_t gemm_gj_update(hipStream_t s, const GemmBatch& g, int batch);

constexpr int kNBI = 8;    // panel width
constexpr int kNBO = kGjNBO;  // outer block width (two GEMM tiles; the update skips its columns)

__global__ void k_gj_init(int* pivstep, int* pivrow, int* info, long long sInt, int n,
                          const int* active) {
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  int* ps = pivstep + (long long)b * sInt;
  int* pr = pivrow + (long long)b * sInt;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    ps[i] = n;
    pr[i] = 0;
  }
}

// Pivot ranking.  A candidate row i is ranked by a 64-bit key: the bit pattern of
// its magnitude |re|+|im| as a double (monotonic for non-negative values) with
// the low 12 bits replaced by 4095 - i, so the larger magnitude wins and ties go
// to the lower row (izamax's rule; magnitudes within 2^-40 relative count as
// ties).  Key 0 means no candidate (a real key is at least 4095 - 3071).
__device__ inline unsigned long long umax64(unsigned long long a, unsigned long long b) {
  return a > b ? a : b;
}
template <int CTRL>
__device__ inline unsigned long long dpp_u64(unsigned long long x) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xf, 0xf, false);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
// Max over the wave, in every lane: within each row of 16 lanes by DPP (quad
// xor 1, quad xor 2, half-row mirror, row mirror), then across rows by xor 16
// and xor 32 shuffles.
__device__ inline unsigned long long wave_max_u64(unsigned long long x) {
  x = umax64(x, dpp_u64<0xb1>(x));   // quad_perm [1,0,3,2]
  x = umax64(x, dpp_u64<0x4e>(x));   // quad_perm [2,3,0,1]
  x = umax64(x, dpp_u64<0x141>(x));  // row_half_mirror
  x = umax64(x, dpp_u64<0x140>(x));  // row_mirror
  x = umax64(x, __shfl_xor(x, 16));
  x = umax64(x, __shfl_xor(x, 32));
  return x;
}

// Two barriers per pivot column: (1) after each wave has published its best key,
// every thread reduces the NT/64 keys itself -- no broadcast barrier -- and the
// thread holding the pivot row publishes the row scaled by 1/pivot; (2) after
// that, every thread eliminates.  The LDS scratch alternates between two buffers
// by column parity, so column jj+1 never overwrites what a slower wave still
// reads for column jj.  512 threads (two waves per SIMD) hold R rows each.
template <typename T, int R, int NT>
__global__ __launch_bounds__(NT) void k_gj_panel(cx<T>* S, long long sS, int n, int ld,
                                                             int c0, int nb, int k0, int kb, GjWork w,
                                                             const int* active) {
  typedef cx<T> V;
  constexpr int NW = NT / 64;
  const int b = active ? active[blockIdx.x] : (int)blockIdx.x;
  V* A = S + (long long)b * sS;
  int* pivstep = w.pivstep + (long long)b * w.sInt;
  int* pivrow = w.pivrow + (long long)b * w.sInt;
  int* info = w.info + (long long)b * w.sInt;

  __shared__ unsigned long long red_k[2][NW];
  __shared__ V psc[2][kNBI + 1];  // pivot row scaled by 1/pivot; [kNBI] = 1/pivot

  const int tid = threadIdx.x, lane = tid & 63;
  V v[R][kNBI];
  bool used[R];
  int pp[kNBI];  // pivot row of each column (the same in every thread), -1 if none
#pragma unroll
  for (int j = 0; j < kNBI; ++j) pp[j] = -1;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid + r * NT;
    used[r] = true;
#pragma unroll
    for (int j = 0; j < kNBI; ++j) v[r][j] = mk<T>(0, 0);
    if (i < n) {
      used[r] = pivstep[i] < n;
#pragma unroll
      for (int j = 0; j < kNBI; ++j)
        if (j < nb) v[r][j] = A[(long long)(c0 + j) * ld + i];
    }
  }

#pragma unroll
  for (int jj = 0; jj < kNBI; ++jj) {
    if (jj < nb) {
      const int par = jj & 1;
      unsigned long long best = 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = tid + r * NT;
        if (i < n && !used[r]) {
          double m = (double)(fabs(v[r][jj].x) + fabs(v[r][jj].y));
          if (m != m) m = (double)INFINITY;
          best = umax64(best, ((unsigned long long)__double_as_longlong(m) & ~0xfffull) |
                                  (unsigned long long)(4095 - i));
        }
      }
      best = wave_max_u64(best);
      if (lane == 0) red_k[par][tid >> 6] = best;
      __syncthreads();  // (1)
      unsigned long long g = red_k[par][0];
#pragma unroll
      for (int q = 1; q < NW; ++q) g = umax64(g, red_k[par][q]);
      if (g == 0) continue;  // no unused row left (the same in every thread)
      const int p = 4095 - (int)(g & 0xfffull);
      pp[jj] = p;
      if (tid == 0) {
        if ((g & ~0xfffull) == 0 && *info == 0) *info = c0 + jj + 1;  // LAPACK-style INFO
        pivstep[p] = c0 + jj;
        pivrow[c0 + jj] = p;
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (tid + r * NT == p) {
          const V inv = cinv(v[r][jj]);
#pragma unroll
          for (int j = 0; j < kNBI; ++j) psc[par][j] = cmul(v[r][j], inv);
          psc[par][kNBI] = inv;
        }
      __syncthreads();  // (2)
      const V inv = psc[par][kNBI];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = tid + r * NT;
        if (i >= n) continue;
        if (i == p) {
#pragma unroll
          for (int j = 0; j < kNBI; ++j) v[r][j] = psc[par][j];
          v[r][jj] = inv;
          used[r] = true;
        } else {
          const V f = v[r][jj];
#pragma unroll
          for (int j = 0; j < kNBI; ++j) v[r][j] = csub(v[r][j], cmul(f, psc[par][j]));
          v[r][jj] = cneg(cmul(f, inv));
        }
      }
    }
  }
  // write the factored panel back
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid + r * NT;
    if (i < n) {
#pragma unroll
      for (int j = 0; j < kNBI; ++j)
        if (j < nb) A[(long long)(c0 + j) * ld + i] = v[r][j];
    }
  }
  // The pivot rows' values in the other columns of the outer block [k0, k0 + kb)
  // -> Prow rows c0 - k0 + m, for the rank-nb update of the block that follows.
  // Those columns are up to date (the previous update has completed) and this
  // kernel does not write them.
  V* prow = (V*)w.prow + (long long)b * w.sProw;
#pragma unroll
  for (int m = 0; m < kNBI; ++m) {
    const int p = pp[m];
    if (m < nb && p >= 0) {
      for (int jc = tid; jc < kb; jc += NT) {
        const int j = k0 + jc;
        if (j < c0 || j >= c0 + nb) prow[(long long)(c0 - k0 + m) * n + j] = A[(long long)j * ld + p];
      }
    }
  }
}

// Rank-nb update of the outer block after panel [c0, c0 + nb):
//     A[i][j] <- [p(i) not in panel] A[i][j] + sum_m A[i][c0 + m] Prow[c0 - k0 + m][j]
// for every column j of [k0, k0 + kb) outside the panel.  It is memory-bound
// (nb = 8), so it runs on the vector units: one thread per row, 16 columns per
// workgroup, the panel row in registers and the 8 x 16 pivot-row block in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_gj_update_block(cx<T>* S, long long sS, int n, int ld, int c0,
                                                         int nb, int k0, int kb, GjWork w,
                                                         const int* active) {
  typedef cx<T> V;
  constexpr int CW = 16;  // columns per workgroup
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  V* A = S + (long long)b * sS;
  const int* pivstep = w.pivstep + (long long)b * w.sInt;
  const V* prow = (const V*)w.prow + (long long)b * w.sProw;
  const int jc0 = blockIdx.y * CW;
  __shared__ V P[kNBI][CW];
  if (threadIdx.x < kNBI * CW) {
    const int m = threadIdx.x / CW, u = threadIdx.x % CW, jc = jc0 + u;
    P[m][u] = (m < nb && jc < kb) ? prow[(long long)(c0 - k0 + m) * n + k0 + jc] : mk<T>(0, 0);
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  V sv[kNBI];
#pragma unroll
  for (int m = 0; m < kNBI; ++m) sv[m] = m < nb ? A[(long long)(c0 + m) * ld + i] : mk<T>(0, 0);
  const int ps = pivstep[i];
  const bool keep = !(ps >= c0 && ps < c0 + nb);
#pragma unroll 4
  for (int u = 0; u < CW; ++u) {
    const int jc = jc0 + u, j = k0 + jc;
    if (jc >= kb) break;
    if (j >= c0 && j < c0 + nb) continue;
    V acc = keep ? A[(long long)j * ld + i] : mk<T>(0, 0);
#pragma unroll
    for (int m = 0; m < kNBI; ++m) acc = cadd(acc, cmul(sv[m], P[m][u]));
    A[(long long)j * ld + i] = acc;
  }
}

// Gathers the pivot rows of the outer block [k0, k0 + kb) in every column outside
// the block into Prow rows 0..kb-1, for the rank-kb update of those columns.  The
// columns outside the block are not modified while the block is factored, so one
// gather after its last panel reads the same values as one per panel would.  In a
// column-major matrix each pivot-row element is its own cache line, so the gather
// is spread over many small workgroups: (64 columns) x (8 pivot rows) x matrix.
template <typename T>
__global__ __launch_bounds__(64) void k_gj_gather(const cx<T>* S, long long sS, int n, int ld, int k0,
                                                  int kb, GjWork w, const int* active) {
  typedef cx<T> V;
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const V* A = S + (long long)b * sS;
  const int* pivrow = w.pivrow + (long long)b * w.sInt;
  V* prow = (V*)w.prow + (long long)b * w.sProw;
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int m0 = blockIdx.y * 8;
  if (j >= n || (j >= k0 && j < k0 + kb)) return;
  V t[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    t[u] = mk<T>(0, 0);
    if (m0 + u < kb) {
      const int p = pivrow[k0 + m0 + u];
      t[u] = A[(long long)j * ld + (p < n ? p : 0)];
    }
  }
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if (m0 + u < kb) prow[(long long)(m0 + u) * n + j] = t[u];
}

// out[a][b] = S[pivrow[a]][pivstep[b]]
template <typename T>
__global__ void k_gj_unpermute(const cx<T>* S, long long sS, int ld, cx<T>* out, long long sOut,
                               int ldo, GjWork w, int n, const int* active) {
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const cx<T>* A = S + (long long)b * sS;
  cx<T>* O = out + (long long)b * sOut;
  const int* pivstep = w.pivstep + (long long)b * w.sInt;
  const int* pivrow = w.pivrow + (long long)b * w.sInt;
  const int col = blockIdx.y;
  const int src = pivstep[col];
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
    const int pr = pivrow[a];
    O[(long long)col * ldo + a] = (src < n && pr < n) ? A[(long long)src * ld + pr]
                                                      : mk<T>((T)NAN, (T)NAN);
  }
}

// In-place batched inverse of the n x n matrices S_b (ld, stride sS); the result
// goes to out_b (ldo, stride sOut), which must not alias S_b.  A zero pivot sets
// info_b to its step (1-based) if info_b is still 0: callers clear info before a
// sequence of inversions and read the first failure after it.
template <typename T>
hipError_t gj_invert(hipStream_t s, cx<T>* S, long long sS, int ld, cx<T>* out, long long sOut,
                     int ldo, int n, int batch, const GjWork& w, const int* active) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (n > 3072) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gj_init, dim3((n + 255) / 256, batch), dim3(256), 0, s, w.pivstep, w.pivrow,
                     w.info, w.sInt, n, active);
  // panel rows live in registers: 512 threads x R rows, R = ceil(n / 512) <= 6
  constexpr int NT = 512;
  const int R = (n + NT - 1) / NT;
  for (int k0 = 0; k0 < n; k0 += kNBO) {
    const int kb = (n - k0 < kNBO) ? n - k0 : kNBO;
    for (int c0 = k0; c0 < k0 + kb; c0 += kNBI) {
      const int nb = (k0 + kb - c0 < kNBI) ? k0 + kb - c0 : kNBI;
#define NEGF_PANEL(RR)                                                                             \
  hipLaunchKernelGGL((k_gj_panel<T, RR, NT>), dim3(batch), dim3(NT), 0, s, S, sS, n, ld, c0, nb, k0, \
                     kb, w, active)
      // sampled timing: the panel's columns read and written, and its pivot rows'
      // values in the outer block copied out (nb x kb elements read and written)
      hipEvent_t ev0, ev1;
      const double bytes = 2.0 * sizeof(cx<T>) * nb * ((double)n + kb) * batch;
      const bool timed = timer_slot(bytes, kTimerPanel64 + (sizeof(T) == sizeof(float) ? 1 : 0), &ev0, &ev1);
      if (timed && hipEventRecord(ev0, s) != hipSuccess) return hipGetLastError();
      switch (R) {
        case 1: NEGF_PANEL(1); break;
        case 2: NEGF_PANEL(2); break;
        case 3: NEGF_PANEL(3); break;
        case 4: NEGF_PANEL(4); break;
        case 5: NEGF_PANEL(5); break;
        default: NEGF_PANEL(6); break;
      }
#undef NEGF_PANEL
      if (timed && hipEventRecord(ev1, s) != hipSuccess) return hipGetLastError();
      if (kb > nb)  // bring the rest of the outer block up to date (rank nb)
        hipLaunchKernelGGL((k_gj_update_block<T>), dim3((n + 255) / 256, (kb + 15) / 16, batch), dim3(256),
                           0, s, S, sS, n, ld, c0, nb, k0, kb, w, active);
    }
    if (n > kb) {  // rank-kb update of every column outside the outer block
      hipLaunchKernelGGL((k_gj_gather<T>), dim3((n + 63) / 64, (kb + 7) / 8, batch), dim3(64), 0, s, S,
                         sS, n, ld, k0, kb, w, active);
      GemmBatch g = {};
      g.A = S + (long long)k0 * ld; g.sA = sS; g.lda = ld;
      g.B = (const cx<T>*)w.prow; g.sB = w.sProw; g.ldb = n;
      g.C = S; g.sC = sS; g.ldc = ld;
      g.m = n; g.n = n; g.k = kb;
      g.active = active; g.pivstep = w.pivstep; g.sPiv = w.sInt;
      g.pv_lo = k0; g.pv_hi = k0 + kb;
      g.skip_lo = k0; g.skip_hi = k0 + kb;
      g.m3 = w.m3;
      hipError_t e = gemm_gj_update<T>(s, g, batch);
      if (e != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL((k_gj_unpermute<T>), dim3((n + 255) / 256, n, batch), dim3(256), 0, s, S, sS,
                     ld, out, sOut, ldo, w, n, active);
  return hipGetLastError();
}

template hipError_t gj_invert<double>(hipStream_t, cx<double>*, long long, int, cx<double>*,
                                      long long, int, int, int, const GjWork&, const int*);
template hipError_t gj_invert<float>(hipStream_t, cx<float>*, long long, int, cx<float>*,
                                     long long, int, int, int, const GjWork&, const int*);

}  // namespace negf
