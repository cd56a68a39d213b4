// Batched dense complex inverse: blocked Gauss-Jordan elimination with partial
// (row) pivoting, for gfx950.
//
// Replaces Invert_Matrix (MKL_Routines_Cluster.F90:2-33, ZGETRF + ZGETRI).
// Gauss-Jordan forms the inverse in n^3 complex multiply-adds, the same count
// as getrf + getri, but every step is a rank-k update of the whole matrix,
// which maps onto the MFMA GEMM kernel.
//
// Pivoting is implicit: rows are never swapped.  Column k is eliminated with
// pivot row p_k (largest |Re|+|Im| among the candidate rows not yet used, as
// izamax; see the blocking below), and
// pivstep[p_k] = k, pivrow[k] = p_k are recorded.  In place, column k then holds
// column p_k of the accumulated row transform S, and after all n steps
//     inv(A)[a][b] = S[pivrow[a]][pivstep[b]],
// which the final unpermute kernel applies while copying to the output.
//
// Blocking.  Columns are processed in outer blocks of kNBO and panels of kNBI.
// The pivots of a panel are chosen by tournament (two-level) partial pivoting,
// so that no step needs one workgroup to see the whole column:
//   candidates -- every 256-row chunk runs partial pivoting (LU steps) on its own
//                 unused rows of the panel and proposes up to kNBI rows (one
//                 workgroup per chunk and matrix);
//   final      -- one small workgroup per matrix runs partial pivoting on the
//                 proposed rows' current values, which fixes the panel's kNBI
//                 pivot rows p_m in step order, runs the panel's column-by-column
//                 Gauss-Jordan steps on those rows alone (each step's pivot-row
//                 values are what every other row's update needs) and copies the
//                 pivot rows' values in the outer block's other columns into Prow;
//   update     -- every row takes the panel's steps with those values (the same
//                 operations, in the same order, as column-by-column elimination
//                 of the whole panel) and then updates the outer block's other
//                 columns, A[i][j] <- [i not a pivot of the panel] A[i][j] + sum_m S[i][m] Prow[m][j].
// With the pivots chosen among all rows (n <= 256: one chunk) this is exactly
// column-by-column partial-pivoting elimination, bit for bit.  After the
// whole outer block, a gather kernel copies the block's pivot rows in every column
// outside it into Prow, and a rank-kNBO update (the GEMM kernel's Gauss-Jordan
// epilogue) is applied to those columns:
//     A[i][j] <- [p(i) not in block] A[i][j] + sum_m S[i][m] Prow[m][j].
// Every kernel of a panel step is many small workgroups, so under concurrent
// streams they find room beside other work on the compute units.
#include "negf_device.h"

namespace negf {

template <typename T>
hipError_t gemm_gj_update(hipStream_t s, const GemmBatch& g, int batch);

constexpr int kNBI = 8;    // panel width
constexpr int kCandSlots = 2;  // candidates per lane of the pivot-selection wave
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

// Loads whose address is always valid (a clamped row or column) and whose value is
// selected afterwards: unconditional loads are issued back to back.
template <typename V>
__device__ inline V load_sel(const V* p, bool keep) {
  const V t = *p;
  V z;
  z.x = keep ? t.x : 0;
  z.y = keep ? t.y : 0;
  return z;
}

// Pivot key of an element of row i (see above).
template <typename V>
__device__ inline unsigned long long pivot_key(V x, int i) {
  double m = (double)(fabs(x.x) + fabs(x.y));
  if (m != m) m = (double)INFINITY;
  return ((unsigned long long)__double_as_longlong(m) & ~0xfffull) | (unsigned long long)(4095 - i);
}

// Candidate pivots of one 256-row chunk (one row per thread): partial pivoting on
// the chunk's unused rows of the panel [c0, c0 + nb), the later columns updated as
// the Gauss-Jordan steps update them (v[j] -= v[m] * pivot row[j] / pivot).  Two barriers
// per column; the LDS scratch alternates by column parity.  Writes the chosen
// rows (global indices, -1 once the chunk has no unused row left) to the
// candidate slots of the chunk.
template <typename T>
__global__ __launch_bounds__(256) void k_gj_cand(const cx<T>* S, long long sS, int n, int ld, int c0,
                                                 int nb, GjWork w, const int* active) {
  typedef cx<T> V;
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  const V* A = S + (long long)b * sS;
  const int* pivstep = w.pivstep + (long long)b * w.sInt;
  int* cand = w.info + (long long)b * w.sInt + 1 + blockIdx.x * kNBI;
  const int tid = threadIdx.x, lane = tid & 63;
  const int i = blockIdx.x * 256 + tid;
  __shared__ unsigned long long red[2][4];
  __shared__ V prs[2][kNBI];
  V v[kNBI];
  bool avail = i < n && pivstep[i] >= n;
  if (tid >= nb && tid < kNBI) cand[tid] = -1;  // slots past a short last panel
#pragma clang loop unroll(full)
  for (int j = 0; j < kNBI; ++j)
    v[j] = load_sel(A + (long long)(c0 + (j < nb ? j : 0)) * ld + (i < n ? i : 0), i < n && j < nb);
  // predicated rather than broken off: full unrolling keeps v[] in registers
  bool more = true;  // the chunk still has unused rows (the same in every thread)
#pragma clang loop unroll(full)
  for (int jj = 0; jj < kNBI; ++jj) {
    if (jj < nb && more) {
      const int par = jj & 1;
      unsigned long long k = 0;
      if (avail) k = pivot_key(v[jj], i);
      k = wave_max_u64(k);
      if (lane == 0) red[par][tid >> 6] = k;
      __syncthreads();
      const unsigned long long g = umax64(umax64(red[par][0], red[par][1]), umax64(red[par][2], red[par][3]));
      if (g == 0) {
        if (tid < kNBI && tid >= jj) cand[tid] = -1;
        more = false;
      } else {
        const int q = 4095 - (int)(g & 0xfffull);
        if (tid == 0) cand[jj] = q;
        if (i == q) {  // the pivot row scaled by 1 / pivot
          const V inv = cinv(v[jj]);
#pragma clang loop unroll(full)
          for (int j = 0; j < kNBI; ++j) prs[par][j] = cmul(v[j], inv);
          avail = false;
        }
        __syncthreads();
        if (avail && (g & ~0xfffull) != 0) {  // the later columns, as the Gauss-Jordan step does them
          const V f = v[jj];
#pragma clang loop unroll(full)
          for (int j = 0; j < kNBI; ++j)
            if (j > jj) v[j] = csub(v[j], cmul(f, prs[par][j]));
        }
      }
    }
  }
}

// The panel's pivots (one wave per matrix, kCandSlots candidates per lane, at most
// 12 x kNBI): partial pivoting, as above, on the candidates' current values, ties
// to the lower row; records pivstep/pivrow and INFO.  Then the column-by-column
// Gauss-Jordan steps of the panel restricted to its pivot rows: step m takes pivot
// row p_m as the earlier steps left it, v, and records
//     T[m][j] = v[j] / v[m] (j != m),  T[m][m] = 1 / v[m]
// -- the pivot row's values after its own step, the same for every row's update --
// in the panel's own columns of Prow, Prow[c0 - k0 + m][c0 + j] = T[m][j].  The
// pivot rows' values in the outer block's other columns go to Prow[c0 - k0 + m][j].
// (A one-wave workgroup: its barriers only order the LDS traffic.)
template <typename T>
__global__ __launch_bounds__(64) void k_gj_pivots(const cx<T>* S, long long sS, int n, int ld, int c0,
                                                  int nb, int k0, int kb, int nchunks, GjWork w,
                                                  const int* active) {
  typedef cx<T> V;
  const int b = active ? active[blockIdx.x] : (int)blockIdx.x;
  const V* A = S + (long long)b * sS;
  int* pivstep = w.pivstep + (long long)b * w.sInt;
  int* pivrow = w.pivrow + (long long)b * w.sInt;
  int* info = w.info + (long long)b * w.sInt;
  const int* cand = info + 1;
  V* prow = (V*)w.prow + (long long)b * w.sProw;
  const int tid = threadIdx.x;
  __shared__ V prs[2][kNBI];
  __shared__ int piv[kNBI];
  __shared__ V Tm[kNBI][kNBI];
  const int ncand = nchunks * kNBI;
  int rr[kCandSlots];
  bool av[kCandSlots];
  V v[kCandSlots][kNBI];
#pragma clang loop unroll(full)
  for (int q = 0; q < kCandSlots; ++q) {
    const int idx = tid + 64 * q;
    rr[q] = idx < ncand ? cand[idx] : -1;
    av[q] = rr[q] >= 0;
#pragma clang loop unroll(full)
    for (int j = 0; j < kNBI; ++j)
      v[q][j] = load_sel(A + (long long)(c0 + (j < nb ? j : 0)) * ld + (av[q] ? rr[q] : 0), av[q] && j < nb);
  }
  // compile-time trip counts keep v[][] and u[] in registers
#pragma clang loop unroll(full)
  for (int jj = 0; jj < kNBI; ++jj) {
    if (jj < nb) {
      const int par = jj & 1;
      unsigned long long k = 0;
#pragma clang loop unroll(full)
      for (int q = 0; q < kCandSlots; ++q)
        if (av[q]) k = umax64(k, pivot_key(v[q][jj], rr[q]));
      const unsigned long long g = wave_max_u64(k);
      // g == 0 cannot happen: every chunk proposes min(nb, its unused rows) rows, the
      // panel has nb <= unused rows, and every candidate's key is positive (row bits);
      // row 0 keeps a broken call in bounds
      const int p = g ? 4095 - (int)(g & 0xfffull) : 0;
      if (tid == 0) {
        piv[jj] = p;
        if ((g & ~0xfffull) == 0 && *info == 0) *info = c0 + jj + 1;  // LAPACK-style INFO
      }
#pragma clang loop unroll(full)
      for (int q = 0; q < kCandSlots; ++q)
        if (av[q] && rr[q] == p) {
          const V inv = cinv(v[q][jj]);
#pragma clang loop unroll(full)
          for (int j = 0; j < kNBI; ++j) prs[par][j] = cmul(v[q][j], inv);
          av[q] = false;
        }
      __syncthreads();
      if ((g & ~0xfffull) != 0) {
#pragma clang loop unroll(full)
        for (int q = 0; q < kCandSlots; ++q)
          if (av[q]) {
            const V f = v[q][jj];
#pragma clang loop unroll(full)
            for (int j = 0; j < kNBI; ++j)
              if (j > jj) v[q][j] = csub(v[q][j], cmul(f, prs[par][j]));
          }
      }
    }
  }
  __syncthreads();
  if (tid < nb) {
    pivstep[piv[tid]] = c0 + tid;
    pivrow[c0 + tid] = piv[tid];
  }
  // the pivot rows' values in the outer block's other columns (a column per lane
  // and pass; all loads of a column before its stores)
  for (int jc = tid; jc < kb; jc += 64) {
    const int j = k0 + jc;
    if (j >= c0 && j < c0 + nb) continue;
    V t[kNBI];
#pragma clang loop unroll(full)
    for (int m = 0; m < kNBI; ++m) t[m] = load_sel(A + (long long)j * ld + piv[m < nb ? m : 0], m < nb);
#pragma clang loop unroll(full)
    for (int m = 0; m < kNBI; ++m)
      if (m < nb) prow[(long long)(c0 - k0 + m) * n + j] = t[m];
  }
  // Gauss-Jordan steps on the pivot rows: lane l < nb follows row p_l
  V u[kNBI];
#pragma clang loop unroll(full)
  for (int j = 0; j < kNBI; ++j)
    u[j] = load_sel(A + (long long)(c0 + (j < nb ? j : 0)) * ld + piv[tid < nb ? tid : 0], tid < nb && j < nb);
#pragma clang loop unroll(full)
  for (int m = 0; m < kNBI; ++m) {
    if (m < nb) {
      if (tid == m) {
        const V inv = cinv(u[m]);
#pragma clang loop unroll(full)
        for (int j = 0; j < kNBI; ++j) Tm[m][j] = cmul(u[j], inv);
        Tm[m][m] = inv;
      }
      __syncthreads();
      if (tid > m && tid < nb) {  // later pivot rows take step m like any other row
        const V f = u[m];
#pragma clang loop unroll(full)
        for (int j = 0; j < kNBI; ++j) u[j] = csub(u[j], cmul(f, Tm[m][j]));
        u[m] = cneg(cmul(f, Tm[m][m]));
      }
    }
  }
  for (int e = tid; e < nb * nb; e += 64) {
    const int m = e / nb, j = e % nb;
    prow[(long long)(c0 - k0 + m) * n + c0 + j] = Tm[m][j];
  }
}

// The panel's Gauss-Jordan steps on every row (32 rows x 8 column groups per
// workgroup), with the pivot rows' step values T from Prow: for m = 0..nb-1, row
// p_m becomes T[m]; every other row takes f = v[m], v[j] -= f T[m][j], v[m] = -f T[m][m].
// The result S[i][0..nb) goes in place of the panel, then the outer block's other
// columns are updated,
//     A[i][j] <- [i not a pivot of the panel] A[i][j] + sum_m S[i][m] Prow[c0 - k0 + m][j].
// A workgroup owns whole rows, so it reads the panel's old values before it writes
// the new ones; the pivot rows' old values come from Prow.
template <typename T>
__global__ __launch_bounds__(256) void k_gj_apply(cx<T>* S, long long sS, int n, int ld, int c0, int nb,
                                                  int k0, int kb, GjWork w, const int* active) {
  typedef cx<T> V;
  constexpr int RW = 32, CG = 8;
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  V* A = S + (long long)b * sS;
  const int* pivstep = w.pivstep + (long long)b * w.sInt;
  const V* prow = (const V*)w.prow + (long long)b * w.sProw;
  __shared__ V Ps[kNBI][kGjNBO];
  for (int e = threadIdx.x; e < kNBI * kb; e += 256) {
    const int m = e / kb, jc = e % kb;
    Ps[m][jc] = m < nb ? prow[(long long)(c0 - k0 + m) * n + k0 + jc] : mk<T>(0, 0);
  }
  __syncthreads();
  const int i = blockIdx.x * RW + (threadIdx.x % RW), g = threadIdx.x / RW;
  const bool live = i < n;
  const int pc = c0 - k0;  // the panel's first column within the block
  V v[kNBI];
  bool keep = true;
  if (live) {
    const int ps = pivstep[i];
    keep = !(ps >= c0 && ps < c0 + nb);
#pragma clang loop unroll(full)
    for (int j = 0; j < kNBI; ++j) v[j] = load_sel(A + (long long)(c0 + (j < nb ? j : 0)) * ld + i, j < nb);
#pragma clang loop unroll(full)
    for (int m = 0; m < kNBI; ++m) {
      if (m < nb) {
        if (ps == c0 + m) {
#pragma clang loop unroll(full)
          for (int j = 0; j < kNBI; ++j) v[j] = Ps[m][pc + j];
        } else {
          const V f = v[m];
#pragma clang loop unroll(full)
          for (int j = 0; j < kNBI; ++j) v[j] = csub(v[j], cmul(f, Ps[m][pc + j]));
          v[m] = cneg(cmul(f, Ps[m][pc + m]));
        }
      }
    }
  }
  __syncthreads();  // every read of the panel's old values is done
  if (live) {
    if (g == 0) {
#pragma clang loop unroll(full)
      for (int m = 0; m < kNBI; ++m)
        if (m < nb) A[(long long)(c0 + m) * ld + i] = v[m];
    }
    // columns g, g + CG, ... in batches of U: every load of a batch is issued
    // before its stores
    constexpr int U = 16;
    for (int jb = g; jb < kb; jb += CG * U) {
      V acc[U];
#pragma clang loop unroll(full)
      for (int u = 0; u < U; ++u) {
        const int jc = jb + u * CG, j = k0 + jc;
        const bool on = jc < kb && (j < c0 || j >= c0 + nb) && keep;
        acc[u] = load_sel(A + (long long)(jc < kb ? j : k0) * ld + i, on);
      }
#pragma clang loop unroll(full)
      for (int u = 0; u < U; ++u) {
        const int jc = jb + u * CG;
        if (jc < kb) {
#pragma clang loop unroll(full)
          for (int m = 0; m < kNBI; ++m)
            if (m < nb) acc[u] = cadd(acc[u], cmul(v[m], Ps[m][jc]));
        }
      }
#pragma clang loop unroll(full)
      for (int u = 0; u < U; ++u) {
        const int jc = jb + u * CG, j = k0 + jc;
        if (jc < kb && (j < c0 || j >= c0 + nb)) A[(long long)j * ld + i] = acc[u];
      }
    }
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
#pragma clang loop unroll(full)
  for (int u = 0; u < 8; ++u) {
    t[u] = mk<T>(0, 0);
    if (m0 + u < kb) {
      const int p = pivrow[k0 + m0 + u];
      t[u] = A[(long long)j * ld + (p < n ? p : 0)];
    }
  }
#pragma clang loop unroll(full)
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
  const int nchunks = (n + 255) / 256;  // <= kGjCandInts / kNBI at n <= 3072
  for (int k0 = 0; k0 < n; k0 += kNBO) {
    const int kb = (n - k0 < kNBO) ? n - k0 : kNBO;
    for (int c0 = k0; c0 < k0 + kb; c0 += kNBI) {
      const int nb = (k0 + kb - c0 < kNBI) ? k0 + kb - c0 : kNBI;
      // sampled timing of the panel step: the outer block read and written once,
      // and the panel's columns read once more for the candidates
      hipEvent_t ev0, ev1;
      const double bytes = sizeof(cx<T>) * (double)n * (2.0 * kb + nb) * batch;
      const bool timed = timer_slot(bytes, kTimerPanel64 + (sizeof(T) == sizeof(float) ? 1 : 0), &ev0, &ev1);
      if (timed && hipEventRecord(ev0, s) != hipSuccess) return hipGetLastError();
      hipLaunchKernelGGL((k_gj_cand<T>), dim3(nchunks, batch), dim3(256), 0, s, S, sS, n, ld, c0, nb, w,
                         active);
      hipLaunchKernelGGL((k_gj_pivots<T>), dim3(batch), dim3(64), 0, s, S, sS, n, ld, c0, nb, k0, kb,
                         nchunks, w, active);
      hipLaunchKernelGGL((k_gj_apply<T>), dim3((n + 31) / 32, batch), dim3(256), 0, s, S, sS, n, ld, c0,
                         nb, k0, kb, w, active);
      if (timed && hipEventRecord(ev1, s) != hipSuccess) return hipGetLastError();
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
      g.m3 = w.m3 ? 1 : 0;
      g.x3 = w.m3 == 2;
      hipError_t e = gemm_gj_update<T>(s, g, batch);
      if (e != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL((k_gj_unpermute<T>), dim3((n + 255) / 256, n, batch), dim3(256), 0, s, S, sS,
                     ld, out, sOut, ldo, w, n, active);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Block Gauss-Jordan with pivoting inside diagonal blocks only (gj_invert_local;
// 128-column blocks in complex64, 64-column blocks in complex128).  For each outer block K = [k0, k0 + kb):
//   diag   -- one 1024-thread workgroup per matrix inverts D = S[K, K] by
//             Gauss-Jordan with partial pivoting among the block's rows (the block
//             held in registers, 16 entries per thread; column m and the pivot row
//             pass through LDS) and writes Dinv (kb x kb, column-major);
//   row    -- R = Dinv S[K, :] (a GEMM; stored transposed, n x kb);
//   cols   -- A'[i] = e_{i-k0} for i in K, -S[i, K] otherwise; S[:, K] = 0; R[:, K] = Dinv;
//   update -- S = keep(S) + A' R with the GEMM's Gauss-Jordan epilogue, keep() zeroing
//             the rows of K (pivstep = identity):
//               rows in K:      R          (the block row; Dinv in the columns of K),
//               rows not in K:  S[i, :] - S[i, K] R  (-S[i, K] Dinv in the columns of K,
//                               which cols cleared first).
// The result is the inverse, in place and unpermuted.  Without row interchanges between
// blocks its accuracy rests on the diagonal blocks being well conditioned: it is
// for callers that check their results (the complex64 block LU whose solves are
// refined in FP64, which falls back to the partially pivoted FP64 inverse when
// the refinement does not contract).  4 launches per 128 columns instead of 50.
// With GjWork::group > 1 (complex64) the updates are delayed over groups of blocks
// (see gj_invert_local).

// Pivot keys of the diagonal-block inverse: the magnitude |re|+|im| with its low
// bits replaced by the row's rank (127 - i), so the larger magnitude wins and ties
// go to the lower row.  complex64: a 32-bit key (magnitudes within 2^-16 relative
// count as ties; a magnitude below 2^-126 counts as zero); complex128: the 64-bit key
// of the partially pivoted inverse (pivot_key).  The row bits of a real candidate
// never make a key zero unless the row is 127 with magnitude zero, which the
// zero-pivot test sees either way.
template <typename T> struct LocalKey;
template <> struct LocalKey<float> {
  typedef unsigned int K;
  __device__ static K make(float2 x, int i) {
    float m = fabsf(x.x) + fabsf(x.y);
    if (m != m) m = INFINITY;
    return (__float_as_uint(m) & ~0x7fu) | (unsigned)(127 - i);
  }
  __device__ static int row(K k) { return 127 - (int)(k & 0x7fu); }
  __device__ static bool zero(K k) { return (k >> 7) == 0; }
};
template <> struct LocalKey<double> {
  typedef unsigned long long K;
  __device__ static K make(double2 x, int i) { return pivot_key(x, i); }
  __device__ static int row(K k) { return 4095 - (int)(k & 0xfffull); }
  __device__ static bool zero(K k) { return (k & ~0xfffull) == 0; }
};

// One 32-bit lane value through DPP; rows outside `RM` keep `old`.
template <int CTRL, int RM>
__device__ inline unsigned dpp32(unsigned old, unsigned x) {
  return (unsigned)__builtin_amdgcn_update_dpp((int)old, (int)x, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ inline unsigned dpp_max(unsigned x) {
  const unsigned y = dpp32<CTRL, RM>(x, x);
  return x > y ? x : y;
}
template <int CTRL, int RM>
__device__ inline unsigned long long dpp_max(unsigned long long x) {
  const unsigned lo = dpp32<CTRL, RM>((unsigned)x, (unsigned)x);
  const unsigned hi = dpp32<CTRL, RM>((unsigned)(x >> 32), (unsigned)(x >> 32));
  const unsigned long long y = ((unsigned long long)hi << 32) | lo;
  return x > y ? x : y;
}
__device__ inline unsigned readlane_u(unsigned x, int l) { return (unsigned)__builtin_amdgcn_readlane((int)x, l); }
__device__ inline unsigned long long readlane_u(unsigned long long x, int l) {
  return ((unsigned long long)readlane_u((unsigned)(x >> 32), l) << 32) | readlane_u((unsigned)x, l);
}
// Max over the wave, returned uniform: within each row of 16 lanes (quad xor 1, quad
// xor 2, half-row mirror, row mirror), then lane 15 of rows 0 and 2 into rows 1 and
// 3, lane 31 into rows 2 and 3, so lane 63 holds the maximum; read from there.
template <typename K>
__device__ inline K wave_max_uniform(K x) {
  x = dpp_max<0xb1, 0xf>(x);
  x = dpp_max<0x4e, 0xf>(x);
  x = dpp_max<0x141, 0xf>(x);
  x = dpp_max<0x140, 0xf>(x);
  x = dpp_max<0x142, 0xa>(x);  // row_bcast:15
  x = dpp_max<0x143, 0xc>(x);  // row_bcast:31
  return readlane_u(x, 63);
}
template <typename V>
__device__ inline V readlane_c(V x, int l);
template <>
__device__ inline float2 readlane_c(float2 x, int l) {
  return mk<float>(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(x.x), l)),
                   __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x.y), l)));
}
template <>
__device__ inline double2 readlane_c(double2 x, int l) {
  auto rd = [&](double d) {
    const long long b = __double_as_longlong(d);
    const unsigned lo = readlane_u((unsigned)b, l), hi = readlane_u((unsigned)((unsigned long long)b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  };
  return mk<double>(rd(x.x), rd(x.y));
}

// The diagonal block's inverse (blocks of KB = 128 columns in complex64, 64 in
// complex128; KB x 8 threads).  Thread (r, g) = (tid % KB, tid / KB) holds row r of D
// in columns g + 8 q (q < KB / 8): one column of every 8-column panel.  Per panel P:
//   factor -- one wave (wave 0) runs the panel's Gauss-Jordan steps on all the
//             block's rows (KB / 64 rows per lane; partial pivoting among the unused
//             rows: a DPP max to lane 63, the pivot row read with readlane) and
//             publishes each row's final panel values S[r][.] and the pivots;
//   update -- every thread takes the steps in its other columns,
//               v[c] <- [r not a pivot of the panel] v[c] + sum_m S[r][m] Prow[m][c],
//             with Prow the pivot rows (staged in LDS), and S[r][.] in the panel's own
//             columns.
// Look-ahead: each thread first updates its column of panel P + 1 with panel P and
// publishes it, so wave 0 factors panel P + 1 while the other waves update the rest
// of the block with panel P.  S and the pivots are double-buffered; three barriers per
// panel.  Finally inv(D)[pstep[r]][prow[c]] = D[r][c] goes to Dinv (kb x kb,
// column-major), and info gets the first zero pivot (1-based) if it has none yet.
template <typename T, int KB>
__global__ __launch_bounds__(KB * 8) void k_gjl_diag(const cx<T>* S, long long sS, int ld, int k0, int kb,
                                                     cx<T>* Dinv, long long sD, int* info, long long sInt,
                                                     const int* active) {
  typedef cx<T> V;
  typedef typename LocalKey<T>::K Key;
  constexpr int NQ = KB / 8, HR = KB / 64;
  const int b = active ? active[blockIdx.x] : (int)blockIdx.x;
  const V* A = S + (long long)b * sS;
  V* Di = Dinv + (long long)b * sD;
  const int tid = threadIdx.x, lane = tid & 63;
  const int r = tid % KB, cg = tid / KB;
  __shared__ V pan[kNBI][KB];            // the next panel's columns
  __shared__ V stab[2][KB][kNBI + 1];    // each row's final panel values (padded)
  __shared__ V prw[kNBI][KB];            // the panel's pivot rows, every column
  __shared__ int pv[2][kNBI];            // the panel's pivot rows
  __shared__ int pstep[KB], prow[KB];
  __shared__ V vstash[NQ][64];           // wave 0's own columns while it factors
  V v[NQ];
#pragma clang loop unroll(full)
  for (int q = 0; q < NQ; ++q) {
    const int c = cg + 8 * q;
    v[q] = load_sel(A + (long long)(k0 + (c < kb ? c : 0)) * ld + k0 + (r < kb ? r : 0), r < kb && c < kb);
  }
  if (tid < KB) pstep[tid] = -1;
  const int npan = (kb + kNBI - 1) / kNBI;
  auto panel_nb = [&](int P) { return (kb - P * kNBI < kNBI) ? kb - P * kNBI : kNBI; };
  int* inf = info + (long long)b * sInt;
  bool used[HR];  // wave 0: rows lane + 64 h
#pragma clang loop unroll(full)
  for (int h = 0; h < HR; ++h) used[h] = lane + 64 * h >= kb;
  // wave 0: the Gauss-Jordan steps of panel P on the panel's columns (from pan)
  auto factor = [&](int P, int buf) {
    const int nb = panel_nb(P);
    V w[HR][kNBI];
#pragma clang loop unroll(full)
    for (int h = 0; h < HR; ++h)
#pragma clang loop unroll(full)
      for (int j = 0; j < kNBI; ++j) w[h][j] = j < nb ? pan[j][lane + 64 * h] : mk<T>(0, 0);
#pragma clang loop unroll(full)
    for (int j = 0; j < kNBI; ++j) {
      if (j < nb) {
        Key k = 0;
#pragma clang loop unroll(full)
        for (int h = 0; h < HR; ++h) {
          const Key kh = used[h] ? (Key)0 : LocalKey<T>::make(w[h][j], lane + 64 * h);
          k = kh > k ? kh : k;
        }
        const Key g = wave_max_uniform(k);
        if (LocalKey<T>::zero(g) && lane == 0 && *inf == 0) *inf = k0 + P * kNBI + j + 1;  // zero pivot
        int p = __builtin_amdgcn_readfirstlane(LocalKey<T>::row(g));
        p = (p >= 0 && p < kb) ? p : 0;
        const int lp = p & 63, hi = p >> 6;
        V t[kNBI];
#pragma clang loop unroll(full)
        for (int c = 0; c < kNBI; ++c) {
          V x = w[0][c];
#pragma clang loop unroll(full)
          for (int h = 1; h < HR; ++h)
            if (hi == h) x = w[h][c];
          t[c] = readlane_c(x, lp);
        }
        const V inva = cinv(t[j]);
#pragma clang loop unroll(full)
        for (int c = 0; c < kNBI; ++c) t[c] = (c == j) ? inva : cmul(t[c], inva);
        // the pivot row becomes t, the others take the step (selects, no branches:
        // every lane runs the same instructions)
#pragma clang loop unroll(full)
        for (int h = 0; h < HR; ++h) {
          const bool pr = lane + 64 * h == p;
          const V f = w[h][j];
#pragma clang loop unroll(full)
          for (int c = 0; c < kNBI; ++c) {
            const V u = (c == j) ? cneg(cmul(f, t[j])) : csub(w[h][c], cmul(f, t[c]));
            w[h][c].x = pr ? t[c].x : u.x;
            w[h][c].y = pr ? t[c].y : u.y;
          }
          used[h] = used[h] || pr;
        }
        // uniform values: every lane writes the same word
        pv[buf][j] = p;
        pstep[p] = P * kNBI + j;
        prow[P * kNBI + j] = p;
      }
    }
#pragma clang loop unroll(full)
    for (int h = 0; h < HR; ++h)
#pragma clang loop unroll(full)
      for (int j = 0; j < kNBI; ++j) stab[buf][lane + 64 * h][j] = w[h][j];
  };
  // this row's place among panel P's pivots (-1: not a pivot)
  auto place = [&](int P, int buf) {
    const int nb = panel_nb(P);
    int ip = -1;
#pragma clang loop unroll(full)
    for (int m = 0; m < kNBI; ++m)
      if (m < nb && pv[buf][m] == r) ip = m;
    return ip;
  };
  // pivot rows of panel P: their values in every column to prw
  auto stage_rows = [&](int P, int buf) {
    const int ip = place(P, buf);
    if (ip >= 0) {
#pragma clang loop unroll(full)
      for (int q = 0; q < NQ; ++q) prw[ip][cg + 8 * q] = v[q];
    }
  };
  // P = -1 only factors panel 0; P >= 0 applies panel P and factors panel P + 1
  for (int P = -1; P < npan; ++P) {
    const int buf = P & 1;
    const bool ahead = P + 1 < npan, have = P >= 0;
    const int nb = have ? panel_nb(P) : 0;
    const int ip = have ? place(P, buf) : -1;
    V sr[kNBI];  // loaded before each use: wave 0 does not hold them through factor()
    auto load_sr = [&]() {
#pragma clang loop unroll(full)
      for (int m = 0; m < kNBI; ++m) sr[m] = stab[buf][r][m];
    };
    // one column q (not of panel P) takes panel P's steps
    auto step_col = [&](int q) {
      V acc = ip >= 0 ? mk<T>(0, 0) : v[q];
#pragma clang loop unroll(full)
      for (int m = 0; m < kNBI; ++m)
        if (m < nb) acc = cadd(acc, cmul(sr[m], prw[m][cg + 8 * q]));
      v[q] = acc;
    };
    if (ahead) {  // look-ahead: panel P + 1's columns first, published for wave 0
      if (have) {
        load_sr();
#pragma clang loop unroll(full)
        for (int q = 0; q < NQ; ++q)
          if (q == P + 1) step_col(q);
      }
      if (cg < panel_nb(P + 1)) {
        V x = v[0];
#pragma clang loop unroll(full)
        for (int q = 1; q < NQ; ++q)
          if (q == P + 1) x = v[q];
        pan[cg][r] = x;
      }
    }
    __syncthreads();
    if (ahead && tid < 64) {  // wave 0 parks its columns in LDS while it factors
#pragma clang loop unroll(full)
      for (int q = 0; q < NQ; ++q) vstash[q][lane] = v[q];
      factor(P + 1, buf ^ 1);
#pragma clang loop unroll(full)
      for (int q = 0; q < NQ; ++q) v[q] = vstash[q][lane];
    }
    // (rows past kb hold zeros and stay zero: no row test, so every branch here is uniform)
    if (have) {
      load_sr();
#pragma clang loop unroll(full)
      for (int q = 0; q < NQ; ++q) {
        if (q == P) {
          if (cg < nb) v[q] = stab[buf][r][cg];
        } else if (q != P + 1 || !ahead) {
          step_col(q);
        }
      }
    }
    __syncthreads();
    if (ahead) stage_rows(P + 1, buf ^ 1);
    __syncthreads();
  }
  if (r < kb) {
    const int a = pstep[r];
#pragma clang loop unroll(full)
    for (int q = 0; q < NQ; ++q) {
      const int c = cg + 8 * q;
      if (c < kb) {
        const int bc = prow[c];
        if (a >= 0 && a < kb && bc >= 0 && bc < kb) Di[(long long)bc * kb + a] = v[q];
      }
    }
  }
}

// A'[i][m] = (i in K) ? delta(i - k0, m) : -S[i][k0 + m];  S[i][k0 + m] = 0;  and for the
// rows of K, R[m][i] = Dinv[m][i - k0] (R stored transposed: R^T (n x kb), ld n).
template <typename T>
__global__ void k_gjl_cols(cx<T>* S, long long sS, int ld, int n, int k0, int kb, cx<T>* Ap, long long sA,
                           cx<T>* Rt, long long sR, const cx<T>* Dinv, long long sD, const int* active) {
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const int m = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cx<T>* Sc = S + (long long)b * sS + (long long)(k0 + m) * ld;
  const bool inK = i >= k0 && i < k0 + kb;
  Ap[(long long)b * sA + (long long)m * n + i] = inK ? mk<T>(i - k0 == m ? 1 : 0, 0) : cneg(Sc[i]);
  Sc[i] = mk<T>(0, 0);
  if (inK) Rt[(long long)b * sR + (long long)m * n + i] = Dinv[(long long)b * sD + (long long)(i - k0) * kb + m];
}

// Bt[j + m n] = S[g0 + m, j] for m < w, j < n: a band of rows of S, transposed (32 x 32
// tiles through LDS, so both the read and the write are coalesced)
template <typename T>
__global__ __launch_bounds__(256) void k_gjl_rows_t(const cx<T>* S, long long sS, int ld, int n, int g0, int w,
                                                    cx<T>* Bt, long long sB, const int* active) {
  __shared__ cx<T> tile[32][33];
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const int j0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const cx<T>* Sb = S + (long long)b * sS;
  cx<T>* Bb = Bt + (long long)b * sB;
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + r, m = m0 + tx;
    if (j < n && m < w) tile[r][tx] = Sb[(long long)(g0 + m) + (long long)j * ld];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int m = m0 + r, j = j0 + tx;
    if (j < n && m < w) Bb[(long long)j + (long long)m * n] = tile[tx][r];
  }
}

// pivstep = identity: the epilogue's row mask for "rows of K" (info is left to the
// caller, which clears it before a sequence of inversions and reads the first failure)
__global__ void k_gjl_init(int* pivstep, long long sInt, int n, const int* active) {
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    pivstep[(long long)b * sInt + i] = i;
}

template <typename T>
hipError_t gj_invert_local(hipStream_t s, cx<T>* S, long long sS, int ld, int n, int batch, const GjWork& w,
                           const int* active) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (!w.tmp) return hipErrorInvalidValue;
  // workspaces: A' (n x kb) at w.tmp, then (grouped) Bt (n x group kb); R^T (n x kb) then
  // Dinv (kb x kb) in the pivot-row panel (kGjNBO x n complex128 holds 2 kGjNBO x n of
  // complex64, and both fit)
  constexpr int KB = LocalBlock<T>::kb;
  cx<T>* Ap = (cx<T>*)w.tmp;
  cx<T>* Bt = Ap + (long long)n * KB;
  cx<T>* Rt = (cx<T>*)w.prow;
  cx<T>* Di = Rt + (long long)n * (n < KB ? n : KB);
  // Delayed updates (complex64, group > 1): the blocks of a group of columns [g0, g0 + wg)
  // update only the group's own columns; the accumulated transform M, which differs from
  // the identity only in those columns and is what they then hold, is applied to the
  // other columns at once: S[:, j] = keep(S[:, j]) + S[:, G] S0[G, j], with S0[G, j] the
  // group's rows as they were before the group (Bt), one rank-wg product.
  const int G = (sizeof(T) == sizeof(float) && w.group > 1) ? w.group : 1;
  hipLaunchKernelGGL(k_gjl_init, dim3((n + 255) / 256, batch), dim3(256), 0, s, w.pivstep, w.sInt, n, active);
  for (int g0 = 0; g0 < n; g0 += G * KB) {
    const int wg = (n - g0 < G * KB) ? n - g0 : G * KB;
    const bool grouped = G > 1 && wg > KB;
    if (grouped && wg < n)
      hipLaunchKernelGGL((k_gjl_rows_t<T>), dim3((n + 31) / 32, (wg + 31) / 32, batch), dim3(256), 0, s, S, sS, ld,
                         n, g0, wg, Bt, w.sTmp, active);
    for (int k0 = g0; k0 < g0 + wg; k0 += KB) {
      const int kb = (g0 + wg - k0 < KB) ? g0 + wg - k0 : KB;
      hipLaunchKernelGGL((k_gjl_diag<T, KB>), dim3(batch), dim3(KB * 8), 0, s, S, sS, ld, k0, kb, Di, w.sProw,
                         w.info, w.sInt, active);
      // R^T = S[K, c]^T Dinv^T over the columns c the update below reaches
      const int c0 = grouped ? g0 : 0, nc = grouped ? wg : n;
      GemmBatch g = {};
      g.A = S + k0 + (long long)c0 * ld; g.sA = sS; g.lda = ld;
      g.B = Di; g.sB = w.sProw; g.ldb = kb;
      g.C = Rt + c0; g.sC = w.sProw; g.ldc = n;
      g.m = nc; g.n = kb; g.k = kb;
      g.active = active;
      g.m3 = w.m3 ? 1 : 0;
      g.x3 = w.m3 == 2;
      hipError_t e = gemm_batched<T>(s, OP_T, OP_T, g, batch, mk<T>(1, 0), mk<T>(0, 0));
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((k_gjl_cols<T>), dim3((n + 255) / 256, kb, batch), dim3(256), 0, s, S, sS, ld, n, k0, kb,
                         Ap, w.sTmp, Rt, w.sProw, Di, w.sProw, active);
      GemmBatch u = {};  // S[:, c] = keep(S[:, c]) + A' R[:, c]
      u.A = Ap; u.sA = w.sTmp; u.lda = n;
      u.B = Rt + c0; u.sB = w.sProw; u.ldb = n;
      u.C = S + (long long)c0 * ld; u.sC = sS; u.ldc = ld;
      u.m = n; u.n = nc; u.k = kb;
      u.active = active; u.pivstep = w.pivstep; u.sPiv = w.sInt;
      u.pv_lo = k0; u.pv_hi = k0 + kb;
      u.skip_lo = u.skip_hi = 0;
      u.m3 = w.m3 ? 1 : 0;
      u.x3 = w.m3 == 2;
      e = gemm_gj_update<T>(s, u, batch);
      if (e != hipSuccess) return e;
    }
    if (grouped && wg < n) {
      GemmBatch u = {};  // S[:, j] = keep(S[:, j]) + S[:, G] Bt^T[G, j] outside the group
      u.A = S + (long long)g0 * ld; u.sA = sS; u.lda = ld;
      u.B = Bt; u.sB = w.sTmp; u.ldb = n;
      u.C = S; u.sC = sS; u.ldc = ld;
      u.m = n; u.n = n; u.k = wg;
      u.active = active; u.pivstep = w.pivstep; u.sPiv = w.sInt;
      u.pv_lo = g0; u.pv_hi = g0 + wg;
      u.skip_lo = g0; u.skip_hi = g0 + wg;
      u.m3 = w.m3 ? 1 : 0;
      u.x3 = w.m3 == 2;
      hipError_t e = gemm_gj_update<T>(s, u, batch);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

template hipError_t gj_invert_local<float>(hipStream_t, cx<float>*, long long, int, int, int, const GjWork&,
                                           const int*);
template hipError_t gj_invert_local<double>(hipStream_t, cx<double>*, long long, int, int, int, const GjWork&,
                                            const int*);

template hipError_t gj_invert<double>(hipStream_t, cx<double>*, long long, int, cx<double>*,
                                      long long, int, int, int, const GjWork&, const int*);
template hipError_t gj_invert<float>(hipStream_t, cx<float>*, long long, int, cx<float>*,
                                     long long, int, int, int, const GjWork&, const int*);

}  // namespace negf
