// Batched complex GEMM on CDNA4 matrix cores (gfx950).
//
// Two forms of the complex product per 16x16x4 step, chosen per launch
// (GemmBatch::m3):
//  * 4-product form: Re C += Re A Re B - Im A Im B, Im C += Re A Im B + Im A Re B.
//    Each component is accurate relative to its own terms, so a small imaginary
//    part (the O(eta) anti-Hermitian part of a lead self-energy inside a gap)
//    keeps its relative accuracy.
//  * 3M form: P1 += Re A Re B, P2 += Im A Im B, P3 += (Re A + Im A)(Re B + Im B),
//    and at the end Re C = P1 - P2, Im C = P3 - P1 - P2: 3/4 of the matrix-core
//    work, sums formed in registers.  Its error is normwise,
//    O(eps (|Re A|+|Im A|)(|Re B|+|Im B|)) per term for both components.
// Uses v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32).
// A 256-thread workgroup owns a 64x64 tile of C; its four wave64s each own a
// 32x32 quadrant (2x2 MFMA blocks; three accumulators in the 3M form, two otherwise).
// Tiles of op(A) (64x16) and op(B) (16x64) are staged through LDS as split
// real/imaginary planes, k-major, so that each MFMA operand read is 16
// consecutive LDS words; the next K slab is fetched into registers while the
// current one is multiplied (LDS: 34 KiB for complex64, 51 KiB for FP64).
//
// This replaces the reference's MKL zgemm call sites
// (Recursive_Greens.F90:18-95, Contact_Self_Energy.F90:147-209): square
// m = n = k products with op in {N, C} on either operand.  The Gauss-Jordan
// epilogue turns the same kernel into the trailing update of the blocked
// inverse (gj_inverse.hip).
#include "negf_device.h"

#include <cstdlib>
#include <mutex>

namespace negf {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Accumulator layout (measured on gfx950): element r of lane l holds D[row][l % 16]
// with row = l/16 + 4r for the f64 16x16x4 op and row = 4*(l/16) + r for f32.
template <> struct Mfma<double> {
  typedef v4d acc;
  static __device__ inline int row(int lk, int r) { return lk + 4 * r; }
  static __device__ inline acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  typedef v4f acc;
  static __device__ inline int row(int lk, int r) { return 4 * lk + r; }
  static __device__ inline acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int kBK = 16, kThreads = 256;
constexpr int kPad = 4;

template <typename T, int OP>
__device__ inline cx<T> load_op(const cx<T>* p, bool ok) {
  cx<T> v = mk<T>((T)0, (T)0);
  if (ok) {
    v = *p;
    if (OP == OP_C) v.y = -v.y;
  }
  return v;
}

// Maps a linear workgroup id to a (column-major) tile index.  Workgroup ids
// congruent mod 8 are dispatched to the same XCD, so each XCD gets a contiguous run
// o of a grouped order: runs of G = 8 tile rows, column-major inside a run.  The
// ~64 workgroups an XCD holds at once then cover about 8 x 8 tiles, which share 8
// A row panels and 8 B column panels in that XCD's L2 (plain column-major runs
// cover all tile rows and two columns: ~38 panels for the same 64 tiles).
// Returns ntiles for ids past the last tile.
__device__ inline int xcd_tile(int bid, int tiles_m, int ntiles) {
  constexpr int G = 8;
  const int per = (ntiles + 7) >> 3;
  const int o = (bid & 7) * per + (bid >> 3);
  if (o >= ntiles) return ntiles;
  const int tiles_n = ntiles / tiles_m;
  const int g = o / (G * tiles_n);
  const int r = o - g * G * tiles_n;
  const int rows = min(G, tiles_m - g * G);
  return (r / rows) * tiles_m + g * G + r % rows;
}

// Workgroup tile (32 SM) x (32 SN): 2 x 2 waves, each owning SM x SN MFMA blocks of
// 16 x 16.  K advances in slabs of 16 through a two-stage LDS ring: the next slab
// is fetched into registers and written to the idle stage while the current
// stage feeds the MFMAs, so each slab costs one barrier (two in FP64, whose B
// tile has one stage).
// PRE = 1 (Gauss-Jordan epilogue with short K): the old C values and pivot flags
// are loaded before the K loop, so their latency overlaps the A/B traffic.
// Two workgroups per CU (launch bound): the FP64 64x64 variant then fits 199 VGPRs
// with no AGPR accumulators and no spills, and the second workgroup hides the
// LDS/global latency that one wave per SIMD leaves exposed (measured 53 -> 70 TF/s
// algorithmic at n = 1024, 55 -> 75 TF/s at n = 2304).
template <typename T, int OPA, int OPB, int EPI, int SM, int SN, int PRE, int M3>
__global__ __launch_bounds__(kThreads, 2) void k_gemm(GemmBatch g, int tiles_m, int ntiles,
                                                   cx<T> alpha, cx<T> beta) {
  typedef typename Mfma<T>::acc acc_t;
  constexpr int BM = 32 * SM, BN = 32 * SN;
  constexpr int LA = BM * kBK / kThreads, LB = BN * kBK / kThreads;
  const int b = g.active ? g.active[blockIdx.z] : (int)blockIdx.z;
  const int tile = xcd_tile(blockIdx.x, tiles_m, ntiles);
  if (tile >= ntiles) return;
  const int m0 = (tile % tiles_m) * BM;
  const int n0 = (tile / tiles_m) * BN;
  if (m0 >= g.m || n0 >= g.n) return;
  if (EPI == EPI_GJ && n0 >= g.skip_lo && n0 + BN <= g.skip_hi) return;
  if (EPI != EPI_GJ && g.ncol && n0 >= g.ncol[(long long)b * g.sNcol]) {
    // the tile lies past this matrix's columns, where B is zero
    const bool b0 = beta.x == (T)0 && beta.y == (T)0, b1 = beta.x == (T)1 && beta.y == (T)0;
    if (b0) {
      cx<T>* C0 = (cx<T>*)g.C + (long long)b * g.sC;
      for (int e = threadIdx.x; e < BM * BN; e += kThreads) {
        const int row = m0 + e % BM, col = n0 + e / BM;
        if (row < g.m && col < g.n) C0[(long long)col * g.ldc + row] = mk<T>(0, 0);
      }
    }
    if (b0 || b1) return;
  }

  const cx<T>* A = (const cx<T>*)g.A + (long long)b * g.sA;
  const cx<T>* B = (const cx<T>*)g.B + (long long)b * g.sB;
  cx<T>* C = (cx<T>*)g.C + (long long)b * g.sC;
  const int K = g.k, M = g.m, N = g.n;

  // B is single-buffered in the FP64 variant: two stages of both tiles would take
  // 68 KiB of LDS; one B stage keeps the kernel at 51 KiB (under 64 KiB) for the
  // cost of a second barrier per slab.
  constexpr int SB = sizeof(T) == sizeof(double) ? 1 : 2;
  __shared__ T sA[2][2][kBK][BM + kPad];   // [stage][re/im][k][m]
  __shared__ T sB[SB][2][kBK][BN + kPad];  // [stage][re/im][k][n]

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 16 * SM, wn = (wave >> 1) * 16 * SN;
  const int lr = lane & 15, lk = lane >> 4;

  acc_t p1[SM][SN], p2[SM][SN], p3[SM][SN];
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j) {
      p1[i][j] = acc_t{0, 0, 0, 0};
      p2[i][j] = acc_t{0, 0, 0, 0};
      p3[i][j] = acc_t{0, 0, 0, 0};
    }

  const int* piv = (EPI == EPI_GJ) ? g.pivstep + (long long)b * g.sPiv : nullptr;
  cx<T> cpre[PRE ? SM : 1][PRE ? SN : 1][PRE ? 4 : 1];
  bool kpre[PRE ? SM : 1][PRE ? 4 : 1];
  if (PRE) {
#pragma unroll
    for (int sm = 0; sm < SM; ++sm)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * sm + Mfma<T>::row(lk, r);
        const int ps = row < M ? piv[row] : 0;
        kpre[sm][r] = !(ps >= g.pv_lo && ps < g.pv_hi);
#pragma unroll
        for (int sn = 0; sn < SN; ++sn) {
          const int col = n0 + wn + 16 * sn + lr;
          const bool ok = row < M && col < N && !(col >= g.skip_lo && col < g.skip_hi);
          cpre[sm][sn][r] = ok ? C[(long long)col * g.ldc + row] : mk<T>(0, 0);
        }
      }
  }

  cx<T> ra[LA], rb[LB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int c = tid + kThreads * i;
      if (OPA == OP_N) {
        const int mm = c % BM, kk = c / BM, r = m0 + mm, q = k0 + kk;
        ra[i] = load_op<T, OPA>(A + (long long)q * g.lda + r, r < M && q < K);
      } else {
        const int kk = c % kBK, mm = c / kBK, r = m0 + mm, q = k0 + kk;
        ra[i] = load_op<T, OPA>(A + (long long)r * g.lda + q, r < M && q < K);
      }
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int c = tid + kThreads * i;
      if (OPB == OP_N) {
        const int kk = c % kBK, nn = c / kBK, q = k0 + kk, col = n0 + nn;
        rb[i] = load_op<T, OPB>(B + (long long)col * g.ldb + q, q < K && col < N);
      } else {
        const int nn = c % BN, kk = c / BN, q = k0 + kk, col = n0 + nn;
        rb[i] = load_op<T, OPB>(B + (long long)q * g.ldb + col, q < K && col < N);
      }
    }
  };
  auto stash = [&](int st) {
    const int sb = SB == 1 ? 0 : st;
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int c = tid + kThreads * i;
      int mm, kk;
      if (OPA == OP_N) { mm = c % BM; kk = c / BM; } else { kk = c % kBK; mm = c / kBK; }
      sA[st][0][kk][mm] = ra[i].x;
      sA[st][1][kk][mm] = ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int c = tid + kThreads * i;
      int nn, kk;
      if (OPB == OP_N) { kk = c % kBK; nn = c / kBK; } else { nn = c % BN; kk = c / BN; }
      sB[sb][0][kk][nn] = rb[i].x;
      sB[sb][1][kk][nn] = rb[i].y;
    }
  };

  const int nk = (K + kBK - 1) / kBK;
  if (nk > 0) {
    fetch(0);
    stash(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int st = kt & 1;
    const int sb = SB == 1 ? 0 : st;
    if (kt + 1 < nk) fetch((kt + 1) * kBK);
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      const int kk = ks * 4 + lk;
      T ar[SM], ai[SM], as[SM], br[SN], bi[SN], bs[SN];
#pragma unroll
      for (int s = 0; s < SM; ++s) {
        ar[s] = sA[st][0][kk][wm + 16 * s + lr];
        ai[s] = sA[st][1][kk][wm + 16 * s + lr];
        as[s] = M3 ? ar[s] + ai[s] : -ai[s];
      }
#pragma unroll
      for (int s = 0; s < SN; ++s) {
        br[s] = sB[sb][0][kk][wn + 16 * s + lr];
        bi[s] = sB[sb][1][kk][wn + 16 * s + lr];
        bs[s] = br[s] + bi[s];
      }
      if (M3) {
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
          for (int sn = 0; sn < SN; ++sn) {
            p1[sm][sn] = Mfma<T>::mma(ar[sm], br[sn], p1[sm][sn]);
            p2[sm][sn] = Mfma<T>::mma(ai[sm], bi[sn], p2[sm][sn]);
            p3[sm][sn] = Mfma<T>::mma(as[sm], bs[sn], p3[sm][sn]);
          }
      } else {
        // 4-product form: p1 = Re C, p3 = Im C (p2 unused)
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
          for (int sn = 0; sn < SN; ++sn) {
            p1[sm][sn] = Mfma<T>::mma(ar[sm], br[sn], p1[sm][sn]);
            p3[sm][sn] = Mfma<T>::mma(ar[sm], bi[sn], p3[sm][sn]);
          }
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
          for (int sn = 0; sn < SN; ++sn) {
            p1[sm][sn] = Mfma<T>::mma(as[sm], bi[sn], p1[sm][sn]);
            p3[sm][sn] = Mfma<T>::mma(ai[sm], br[sn], p3[sm][sn]);
          }
      }
    }
    if (SB == 1 && kt + 1 < nk) __syncthreads();  // every wave is done with the B stage
    if (kt + 1 < nk) stash(st ^ 1);
    __syncthreads();
  }

  // Epilogue: accumulator element r of lane (lk, lr) is C[Mfma<T>::row(lk, r)][lr]
  // of its 16x16 block.
#pragma unroll
  for (int sm = 0; sm < SM; ++sm)
#pragma unroll
    for (int sn = 0; sn < SN; ++sn) {
      const int col = n0 + wn + 16 * sn + lr;
      if (col >= N) continue;
      if (EPI == EPI_GJ && col >= g.skip_lo && col < g.skip_hi) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * sm + Mfma<T>::row(lk, r);
        if (row >= M) continue;
        cx<T>* pc = C + (long long)col * g.ldc + row;
        const T q1 = p1[sm][sn][r], q2 = p2[sm][sn][r];
        cx<T> v = M3 ? mk<T>(q1 - q2, p3[sm][sn][r] - q1 - q2) : mk<T>(q1, p3[sm][sn][r]);
        if (EPI == EPI_GJ && PRE) {
          if (kpre[sm][r]) v = cadd(v, cpre[sm][sn][r]);
        } else if (EPI == EPI_GJ) {
          const int ps = piv[row];
          if (!(ps >= g.pv_lo && ps < g.pv_hi)) v = cadd(v, *pc);
        } else {
          v = cmul(alpha, v);
          if (beta.x != (T)0 || beta.y != (T)0) v = cadd(v, cmul(beta, *pc));
        }
        *pc = v;
      }
    }
}

GemmTimer& gemm_timer() {
  static GemmTimer t;
  return t;
}

static std::mutex g_timer_mu;

// Reserves an event pair for one launch; the handles are copied out under the lock
// (the event array may grow while another host thread launches).
// Tuning table: per element type, rules (n_max, tile) in increasing n_max; the
// first rule with n <= n_max picks the tile, none gives the 64x64 tile.  Set once
// by negf_set_gemm_tiles (before solves start; not synchronised with launches).
constexpr int kTile64 = 0, kTile32 = 1, kMaxTileRules = 16;
static int g_rules[2][kMaxTileRules][2];
static int g_nrules[2] = {0, 0};
static thread_local int t_force_tile = -1;  // negf_bench_gemm_tile: one tile for this thread

int gemm_tile(int fp32, int n) {
  if (t_force_tile >= 0) return t_force_tile;
  for (int i = 0; i < g_nrules[fp32]; ++i)
    if (n <= g_rules[fp32][i][0]) return g_rules[fp32][i][1];
  return kTile64;
}

int set_gemm_tiles(int fp32, int count, const int* n_max, const int* tile) {
  if (fp32 < 0 || fp32 > 1 || count < 0 || count > kMaxTileRules || (count && (!n_max || !tile)))
    return -1;
  for (int i = 0; i < count; ++i) {
    if (tile[i] != kTile64 && tile[i] != kTile32) return -1;
    if (i && n_max[i] <= n_max[i - 1]) return -1;
  }
  for (int i = 0; i < count; ++i) {
    g_rules[fp32][i][0] = n_max[i];
    g_rules[fp32][i][1] = tile[i];
  }
  g_nrules[fp32] = count;
  return 0;
}

void force_gemm_tile(int tile) { t_force_tile = tile; }

static thread_local bool t_capturing = false;
void timer_capture(bool on) { t_capturing = on; }

bool timer_slot(double flops, int kind, hipEvent_t* e0, hipEvent_t* e1) {
  GemmTimer& t = gemm_timer();
  if (!t.on || t_capturing) return false;
  std::lock_guard<std::mutex> lock(g_timer_mu);
  if (t.seen[kind]++ % (unsigned long long)t.every != 0) return false;
  if (t.used == t.cap) {
    size_t nc = t.cap ? 2 * t.cap : 4096;
    hipEvent_t* ev = (hipEvent_t*)realloc(t.ev, sizeof(hipEvent_t) * 2 * nc);
    if (!ev) return false;
    t.ev = ev;
    double* fl = (double*)realloc(t.flops, sizeof(double) * nc);
    if (!fl) return false;
    t.flops = fl;
    int* kd = (int*)realloc(t.kind, sizeof(int) * nc);
    if (!kd) return false;
    t.kind = kd;
    for (size_t i = 2 * t.cap; i < 2 * nc; ++i)
      if (hipEventCreate(&ev[i]) != hipSuccess) {
        for (size_t k = 2 * t.cap; k < i; ++k) (void)hipEventDestroy(ev[k]);
        return false;  // sampling stops; the launch itself is unaffected
      }
    t.cap = nc;
  }
  t.flops[t.used] = flops;
  t.kind[t.used] = kind;
  *e0 = t.ev[2 * t.used];
  *e1 = t.ev[2 * t.used + 1];
  ++t.used;
  return true;
}

template <typename T, int OPA, int OPB, int EPI, int SM, int SN, int M3>
static hipError_t launch_cfg(hipStream_t s, const GemmBatch& g, int batch, cx<T> alpha, cx<T> beta) {
  constexpr int BM = 32 * SM, BN = 32 * SN;
  int tm = (g.m + BM - 1) / BM, tn = (g.n + BN - 1) / BN;
  int nt = tm * tn;
  dim3 grid(((nt + 7) / 8) * 8, 1, batch);
  // C prefetch for short K (Gauss-Jordan updates of a last block of <= 32 columns);
  // the FP64 3M variant has no registers left for it (it would spill), so it skips it
  constexpr bool kPre = EPI == EPI_GJ && !(M3 && sizeof(T) == sizeof(double));
  if constexpr (kPre) {
    if (g.k <= 32) {
      hipLaunchKernelGGL((k_gemm<T, OPA, OPB, EPI, SM, SN, 1, M3>), grid, dim3(kThreads), 0, s, g, tm,
                         nt, alpha, beta);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_gemm<T, OPA, OPB, EPI, SM, SN, 0, M3>), grid, dim3(kThreads), 0, s, g, tm, nt,
                     alpha, beta);
  return hipGetLastError();
}

template <typename T, int OPA, int OPB, int EPI>
static hipError_t launch(hipStream_t s, const GemmBatch& g, int batch, cx<T> alpha, cx<T> beta) {
  if (batch <= 0 || g.m <= 0 || g.n <= 0) return hipSuccess;
  const int ncols = (EPI == EPI_GJ) ? g.n - (g.skip_hi - g.skip_lo) : g.n;
  hipEvent_t ev0, ev1;
  // Workgroup tile from the tuning table (gemm_tile): 64x64 by default -- 128x64 and
  // 64x128 measured within 3% of it at blocks of 1024 and more, and their FP64
  // variants do not fit two workgroups per CU without spills -- or 32x32, four
  // times the workgroups for small blocks.  Complex64 products that ask for the bf16
  // split form take it with 64x64 tiles (zgemm_x3.hip).
  const int tile = gemm_tile(sizeof(T) == sizeof(float), g.m > g.n ? g.m : g.n);
  const bool x3 = sizeof(T) == sizeof(float) && g.x3 && tile != kTile32;
  const int kind = x3 ? kTimerX3 + (EPI == EPI_GJ ? 1 : 0)
                      : (EPI == EPI_GJ ? 1 : 0) + (sizeof(T) == sizeof(float) ? 2 : 0) + (g.m3 ? 4 : 0);
  // a kind's samples are one kernel instance: the plain products (both operands
  // untransposed) and the Gauss-Jordan updates; the transposed and adjoint products
  // (small block-row products and traces) are not sampled
  const bool sampled = EPI == EPI_GJ || (OPA == OP_N && OPB == OP_N);
  const bool timed = sampled && timer_slot(8.0 * g.m * (double)ncols * g.k * batch * g.work, kind, &ev0, &ev1);
  if (timed) {
    const hipError_t er = hipEventRecord(ev0, s);
    if (er != hipSuccess) return er;
  }
  hipError_t e;
  if constexpr (sizeof(T) == sizeof(float)) {
    if (x3) {
      e = gemm_x3(s, OPA, OPB, EPI, g, batch, alpha, beta);
      if (timed && e == hipSuccess) return hipEventRecord(ev1, s);
      return e;
    }
  }
  if (tile == kTile32)
    e = g.m3 ? launch_cfg<T, OPA, OPB, EPI, 1, 1, 1>(s, g, batch, alpha, beta)
             : launch_cfg<T, OPA, OPB, EPI, 1, 1, 0>(s, g, batch, alpha, beta);
  else
    e = g.m3 ? launch_cfg<T, OPA, OPB, EPI, 2, 2, 1>(s, g, batch, alpha, beta)
             : launch_cfg<T, OPA, OPB, EPI, 2, 2, 0>(s, g, batch, alpha, beta);
  if (timed && e == hipSuccess) return hipEventRecord(ev1, s);
  return e;
}

template <typename T>
hipError_t gemm_batched(hipStream_t s, int opa, int opb, const GemmBatch& g, int batch,
                        cx<T> alpha, cx<T> beta) {
#define NEGF_CASE(a, b)                                              \
  if (opa == a && opb == b) return launch<T, a, b, EPI_STD>(s, g, batch, alpha, beta);
  NEGF_CASE(OP_N, OP_N)
  NEGF_CASE(OP_N, OP_C)
  NEGF_CASE(OP_C, OP_N)
  NEGF_CASE(OP_N, OP_T)
  NEGF_CASE(OP_T, OP_N)
  NEGF_CASE(OP_C, OP_C)
  NEGF_CASE(OP_T, OP_T)
#undef NEGF_CASE
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t gemm_gj_update(hipStream_t s, const GemmBatch& g, int batch) {
  return launch<T, OP_N, OP_T, EPI_GJ>(s, g, batch, mk<T>(1, 0), mk<T>(0, 0));
}

template hipError_t gemm_batched<double>(hipStream_t, int, int, const GemmBatch&, int, cx<double>, cx<double>);
template hipError_t gemm_batched<float>(hipStream_t, int, int, const GemmBatch&, int, cx<float>, cx<float>);
template hipError_t gemm_gj_update<double>(hipStream_t, const GemmBatch&, int);
template hipError_t gemm_gj_update<float>(hipStream_t, const GemmBatch&, int);

}  // namespace negf
