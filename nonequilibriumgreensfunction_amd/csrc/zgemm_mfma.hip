// Batched complex GEMM on CDNA4 matrix cores (gfx950).
//
// The complex product is built from four real MFMA products per 16x16x4 step:
//   Re C += Re A Re B - Im A Im B,   Im C += Re A Im B + Im A Re B
// using v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32).
// A 256-thread workgroup owns a 64x64 tile of C; its four wave64s each own a
// 32x32 quadrant (2x2 MFMA blocks, separate real and imaginary accumulators).
// Tiles of op(A) (64x16) and op(B) (16x64) are staged through LDS as split
// real/imaginary planes, k-major, so that each MFMA operand read is 16
// consecutive LDS words; the next K slab is fetched into registers while the
// current one is multiplied.
//
// This replaces the reference's MKL zgemm call sites
// (Recursive_Greens.F90:18-95, Contact_Self_Energy.F90:147-209): square
// m = n = k products with op in {N, C} on either operand.  The Gauss-Jordan
// epilogue turns the same kernel into the trailing update of the blocked
// inverse (gj_inverse.hip).
#include "negf_device.h"

#include <cstdlib>

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

constexpr int kBM = 64, kBN = 64, kBK = 16, kThreads = 256;
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

// Maps a linear workgroup id to a tile so that the tiles of one XCD (ids
// congruent mod 8 are dispatched to the same XCD) form a contiguous run of
// column-major tiles: neighbouring tiles share A/B panels in that XCD's L2.
__device__ inline int xcd_tile(int bid, int ntiles) {
  int per = (ntiles + 7) >> 3;
  return (bid & 7) * per + (bid >> 3);
}

template <typename T, int OPA, int OPB, int EPI>
__global__ __launch_bounds__(kThreads) void k_gemm(GemmBatch g, int tiles_m, int ntiles,
                                                   cx<T> alpha, cx<T> beta) {
  typedef typename Mfma<T>::acc acc_t;
  const int b = g.active ? g.active[blockIdx.z] : (int)blockIdx.z;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int m0 = (tile % tiles_m) * kBM;
  const int n0 = (tile / tiles_m) * kBN;
  if (m0 >= g.m || n0 >= g.n) return;
  if (EPI == EPI_GJ && n0 >= g.skip_lo && n0 + kBN <= g.skip_hi) return;

  const cx<T>* A = (const cx<T>*)g.A + (long long)b * g.sA;
  const cx<T>* B = (const cx<T>*)g.B + (long long)b * g.sB;
  cx<T>* C = (cx<T>*)g.C + (long long)b * g.sC;
  const int K = g.k, M = g.m, N = g.n;

  __shared__ T sA[2][kBK][kBM + kPad];
  __shared__ T sB[2][kBK][kBN + kPad];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  acc_t cr[2][2], ci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = acc_t{0, 0, 0, 0};
      ci[i][j] = acc_t{0, 0, 0, 0};
    }

  cx<T> ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int c = tid + kThreads * i;
      if (OPA == OP_N) {
        int mm = c & 63, kk = c >> 6;
        int r = m0 + mm, q = k0 + kk;
        ra[i] = load_op<T, OPA>(A + (long long)q * g.lda + r, r < M && q < K);
      } else {
        int kk = c & 15, mm = c >> 4;
        int r = m0 + mm, q = k0 + kk;
        ra[i] = load_op<T, OPA>(A + (long long)r * g.lda + q, r < M && q < K);
      }
      if (OPB == OP_N) {
        int kk = c & 15, nn = c >> 4;
        int q = k0 + kk, col = n0 + nn;
        rb[i] = load_op<T, OPB>(B + (long long)col * g.ldb + q, q < K && col < N);
      } else {
        int nn = c & 63, kk = c >> 6;
        int q = k0 + kk, col = n0 + nn;
        rb[i] = load_op<T, OPB>(B + (long long)q * g.ldb + col, q < K && col < N);
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int c = tid + kThreads * i;
      int mm, kk;
      if (OPA == OP_N) { mm = c & 63; kk = c >> 6; } else { kk = c & 15; mm = c >> 4; }
      sA[0][kk][mm] = ra[i].x;
      sA[1][kk][mm] = ra[i].y;
      int nn;
      if (OPB == OP_N) { kk = c & 15; nn = c >> 4; } else { nn = c & 63; kk = c >> 6; }
      sB[0][kk][nn] = rb[i].x;
      sB[1][kk][nn] = rb[i].y;
    }
  };

  fetch(0);
  for (int k0 = 0; k0 < K; k0 += kBK) {
    stash();
    __syncthreads();
    if (k0 + kBK < K) fetch(k0 + kBK);
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      const int kk = ks * 4 + lk;
      T ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        ar[s] = sA[0][kk][wm + 16 * s + lr];
        ai[s] = sA[1][kk][wm + 16 * s + lr];
        nai[s] = -ai[s];
        br[s] = sB[0][kk][wn + 16 * s + lr];
        bi[s] = sB[1][kk][wn + 16 * s + lr];
      }
#pragma unroll
      for (int sm = 0; sm < 2; ++sm)
#pragma unroll
        for (int sn = 0; sn < 2; ++sn) {
          cr[sm][sn] = Mfma<T>::mma(ar[sm], br[sn], cr[sm][sn]);
          ci[sm][sn] = Mfma<T>::mma(ar[sm], bi[sn], ci[sm][sn]);
          cr[sm][sn] = Mfma<T>::mma(nai[sm], bi[sn], cr[sm][sn]);
          ci[sm][sn] = Mfma<T>::mma(ai[sm], br[sn], ci[sm][sn]);
        }
    }
    __syncthreads();
  }

  // Epilogue: accumulator element r of lane (lk, lr) is C[Mfma<T>::row(lk, r)][lr]
  // of its 16x16 block.
  const int* piv = (EPI == EPI_GJ) ? g.pivstep + (long long)b * g.sPiv : nullptr;
#pragma unroll
  for (int sm = 0; sm < 2; ++sm)
#pragma unroll
    for (int sn = 0; sn < 2; ++sn) {
      const int col = n0 + wn + 16 * sn + lr;
      if (col >= N) continue;
      if (EPI == EPI_GJ && col >= g.skip_lo && col < g.skip_hi) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * sm + Mfma<T>::row(lk, r);
        if (row >= M) continue;
        cx<T>* pc = C + (long long)col * g.ldc + row;
        cx<T> v = mk<T>(cr[sm][sn][r], ci[sm][sn][r]);
        if (EPI == EPI_GJ) {
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

static hipEvent_t* timer_slot(double flops) {
  GemmTimer& t = gemm_timer();
  if (!t.on) return nullptr;
  if (t.used == t.cap) {
    size_t nc = t.cap ? 2 * t.cap : 4096;
    hipEvent_t* ev = (hipEvent_t*)realloc(t.ev, sizeof(hipEvent_t) * 2 * nc);
    double* fl = (double*)realloc(t.flops, sizeof(double) * nc);
    if (!ev || !fl) return nullptr;
    for (size_t i = 2 * t.cap; i < 2 * nc; ++i) hipEventCreate(&ev[i]);
    t.ev = ev;
    t.flops = fl;
    t.cap = nc;
  }
  t.flops[t.used] = flops;
  return &t.ev[2 * t.used++];
}

template <typename T, int OPA, int OPB, int EPI>
static hipError_t launch(hipStream_t s, const GemmBatch& g, int batch, cx<T> alpha, cx<T> beta) {
  if (batch <= 0 || g.m <= 0 || g.n <= 0) return hipSuccess;
  int tm = (g.m + kBM - 1) / kBM, tn = (g.n + kBN - 1) / kBN;
  int nt = tm * tn;
  dim3 grid(((nt + 7) / 8) * 8, 1, batch);
  const int ncols = (EPI == EPI_GJ) ? g.n - (g.skip_hi - g.skip_lo) : g.n;
  hipEvent_t* ev = timer_slot(8.0 * g.m * (double)ncols * g.k * batch);
  if (ev) hipEventRecord(ev[0], s);
  hipLaunchKernelGGL((k_gemm<T, OPA, OPB, EPI>), grid, dim3(kThreads), 0, s, g, tm, nt, alpha, beta);
  if (ev) hipEventRecord(ev[1], s);
  return hipGetLastError();
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
