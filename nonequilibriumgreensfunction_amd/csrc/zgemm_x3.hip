// Complex64 products on the bf16 matrix cores of gfx950, accurate to FP32.
//
// Each FP32 operand value x is split into three bf16 values, x = h + m + l with
// h = bf16(x), m = bf16(x - h), l = bf16(x - h - m) (round to nearest): |m| <= 2^-8 |x|,
// |l| <= 2^-16 |x|, and x - h - m - l is below 2^-24 |x|.  A real product X Y is formed
// from the six split products of order 2^-16 and above,
//   X Y ~ Xh Yh + Xh Ym + Xm Yh + Xm Ym + Xh Yl + Xl Yh,
// accumulated in FP32 by v_mfma_f32_16x16x32_bf16; the dropped terms (Xm Yl, Xl Ym,
// Xl Yl) and the split remainders are of order 2^-24 relative, the FP32 rounding
// level.  The complex product takes four real products (Re A Re B, Im A Im B and
// Re A Im B + Im A Re B into three accumulators; Re C = P1 - P2 at the end), so
// each component keeps the accuracy of the four-product FP32 form.  The bf16 units
// run 16 times the FP32 matrix rate; with 24 bf16 products per complex product
// (against 3 FP32 products in the 3M form) the ceiling is 2.7 times the FP32 3M
// kernel's.
//
// A 256-thread workgroup owns a 64x64 tile of C; its four wave64s own 32x32
// quadrants (2 x 2 blocks of 16 x 16).  K advances in slabs of 32.  Each thread loads
// 8 consecutive k of one row of op(A) and of one column of op(B) (FP32, into
// registers while the previous slab is multiplied), splits them once and stores the
// six bf16 planes (real and imaginary parts, three pieces each) with k contiguous, so
// that an MFMA operand (8 consecutive k of a row or column) is one 16-byte LDS read.
// One LDS stage of 60 KiB; two barriers per slab.
// The epilogues are those of k_gemm (zgemm_mfma.hip): alpha op(A) op(B) + beta C,
// or the Gauss-Jordan update keep(C) + op(A) op(B), and the per-matrix column limit.
#include "negf_device.h"

#include <utility>

namespace negf {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kXBK = 32, kXThreads = 256, kXBM = 64, kXBN = 64;
constexpr int kXLd = kXBK + 8;  // LDS row stride in bf16 (80 bytes: 16-byte aligned, no bank conflicts)

// x = h + m + l for 8 values (round to nearest at every step; the pieces are exact)
struct Split8 {
  v8bf h, m, l;
};
__device__ inline Split8 split8(const float* x) {
  Split8 s;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const __bf16 h = (__bf16)x[i];
    const float r1 = x[i] - (float)h;
    const __bf16 m = (__bf16)r1;
    s.h[i] = h;
    s.m[i] = m;
    s.l[i] = (__bf16)(r1 - (float)m);
  }
  return s;
}

__device__ inline v4f mma(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

template <int OP>
__device__ inline float2 xload(const float2* p, bool ok) {
  float2 v = make_float2(0.f, 0.f);
  if (ok) {
    v = *p;
    if (OP == OP_C) v.y = -v.y;
  }
  return v;
}

// Workgroup id -> tile, XCD-aware (the grouped order of k_gemm).
__device__ inline int x3_tile(int bid, int tiles_m, int ntiles) {
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

template <int OPA, int OPB, int EPI>
__global__ __launch_bounds__(kXThreads, 2) void k_gemm_x3(GemmBatch g, int tiles_m, int ntiles, float2 alpha,
                                                       float2 beta) {
  constexpr int BM = kXBM, BN = kXBN;
  const int b = g.active ? g.active[blockIdx.z] : (int)blockIdx.z;
  const int tile = x3_tile(blockIdx.x, tiles_m, ntiles);
  if (tile >= ntiles) return;
  const int m0 = (tile % tiles_m) * BM;
  const int n0 = (tile / tiles_m) * BN;
  if (m0 >= g.m || n0 >= g.n) return;
  if (EPI == EPI_GJ && n0 >= g.skip_lo && n0 + BN <= g.skip_hi) return;
  if (EPI != EPI_GJ && g.ncol && n0 >= g.ncol[(long long)b * g.sNcol]) {
    // the tile lies past this matrix's columns, where B is zero
    const bool b0 = beta.x == 0.f && beta.y == 0.f, b1 = beta.x == 1.f && beta.y == 0.f;
    if (b0) {
      float2* C0 = (float2*)g.C + (long long)b * g.sC;
      for (int e = threadIdx.x; e < BM * BN; e += kXThreads) {
        const int row = m0 + e % BM, col = n0 + e / BM;
        if (row < g.m && col < g.n) C0[(long long)col * g.ldc + row] = make_float2(0.f, 0.f);
      }
    }
    if (b0 || b1) return;
  }
  const float2* A = (const float2*)g.A + (long long)b * g.sA;
  const float2* B = (const float2*)g.B + (long long)b * g.sB;
  float2* C = (float2*)g.C + (long long)b * g.sC;
  const int K = g.k, M = g.m, N = g.n;

  // planes: 0..2 the real part's pieces h, m, l; 3..5 the imaginary part's
  __shared__ __align__(16) __bf16 sA[6][BM][kXLd];  // [plane][m][k]
  __shared__ __align__(16) __bf16 sB[6][BN][kXLd];  // [plane][n][k]

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  // p1 = Re A Re B, p2 = Im A Im B, p3 = Re A Im B, p4 = Im A Re B per 16x16 block
  v4f p1[2][2], p2[2][2], p3[2][2], p4[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) p1[i][j] = p2[i][j] = p3[i][j] = p4[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // This thread's row of op(A) and column of op(B), and its 8 k of each slab.  Where
  // consecutive k are strided in memory (op(A) N, op(B) T) a wave takes 64
  // consecutive rows (columns) of one k; where they are contiguous (op(A) T/C,
  // op(B) N) each thread reads its 8 contiguous k.
  const int am = OPA == OP_N ? tid % BM : tid / 4, ak = OPA == OP_N ? tid / BM : tid % 4;
  const int bn = OPB == OP_N ? tid / 4 : tid % BN, bk = OPB == OP_N ? tid % 4 : tid / BN;
  float2 ra[8], rb[8];
  auto fetch = [&](int k0) {
    const int r = m0 + am, col = n0 + bn;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int qa = k0 + 8 * ak + i, qb = k0 + 8 * bk + i;
      ra[i] = xload<OPA>(OPA == OP_N ? A + (long long)qa * g.lda + r : A + (long long)r * g.lda + qa, r < M && qa < K);
      rb[i] = xload<OPB>(OPB == OP_N ? B + (long long)col * g.ldb + qb : B + (long long)qb * g.ldb + col,
                         qb < K && col < N);
    }
  };
  // split into registers (VALU work that overlaps the MFMAs in flight), then store
  Split8 sar, sai, sbr, sbi;
  auto split = [&]() {
    float xr[8], xi[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) { xr[i] = ra[i].x; xi[i] = ra[i].y; }
    sar = split8(xr);
    sai = split8(xi);
#pragma unroll
    for (int i = 0; i < 8; ++i) { xr[i] = rb[i].x; xi[i] = rb[i].y; }
    sbr = split8(xr);
    sbi = split8(xi);
  };
  auto stash = [&]() {
    *(v8bf*)&sA[0][am][8 * ak] = sar.h;
    *(v8bf*)&sA[1][am][8 * ak] = sar.m;
    *(v8bf*)&sA[2][am][8 * ak] = sar.l;
    *(v8bf*)&sA[3][am][8 * ak] = sai.h;
    *(v8bf*)&sA[4][am][8 * ak] = sai.m;
    *(v8bf*)&sA[5][am][8 * ak] = sai.l;
    *(v8bf*)&sB[0][bn][8 * bk] = sbr.h;
    *(v8bf*)&sB[1][bn][8 * bk] = sbr.m;
    *(v8bf*)&sB[2][bn][8 * bk] = sbr.l;
    *(v8bf*)&sB[3][bn][8 * bk] = sbi.h;
    *(v8bf*)&sB[4][bn][8 * bk] = sbi.m;
    *(v8bf*)&sB[5][bn][8 * bk] = sbi.l;
  };
  auto rdA = [&](int s) {
    Split8 re, im;
    const int row = wm + 16 * s + lr;
    re.h = *(const v8bf*)&sA[0][row][8 * lk];
    re.m = *(const v8bf*)&sA[1][row][8 * lk];
    re.l = *(const v8bf*)&sA[2][row][8 * lk];
    im.h = *(const v8bf*)&sA[3][row][8 * lk];
    im.m = *(const v8bf*)&sA[4][row][8 * lk];
    im.l = *(const v8bf*)&sA[5][row][8 * lk];
    return std::make_pair(re, im);
  };
  auto rdB = [&](int s) {
    Split8 re, im;
    const int col = wn + 16 * s + lr;
    re.h = *(const v8bf*)&sB[0][col][8 * lk];
    re.m = *(const v8bf*)&sB[1][col][8 * lk];
    re.l = *(const v8bf*)&sB[2][col][8 * lk];
    im.h = *(const v8bf*)&sB[3][col][8 * lk];
    im.m = *(const v8bf*)&sB[4][col][8 * lk];
    im.l = *(const v8bf*)&sB[5][col][8 * lk];
    return std::make_pair(re, im);
  };

  // Gauss-Jordan update (short K): the old C values and the row mask are loaded
  // while the first slab is in flight, so their latency overlaps the products
  const int* piv = (EPI == EPI_GJ) ? g.pivstep + (long long)b * g.sPiv : nullptr;
  float2 cpre[EPI == EPI_GJ ? 2 : 1][EPI == EPI_GJ ? 2 : 1][EPI == EPI_GJ ? 4 : 1];
  bool kpre[EPI == EPI_GJ ? 2 : 1][EPI == EPI_GJ ? 4 : 1];
  const int nk = (K + kXBK - 1) / kXBK;
  if (nk > 0) fetch(0);
  if (EPI == EPI_GJ) {
#pragma unroll
    for (int sm = 0; sm < 2; ++sm)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * sm + 4 * lk + r;
        const int ps = row < M ? piv[row] : 0;
        kpre[sm][r] = !(ps >= g.pv_lo && ps < g.pv_hi);
#pragma unroll
        for (int sn = 0; sn < 2; ++sn) {
          const int col = n0 + wn + 16 * sn + lr;
          const bool ok = row < M && col < N && !(col >= g.skip_lo && col < g.skip_hi);
          cpre[sm][sn][r] = ok ? C[(long long)col * g.ldc + row] : make_float2(0.f, 0.f);
        }
      }
  }
  if (nk > 0) {
    split();
    stash();
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) fetch((kt + 1) * kXBK);
    const auto a0 = rdA(0), a1 = rdA(1), b0 = rdB(0), b1 = rdB(1);
    // the six split products, each over all 16 accumulators (independent MFMAs back
    // to back): Xl Yh, Xh Yl, Xm Ym, Xm Yh, Xh Ym, Xh Yh
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      auto pa = [&](const Split8& x) { return t == 0 ? x.l : (t == 2 || t == 3) ? x.m : x.h; };
      auto pb = [&](const Split8& y) { return t == 1 ? y.l : (t == 2 || t == 4) ? y.m : y.h; };
#pragma unroll
      for (int sm = 0; sm < 2; ++sm) {
        const auto& a = sm ? a1 : a0;
#pragma unroll
        for (int sn = 0; sn < 2; ++sn) {
          const auto& bb = sn ? b1 : b0;
          p1[sm][sn] = mma(pa(a.first), pb(bb.first), p1[sm][sn]);
          p2[sm][sn] = mma(pa(a.second), pb(bb.second), p2[sm][sn]);
          p3[sm][sn] = mma(pa(a.first), pb(bb.second), p3[sm][sn]);
          p4[sm][sn] = mma(pa(a.second), pb(bb.first), p4[sm][sn]);
        }
      }
    }
    if (kt + 1 < nk) {
      split();
      __syncthreads();  // every wave is done reading the slab
      stash();
    }
    __syncthreads();
  }

  // Epilogue: accumulator element r of lane (lk, lr) is C[4 lk + r][lr] of its block.
#pragma unroll
  for (int sm = 0; sm < 2; ++sm)
#pragma unroll
    for (int sn = 0; sn < 2; ++sn) {
      const int col = n0 + wn + 16 * sn + lr;
      if (col >= N) continue;
      if (EPI == EPI_GJ && col >= g.skip_lo && col < g.skip_hi) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * sm + 4 * lk + r;
        if (row >= M) continue;
        float2* pc = C + (long long)col * g.ldc + row;
        float2 v = make_float2(p1[sm][sn][r] - p2[sm][sn][r], p3[sm][sn][r] + p4[sm][sn][r]);
        if (EPI == EPI_GJ) {
          if (kpre[sm][r]) v = cadd(v, cpre[sm][sn][r]);
        } else {
          v = cmul(alpha, v);
          if (beta.x != 0.f || beta.y != 0.f) v = cadd(v, cmul(beta, *pc));
        }
        *pc = v;
      }
    }
}

template <int OPA, int OPB, int EPI>
static hipError_t launch_x3(hipStream_t s, const GemmBatch& g, int batch, float2 alpha, float2 beta) {
  const int tm = (g.m + kXBM - 1) / kXBM, tn = (g.n + kXBN - 1) / kXBN;
  const int nt = tm * tn;
  hipLaunchKernelGGL((k_gemm_x3<OPA, OPB, EPI>), dim3(((nt + 7) / 8) * 8, 1, batch), dim3(kXThreads), 0, s, g,
                     tm, nt, alpha, beta);
  return hipGetLastError();
}

hipError_t gemm_x3(hipStream_t s, int opa, int opb, int epi, const GemmBatch& g, int batch, float2 alpha,
                   float2 beta) {
  if (batch <= 0 || g.m <= 0 || g.n <= 0) return hipSuccess;
  if (epi == EPI_GJ) {
    if (opa == OP_N && opb == OP_T) return launch_x3<OP_N, OP_T, EPI_GJ>(s, g, batch, alpha, beta);
    return hipErrorInvalidValue;
  }
#define NEGF_X3(a, b) \
  if (opa == a && opb == b) return launch_x3<a, b, EPI_STD>(s, g, batch, alpha, beta);
  NEGF_X3(OP_N, OP_N)
  NEGF_X3(OP_N, OP_C)
  NEGF_X3(OP_C, OP_N)
  NEGF_X3(OP_N, OP_T)
  NEGF_X3(OP_T, OP_N)
  NEGF_X3(OP_C, OP_C)
  NEGF_X3(OP_T, OP_T)
#undef NEGF_X3
  return hipErrorInvalidValue;
}

}  // namespace negf
