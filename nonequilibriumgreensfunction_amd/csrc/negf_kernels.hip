// Bandwidth-bound kernels of the per-energy NEGF solve (gfx950).
//
// The dense n^3 work runs in the MFMA GEMM and the Gauss-Jordan inverse; what
// remains is O(n^2) per slice and is HBM/L2-bound:
//  * k_assemble   -- the RGF slice matrix  z I - H_x - Sigma - X' g X,  fused:
//                    the reference forms it from two dense zgemm with -HopX
//                    (Recursive_Greens.F90:18-20, :70-72) plus the InvGRd
//                    assembly (Energy_Sweep.F90:31-38); HopX is site-block-diagonal
//                    (Build_Hamiltonian.F90:62-64), so the sandwich costs
//                    norb^2 multiply-adds per element instead of 2n;
//  * k_bd_mul     -- HopX * G and HopX' * G for the block-column recursions
//                    (Recursive_Greens.F90:61-62, :94-95);
//  * k_siteblocks -- the norb x norb site-diagonal blocks of B G' (in the solve
//                    B = G = V = G W with Gamma = W W', so B G' = G Gamma G'),
//                    reduced to the traces the density (Calculate_Gnp.F90:118-124)
//                    and spin (Calculate_Observables.F90:10-30) need;
//  * k_inner      -- Re sum_ij P_ij conj(Q_ij), the trace form of T12 and the
//                    injected modes (Calculate_Observables.F90:48-53, :86-100);
//  * Sancho-Rubio helpers (Contact_Self_Energy.F90:110-223).
// Every reduction runs in a fixed order inside one workgroup, so an energy's
// result does not depend on what else is in its batch.
#include "negf_device.h"

#include <algorithm>

namespace negf {

// ---------------------------------------------------------------------------
// Slice assembly.  For batch item b (an energy):
//   M_b = hsign*H + [diag] z_b I - Sig1_b - Sig2_b - D_b
// with z_b = energy[b] + i*eta and, by MODE,
//   MODE 0: D = 0,  MODE 1: D = X' g_b X,  MODE 2: D = X g_b X',  MODE 3: D = P_b.
// X is site-block-diagonal with norb x norb blocks (MODE 1/2 read only those).
// V (optional): per-site potentials of this slice, added to the diagonal of H on
// every orbital of the site (compact Hamiltonian: slice = template + potentials).
template <typename T, int NORB, int MODE>
__global__ void k_assemble(cx<T>* M, long long sM, const cx<T>* H, int add_diag, T hsign,
                           const double* energy, double eta, const cx<T>* S1, long long sS1,
                           const cx<T>* S2, long long sS2, const cx<T>* g, long long sg,
                           const cx<T>* X, int n, const int* active, const double* V) {
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const int c = blockIdx.y;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const long long e = (long long)c * n + a;
  cx<T> v = H[e];
  if (V && a == c) v.x += (T)V[a / NORB];
  v.x *= hsign;
  v.y *= hsign;
  if (add_diag && a == c) {
    v.x += (T)energy[b];
    v.y += (T)eta;
  }
  if (S1) v = csub(v, S1[(long long)b * sS1 + e]);
  if (S2) v = csub(v, S2[(long long)b * sS2 + e]);
  if (MODE == 1 || MODE == 2) {
    const cx<T>* G = g + (long long)b * sg;
    const int a0 = (a / NORB) * NORB, r = a - a0;
    const int c0 = (c / NORB) * NORB, s = c - c0;
    cx<T> d = mk<T>(0, 0);
#pragma unroll
    for (int p = 0; p < NORB; ++p) {
      // t = (g X)[a0+p][c] (MODE 1)  or  (g X')[a0+p][c] (MODE 2)
      cx<T> t = mk<T>(0, 0);
#pragma unroll
      for (int q = 0; q < NORB; ++q) {
        const cx<T> gv = G[(long long)(c0 + q) * n + a0 + p];
        if (MODE == 1)
          t = cadd(t, cmul(gv, X[(long long)(c0 + s) * n + c0 + q]));
        else
          t = cadd(t, cmulc(gv, X[(long long)(c0 + q) * n + c0 + s]));
      }
      if (MODE == 1)
        d = cadd(d, cmulc(t, X[(long long)(a0 + r) * n + a0 + p]));  // conj(X[p][r]) t
      else
        d = cadd(d, cmul(X[(long long)(a0 + p) * n + a0 + r], t));   // X[r][p] t
    }
    v = csub(v, d);
  } else if (MODE == 3) {
    v = csub(v, g[(long long)b * sg + e]);
  }
  M[(long long)b * sM + e] = v;
}

// The same slice matrix as k_assemble MODE 1 (z I - H_x - Sigma - X' g X), one thread
// per NORB x NORB site block (a, c): the g block is read once (complex64 or complex128:
// the widening to FP64 is exact, so both give the bits of k_assemble on the
// complex128 copy of g) and X' g X takes 2 NORB^3 complex multiply-adds per block
// instead of 4 per element per orbital pair; the column site's X block is staged in
// LDS (one site column per workgroup row).  Each element's operations and their order
// are those of k_assemble.
template <int NORB, typename TG>
__global__ __launch_bounds__(128) void k_assemble_blk(double2* M, long long sM, const double2* H, int add_diag, double hsign,
                               const double* energy, double eta, const double2* S1, long long sS1,
                               const double2* S2, long long sS2, const TG* g, long long sg, const double2* X,
                               int n, int nsite, const double* V) {
  const int b = blockIdx.z;
  const int sc = blockIdx.y;
  const int sa = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double2 xc[NORB][NORB];  // X[c0 + q][c0 + s] as xc[q][s]
  const int c0 = sc * NORB;
  if (threadIdx.x < NORB * NORB) {
    const int q = threadIdx.x % NORB, t = threadIdx.x / NORB;
    xc[q][t] = X[(long long)(c0 + t) * n + c0 + q];
  }
  __syncthreads();
  if (sa >= nsite) return;
  const int a0 = sa * NORB;
  const TG* G = g + (long long)b * sg;
  // t[p][s] = (g X)[a0 + p][c0 + s]
  double2 t[NORB][NORB];
#pragma unroll
  for (int p = 0; p < NORB; ++p)
#pragma unroll
    for (int s2 = 0; s2 < NORB; ++s2) t[p][s2] = mk<double>(0, 0);
#pragma unroll
  for (int q = 0; q < NORB; ++q) {
    double2 gq[NORB];
#pragma unroll
    for (int p = 0; p < NORB; ++p) {
      const TG v = G[(long long)(c0 + q) * n + a0 + p];
      gq[p] = mk<double>((double)v.x, (double)v.y);
    }
#pragma unroll
    for (int p = 0; p < NORB; ++p)
#pragma unroll
      for (int s2 = 0; s2 < NORB; ++s2) t[p][s2] = cadd(t[p][s2], cmul(gq[p], xc[q][s2]));
  }
#pragma unroll
  for (int r = 0; r < NORB; ++r) {
    const int a = a0 + r;
    double2 xa[NORB];  // X[a0 + p][a0 + r]
#pragma unroll
    for (int p = 0; p < NORB; ++p) xa[p] = X[(long long)(a0 + r) * n + a0 + p];
#pragma unroll
    for (int s2 = 0; s2 < NORB; ++s2) {
      const int c = c0 + s2;
      const long long e = (long long)c * n + a;
      double2 v = H[e];
      if (V && a == c) v.x += V[sa];
      v.x *= hsign;
      v.y *= hsign;
      if (add_diag && a == c) {
        v.x += energy[b];
        v.y += eta;
      }
      if (S1) v = csub(v, S1[(long long)b * sS1 + e]);
      if (S2) v = csub(v, S2[(long long)b * sS2 + e]);
      double2 d = mk<double>(0, 0);
#pragma unroll
      for (int p = 0; p < NORB; ++p) d = cadd(d, cmulc(t[p][s2], xa[p]));  // conj(X[p][r]) t
      M[(long long)b * sM + e] = csub(v, d);
    }
  }
}

// Y_b = X G_b (CONJ = 0) or X' G_b (CONJ = 1), X site-block-diagonal; accum: Y_b +=.
template <typename T, int NORB, int CONJ>
__global__ void k_bd_mul(cx<T>* Y, long long sY, const cx<T>* G, long long sG, const cx<T>* X,
                         int n, const int* active, int accum) {
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const int col = blockIdx.y;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int a0 = (a / NORB) * NORB, r = a - a0;
  const cx<T>* Gb = G + (long long)b * sG + (long long)col * n;
  cx<T> v = mk<T>(0, 0);
#pragma unroll
  for (int q = 0; q < NORB; ++q) {
    if (CONJ)
      v = cadd(v, cmul(cconj(X[(long long)(a0 + r) * n + a0 + q]), Gb[a0 + q]));
    else
      v = cadd(v, cmul(X[(long long)(a0 + q) * n + a0 + r], Gb[a0 + q]));
  }
  cx<T>& y = Y[(long long)b * sY + (long long)col * n + a];
  y = accum ? cadd(y, v) : v;
}

// Site-diagonal blocks D_s = (B G')[site s], over the columns [kfirst, min(kc[b * skc],
// klast)) of B and G (kc null: n).
// Writes (accum: adds), per site s of slice x:  out[0] = Re tr D,  out[1+a] =
// Re tr(D sigma_a), out[4+r] = Re D_rr (r < NORB).  out layout: out_b + q*qstride + s.
// A workgroup takes 16 rows (16 / NORB sites) and splits the sum over k into 16
// chunks of the matrix's own column range, one per 16-thread row group: with 16 rows
// per workgroup there are n / 16 workgroups per matrix (144 at n = 2304), enough to
// spread the stream of B and G over the whole GPU.  The 16 chunk sums are added in
// chunk order by one thread per site, so the result does not depend on the batch.
template <typename T, int NORB>
__global__ __launch_bounds__(256) void k_siteblocks(const cx<T>* Bm, long long sB,
                                                    const cx<T>* G, long long sG, int n,
                                                    double* out, long long sOut, long long qstride,
                                                    SpinMats sp, const int* kc, long long skc,
                                                    const int* active, int kfirst, int klast,
                                                    int accum) {
  constexpr int RW = 16, KQ = 16;
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const int kcols = min(kc ? kc[(long long)b * skc] : n, klast);  // columns of B and G taken
  const cx<T>* Bb = Bm + (long long)b * sB;
  const cx<T>* Gb = G + (long long)b * sG;
  const int row0 = blockIdx.x * RW;
  const int tr = threadIdx.x % RW, kq = threadIdx.x / RW;
  const int a = row0 + tr;
  const int a0 = (a / NORB) * NORB;
  double accr[NORB], acci[NORB];
#pragma unroll
  for (int q = 0; q < NORB; ++q) accr[q] = acci[q] = 0.0;
  if (a < n) {
    // the columns [kfirst, kcols) in 16 chunks fixed by this matrix's own range, so
    // the sum order does not depend on the rest of the batch
    const int k_lo = kfirst, k_n = kcols > kfirst ? kcols - kfirst : 0;
    const int kbeg = k_lo + (k_n * kq) / KQ, kend = k_lo + (k_n * (kq + 1)) / KQ;
    // U columns' loads issued together, then added in column order (the same sums as
    // one column at a time): the loop is latency-bound, not bandwidth-bound
    constexpr int U = sizeof(T) == sizeof(float) ? 8 : 4;
    int k = kbeg;
    for (; k + U <= kend; k += U) {
      cx<T> bv[U], gv[U][NORB];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        bv[u] = Bb[(long long)(k + u) * n + a];
#pragma unroll
        for (int q = 0; q < NORB; ++q) gv[u][q] = Gb[(long long)(k + u) * n + a0 + q];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < NORB; ++q) {
          accr[q] += (double)bv[u].x * gv[u][q].x + (double)bv[u].y * gv[u][q].y;
          acci[q] += (double)bv[u].y * gv[u][q].x - (double)bv[u].x * gv[u][q].y;
        }
    }
    for (; k < kend; ++k) {
      const cx<T> bv = Bb[(long long)k * n + a];
#pragma unroll
      for (int q = 0; q < NORB; ++q) {
        const cx<T> gv = Gb[(long long)k * n + a0 + q];
        accr[q] += (double)bv.x * gv.x + (double)bv.y * gv.y;
        acci[q] += (double)bv.y * gv.x - (double)bv.x * gv.y;
      }
    }
  }
  __shared__ double sr[KQ][RW][NORB], si[KQ][RW][NORB];
#pragma unroll
  for (int q = 0; q < NORB; ++q) {
    sr[kq][tr][q] = accr[q];
    si[kq][tr][q] = acci[q];
  }
  __syncthreads();
  // one thread per site of this row group
  constexpr int nsite_grp = RW / NORB;
  if (threadIdx.x < nsite_grp) {
    const int sl = threadIdx.x;
    const int s = (row0 / NORB) + sl;
    if (s * NORB < n) {
      // each entry D(r, q) once, in chunk order, into the trace, the per-orbital
      // diagonal and the three spin traces (summed over r, then q, as written out)
      double* o = out + (long long)b * sOut + s;
      auto put = [&](long long off, double v) { o[off] = accum ? o[off] + v : v; };
      double trace = 0.0, spin[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < NORB; ++r)
#pragma unroll
        for (int q = 0; q < NORB; ++q) {
          const int t = sl * NORB + r;
          double xr = 0.0, xi = 0.0;
#pragma unroll 1
          for (int c = 0; c < KQ; ++c) {
            xr += sr[c][t][q];
            xi += si[c][t][q];
          }
          if (r == q) {
            trace += xr;
            put((4 + r) * qstride, xr);  // per orbital (the orbital-resolved density)
          }
#pragma unroll
          for (int m = 0; m < 3; ++m) spin[m] += xr * sp.re[m][q][r] - xi * sp.im[m][q][r];
        }
      put(0, trace);
#pragma unroll
      for (int m = 0; m < 3; ++m) put((m + 1) * qstride, spin[m]);
    }
  }
}

// Bond-current terms of one slice from the neighbour site blocks of A = V V'
// (V = G W, Gamma = W W', so A = G Gamma G' is one lead's Gn/Gp up to the
// occupation factor), Calculate_Observables.F90:111-168 before the q^2/hbar and
// occupation factors.  For site s (orbitals a = NORB s + ii) and a bond to b:
//   y: b on site s + 1 (y < Npy-1), z: b on site s + Npy (z < Npz-1), both inside
//      the slice, H = this slice's block;
//   x: b the same orbitals of slice x+1 (Vn, when given), H_ab = HopX(a, b'),
//      H_ba = conj(HopX(a, b')) as the reference writes it;
// with M = A_ba (row b, column a):  out = Re i (H_ab M - conj(M) H_ba) = -Im(...),
// summed over the orbital pairs at out_b + (8 + d) * qstride + s, d = 0 x, 1 y, 2 z,
// and, when `pairs`, per orbital pair at out_b + (11 + 16 d + NORB ii + jj) * qstride + s
// (the orbital-resolved currents); 0 where the bond does not exist.
// Reduction as in k_siteblocks (over the columns [kfirst, min(kc[b * skc], klast)),
// written or, with accum, added): 16 rows per workgroup, the k sum in 16 chunks
// added in chunk order, so results do not depend on the batch.
template <typename T, int NORB>
__global__ __launch_bounds__(256) void k_bondblocks(const cx<T>* V, long long sV, const cx<T>* Vn,
                                                    long long sVn, const double2* H, const double2* X,
                                                    int n, int npy, int npz, double* out,
                                                    long long sOut, long long qstride, int pairs,
                                                    const int* kc, long long skc, const int* active,
                                                    int kfirst, int klast, int accum) {
  constexpr int RW = 16, KQ = 16, NQ = 3 * NORB;
  const int b = active ? active[blockIdx.z] : (int)blockIdx.z;
  const int kcols = min(kc ? kc[(long long)b * skc] : n, klast);  // columns of V and Vn taken
  const cx<T>* Vb = V + (long long)b * sV;
  const cx<T>* Vnb = Vn ? Vn + (long long)b * sVn : nullptr;
  const int NN = NORB * npy;
  const int row0 = blockIdx.x * RW;
  const int tr = threadIdx.x % RW, kq = threadIdx.x / RW;
  const int a = row0 + tr;
  const int site = a / NORB, a0 = site * NORB;
  const int y = site % npy, z = site / npy;
  const bool hy = y < npy - 1, hz = z < npz - 1, hx = Vnb != nullptr;
  // acc[d * NORB + q] = sum_k V_nb(b0_d + q, k) conj(V(a, k))
  double accr[NQ], acci[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) accr[q] = acci[q] = 0.0;
  if (a < n) {
    const int k_lo = kfirst, k_n = kcols > kfirst ? kcols - kfirst : 0;
    const int kbeg = k_lo + (k_n * kq) / KQ, kend = k_lo + (k_n * (kq + 1)) / KQ;
    // one column's terms, in the order of the sums (as k_siteblocks, U columns' loads
    // are issued together and their terms then added in column order)
    auto terms = [&](const cx<T>& va, const cx<T>* wx, const cx<T>* wy, const cx<T>* wz) {
#pragma unroll
      for (int q = 0; q < NORB; ++q) {
        if (hx) {
          accr[q] += (double)wx[q].x * va.x + (double)wx[q].y * va.y;
          acci[q] += (double)wx[q].y * va.x - (double)wx[q].x * va.y;
        }
        if (hy) {
          accr[NORB + q] += (double)wy[q].x * va.x + (double)wy[q].y * va.y;
          acci[NORB + q] += (double)wy[q].y * va.x - (double)wy[q].x * va.y;
        }
        if (hz) {
          accr[2 * NORB + q] += (double)wz[q].x * va.x + (double)wz[q].y * va.y;
          acci[2 * NORB + q] += (double)wz[q].y * va.x - (double)wz[q].x * va.y;
        }
      }
    };
    auto load = [&](int k, cx<T>& va, cx<T>* wx, cx<T>* wy, cx<T>* wz) {
      const long long col = (long long)k * n;
      va = Vb[col + a];
#pragma unroll
      for (int q = 0; q < NORB; ++q) {
        if (hx) wx[q] = Vnb[col + a0 + q];
        if (hy) wy[q] = Vb[col + a0 + NORB + q];
        if (hz) wz[q] = Vb[col + a0 + NN + q];
      }
    };
    constexpr int U = sizeof(T) == sizeof(float) ? 2 : 1;  // (more would halve the occupancy)
    int k = kbeg;
    for (; k + U <= kend; k += U) {
      cx<T> va[U], wx[U][NORB], wy[U][NORB], wz[U][NORB];
#pragma unroll
      for (int u = 0; u < U; ++u) load(k + u, va[u], wx[u], wy[u], wz[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) terms(va[u], wx[u], wy[u], wz[u]);
    }
    for (; k < kend; ++k) {
      cx<T> va, wx[NORB], wy[NORB], wz[NORB];
      load(k, va, wx, wy, wz);
      terms(va, wx, wy, wz);
    }
  }
  __shared__ double sr[KQ][RW][NQ], si[KQ][RW][NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    sr[kq][tr][q] = accr[q];
    si[kq][tr][q] = acci[q];
  }
  __syncthreads();
  constexpr int nsite_grp = RW / NORB;
  if (threadIdx.x < nsite_grp) {
    const int sl = threadIdx.x;
    const int s = (row0 / NORB) + sl;
    if (s * NORB < n) {
      const int sy = s % npy, sz = s / npy;
      const bool has[3] = {hx, sy < npy - 1, sz < npz - 1};
      const int sa0 = s * NORB;
      double* o = out + (long long)b * sOut + s;
      for (int d = 0; d < 3; ++d) {
        double sum = 0.0;
        for (int ii = 0; ii < NORB; ++ii)
          for (int jj = 0; jj < NORB; ++jj) {
            double val = 0.0;
            if (has[d]) {
              double mr = 0.0, mi = 0.0;
              for (int c = 0; c < KQ; ++c) {
                mr += sr[c][sl * NORB + ii][d * NORB + jj];
                mi += si[c][sl * NORB + ii][d * NORB + jj];
              }
              const int ai = sa0 + ii;
              double2 hab, hba;
              if (d == 0) {
                hab = X[(long long)(sa0 + jj) * n + ai];
                hba = make_double2(hab.x, -hab.y);
              } else {
                const int bj = (d == 1 ? sa0 + NORB : sa0 + NN) + jj;
                hab = H[(long long)bj * n + ai];
                hba = H[(long long)ai * n + bj];
              }
              // Im(H_ab M - conj(M) H_ba)
              const double im = (hab.x * mi + hab.y * mr) - (mr * hba.y - mi * hba.x);
              val = -im;
            }
            sum += val;
            if (pairs) {
              double& op = o[(11 + 16 * d + NORB * ii + jj) * qstride];
              op = accum ? op + val : val;
            }
          }
        double& os = o[(8 + d) * qstride];
        os = accum ? os + sum : sum;
      }
    }
  }
}

// out[b*sOut] = Re sum_{ij} P_ij conj(Q_ij) (accum: +=) over the leading m x ncols
// part of blocks with leading dimension ld (one workgroup).  Element (i, j) goes to thread
// (j ld + i) mod 1024 whatever m and ncols are, so entries that are zero leave the
// sum's bits unchanged when m or ncols grow.
template <typename T>
__global__ __launch_bounds__(1024) void k_inner(const cx<T>* P, long long sP, const cx<T>* Q,
                                                long long sQ, int ld, int m, int ncols, double* out,
                                                long long sOut, const int* active, int accum) {
  const int b = active ? active[blockIdx.x] : (int)blockIdx.x;
  const cx<T>* Pb = P + (long long)b * sP;
  const cx<T>* Qb = Q + (long long)b * sQ;
  double acc = 0.0;
  const long long end = (long long)ld * ncols;
  for (long long e = threadIdx.x; e < end; e += 1024) {
    if ((int)(e % ld) >= m) continue;
    const cx<T> p = Pb[e], q = Qb[e];
    acc += (double)p.x * q.x + (double)p.y * q.y;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  __shared__ double red[16];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += red[w];
    out[(long long)b * sOut] = accum ? out[(long long)b * sOut] + t : t;
  }
}

// ---------------------------------------------------------------------------
// Batched elementwise helpers over n x n blocks (blockIdx.y = batch slot).

// Y = I
template <typename T>
__global__ void k_eye(cx<T>* Y, long long sY, int n, const int* active) {
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  const long long nn = (long long)n * n;
  cx<T>* Yb = Y + (long long)b * sY;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i % n), c = (int)(i / n);
    Yb[i] = mk<T>(r == c ? (T)1 : (T)0, 0);
  }
}

// Y = X + Z  (Z may alias Y)
template <typename T>
__global__ void k_add(cx<T>* Y, long long sY, const cx<T>* X, long long sX, const cx<T>* Z,
                      long long sZ, long long nn, const int* active) {
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x)
    Y[(long long)b * sY + i] = cadd(X[(long long)b * sX + i], Z[(long long)b * sZ + i]);
}

// Mirror sectors (SymMap): for matrix b the sector blocks M+ at M + b*sM and M- at
// M + b*sM + off1 (d x d each) give the n x n matrix V+ M+ V+' + V- M- V-'.  With
// S = (M+ + M-)/2 and D = (M+ - M-)/2 at (a, c): Y(r_a, r_c) = S, Y(r_a, m_c) =
// i^-ph_c D, Y(m_a, r_c) = i^ph_a D, Y(m_a, m_c) = i^(ph_a - ph_c) S.  The halving
// and the unit phases are exact, so every entry is one rounded sum or difference.
// With `pair` (sectors related by a second mirror) M- is not stored: M-(a, c) =
// i^(k_a - k_c) M+(s_a, s_c) with (s_a, k_a) = pair[a].
__device__ inline double2 sym_minus(const double2* M0, const double2* M1, const int2* pair, long long i, int a,
                                    int cc, int d) {
  if (!pair) return M1[i];
  const int2 pa = pair[a], pc = pair[cc];
  return rot_i(M0[pa.x + (long long)pc.x * d], pa.y + 4 - pc.y);
}

__global__ __launch_bounds__(256) void k_sym_expand(const double2* M, long long sM, long long off1,
                                                    const int2* pair, int d, double2* Y, long long sY, int n,
                                                    SymMap map) {
  const int b = blockIdx.y;
  const double2* M0 = M + (long long)b * sM;
  const double2* M1 = M0 + off1;
  double2* Yb = Y + (long long)b * sY;
  const long long dd = (long long)d * d;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < dd; i += (long long)gridDim.x * 256) {
    const int a = (int)(i % d), cc = (int)(i / d);
    const double2 p = M0[i], q = sym_minus(M0, M1, pair, i, a, cc, d);
    double2 S, D;
    S.x = 0.5 * (p.x + q.x); S.y = 0.5 * (p.y + q.y);
    D.x = 0.5 * (p.x - q.x); D.y = 0.5 * (p.y - q.y);
    int ra, ma, pa, rc, mc, pc;
    sym_index(map, a, ra, ma, pa);
    sym_index(map, cc, rc, mc, pc);
    Yb[ra + (long long)rc * n] = S;
    Yb[ra + (long long)mc * n] = rot_i(D, 4 - pc);
    Yb[ma + (long long)rc * n] = rot_i(D, pa);
    Yb[ma + (long long)mc * n] = rot_i(S, pa + 4 - pc);
  }
}

// out[b] = max |entry| of the matrix k_sym_expand would form from the sector blocks
// of matrix b (through `active`): the phases have unit modulus, so it is the larger
// of max|S| and max|D|; NaN as k_maxabs.  out[b] cleared by k_maxabs_clear first.
__global__ __launch_bounds__(256) void k_sym_maxabs(const double2* M, long long sM, long long off1,
                                                    const int2* pair, int d, double* out, const int* active) {
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  const double2* M0 = M + (long long)b * sM;
  const double2* M1 = M0 + off1;
  const long long dd = (long long)d * d;
  double m = 0.0;
  bool nan = false;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < dd; i += (long long)gridDim.x * 256) {
    const double2 p = M0[i], q = sym_minus(M0, M1, pair, i, (int)(i % d), (int)(i / d), d);
    const double vs = hypot(0.5 * (p.x + q.x), 0.5 * (p.y + q.y));
    const double vd = hypot(0.5 * (p.x - q.x), 0.5 * (p.y - q.y));
    if (vs != vs || vd != vd) nan = true;
    m = fmax(m, fmax(vs, vd));
  }
  for (int off = 32; off > 0; off >>= 1) {
    m = fmax(m, __shfl_xor(m, off));
    nan = nan || __shfl_xor((int)nan, off);
  }
  __shared__ double red[4];
  __shared__ int rn[4];
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = m;
    rn[threadIdx.x >> 6] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    const bool an = rn[0] | rn[1] | rn[2] | rn[3];
    const unsigned long long bits = an ? 0x7ff8000000000000ull : (unsigned long long)__double_as_longlong(t);
    atomicMax((unsigned long long*)(out + b), bits);
  }
}

// Site-block stencil of a slice (site = y + Npy z): the site itself and its y and z
// neighbours, periodic wrap included.  Slice Hamiltonians built from on-site blocks
// and nearest-neighbour y/z hoppings (Build_Hamiltonian.F90:66-95, with or without
// periodic boundaries, Peierls phases and Zeeman terms) have no other nonzero blocks.
__device__ inline bool stencil_pair(int si, int sj, int npy, int npz) {
  if (si == sj) return true;
  const int yi = si % npy, zi = si / npy, yj = sj % npy, zj = sj / npy;
  if (zi == zj) {
    const int d = yi > yj ? yi - yj : yj - yi;
    return d == 1 || d == npy - 1;
  }
  if (yi == yj) {
    const int d = zi > zj ? zi - zj : zj - zi;
    return d == 1 || d == npz - 1;
  }
  return false;
}

// flag[0] = 1 when some slice x < nslices (stride sH) has a nonzero entry outside the
// stencil (a plain store of the same value by every finder; flag cleared before).
__global__ __launch_bounds__(256) void k_stencil_check(const double2* H, long long sH, int n, int npy, int npz,
                                                       int norb, int* flag) {
  const double2* Hx = H + (long long)blockIdx.y * sH;
  const int nn = n * n;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < nn; e += gridDim.x * 256) {
    const double2 h = Hx[e];
    const int i = e % n, j = e / n;
    if ((h.x != 0.0 || h.y != 0.0) && !stencil_pair(i / norb, j / norb, npy, npz)) flag[0] = 1;
  }
}

// P_x = H_x V_x (n x R, FP64) for stencil-sparse slices: grid z = b L + x over the slices
// of every matrix b; each thread holds its row's stencil entries of H_x in registers and
// takes kStencilCols columns.  H_x at H + x sH (sH = 0: one slice for all), V_x at V + b sVb
// + x sVx, P_x likewise, both n x R with leading dimension n.
constexpr int kStencilCols = 16;
template <int NORB>
__global__ __launch_bounds__(256) void k_stencil_mul(const double2* H, long long sH, const double2* V,
                                                     long long sVb, long long sVx, double2* P, long long sPb,
                                                     long long sPx, int n, int R, int L, int npy, int npz) {
  const int b = blockIdx.z / L, x = blockIdx.z % L;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2* Hx = H + (long long)x * sH;
  const double2* Vx = V + (long long)b * sVb + (long long)x * sVx;
  double2* Px = P + (long long)b * sPb + (long long)x * sPx;
  // the distinct stencil sites of row i's site (wrap-around neighbours may coincide)
  const int si = i / NORB, yi = si % npy, zi = si / npy;
  int nb[5];
  int cnt = 0;
  auto add = [&](int t) {
    for (int k = 0; k < cnt; ++k)
      if (nb[k] == t) return;
    nb[cnt++] = t;
  };
  add(si);
  add(zi * npy + (yi + 1) % npy);
  add(zi * npy + (yi + npy - 1) % npy);
  add(((zi + 1) % npz) * npy + yi);
  add(((zi + npz - 1) % npz) * npy + yi);
  double2 h[5][NORB];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int o = 0; o < NORB; ++o)
      h[k][o] = k < cnt ? Hx[i + (long long)(nb[k] * NORB + o) * n] : make_double2(0.0, 0.0);
  const int c0 = blockIdx.y * kStencilCols, c1 = min(R, c0 + kStencilCols);
  for (int c = c0; c < c1; ++c) {
    const double2* v = Vx + (long long)c * n;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if (k < cnt)
#pragma unroll
        for (int o = 0; o < NORB; ++o) acc = cadd(acc, cmul(h[k][o], v[nb[k] * NORB + o]));
    Px[i + (long long)c * n] = acc;
  }
}

// up to three batched copies Y_k = X_k in one pass
template <typename T>
__global__ void k_copy3(cx<T>* Y0, const cx<T>* X0, cx<T>* Y1, const cx<T>* X1, cx<T>* Y2,
                        const cx<T>* X2, long long stride, long long nn, const int* active) {
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  const long long o = (long long)b * stride;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    Y0[o + i] = X0[o + i];
    if (Y1) Y1[o + i] = X1[o + i];
    if (Y2) Y2[o + i] = X2[o + i];
  }
}

// Sancho-Rubio seed (Contact_Self_Energy.F90:71-76, :83-88, :141): for chain c
// (lead = c / nb, energy = c % nb)  alpha = (E + i eta) I - Hend,
// work = alpha + extra * i eta I.
template <typename T>
__global__ void k_sr_alpha(cx<T>* alpha, cx<T>* work, long long stride, const cx<T>* H0,
                           const cx<T>* H1, const double* energy, int nb, double eta, int extra,
                           int n) {
  const int c = blockIdx.y;
  const int lead = c / nb, e = c % nb;
  const cx<T>* H = lead ? H1 : H0;
  const long long nn = (long long)n * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i % n), col = (int)(i / n);
    cx<T> v = cneg(H[i]);
    if (r == col) {
      v.x += (T)energy[e];
      v.y += (T)eta;
    }
    alpha[(long long)c * stride + i] = v;
    if (r == col && extra) v.y += (T)eta;
    work[(long long)c * stride + i] = v;
  }
}

// out[c] = max_ij |X_ij| (NaN if any entry is NaN).  A (chunks x chains) grid:
// each workgroup reduces its chunk and folds the result into out[c] with a 64-bit
// unsigned atomic max on the double's bit pattern, which orders non-negative
// doubles as their values; NaN is stored as the positive quiet NaN, whose pattern
// exceeds +inf's.  k_maxabs_clear zeroes out[c] first (stream order).
__global__ void k_maxabs_clear(double* out, const int* active, int batch) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < batch) out[active ? active[t] : t] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_maxabs(const cx<T>* X, long long sX, long long nn,
                                                double* out, const int* active) {
  const int c = active ? active[blockIdx.y] : (int)blockIdx.y;
  const cx<T>* Xc = X + (long long)c * sX;
  double m = 0.0;
  bool nan = false;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn; i += (long long)gridDim.x * 256) {
    const double v = hypot((double)Xc[i].x, (double)Xc[i].y);
    if (v != v) nan = true;
    m = fmax(m, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    m = fmax(m, __shfl_xor(m, off));
    nan = nan || __shfl_xor((int)nan, off);
  }
  __shared__ double red[4];
  __shared__ int rn[4];
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = m;
    rn[threadIdx.x >> 6] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    const bool an = rn[0] | rn[1] | rn[2] | rn[3];
    const unsigned long long bits =
        an ? 0x7ff8000000000000ull : (unsigned long long)__double_as_longlong(t);
    atomicMax((unsigned long long*)(out + c), bits);
  }
}

// Gamma = i (Sigma - Sigma')  (Contact_Self_Energy.F90:100-101) and
// dos[c] = -sum_i Im Sigma_ii / pi  (:94-97); one workgroup per (column block, chain).
template <typename T>
__global__ void k_gamma(const cx<T>* Sig, long long sS, cx<T>* Gam, long long sG, int n) {
  const int c = blockIdx.z;
  const int col = blockIdx.y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const cx<T>* S = Sig + (long long)c * sS;
  const cx<T> a = S[(long long)col * n + r], bt = S[(long long)r * n + col];
  // i (a - conj(bt))
  const T dr = a.x - bt.x, di = a.y + bt.y;
  Gam[(long long)c * sG + (long long)col * n + r] = mk<T>(-di, dr);
}

// ---------------------------------------------------------------------------
// Pivoted Cholesky factor Gam = W W' of a Hermitian positive semidefinite matrix
// (the contact broadening), one launch per column, with implicit pivoting: W keeps
// the original row order and column j belongs to pivot p_j, the unused row of
// largest remaining diagonal d_i = Gam_ii - sum_{l<j} |W_il|^2 (ties to the lower
// row):
//     W_pj = sqrt(d_p),  W_ij = (Gam_ip - sum_{l<j} W_il conj(W_pl)) / W_pj  (i unused),
//     W_ij = 0 for rows already used.
// Blocked by panels of kPcholNB columns: after each panel Gam -= W_panel W_panel'
// (one MFMA product, in place), so a column sums only over its own panel.
// Once the largest remaining d_p is at most kPcholTol * max_i Gam_ii the remaining
// columns are zero (the dropped part is positive semidefinite with trace below
// n * kPcholTol * max Gam_ii).  Before each column the remaining trace
// sum_{i unused} d_i -- the trace of the part the columns so far leave out, which
// is positive semidefinite -- is formed (per workgroup, then over the workgroups
// in a fixed order, so the same bits on every run); the first column count at
// which it is at most a.tau * tr(Gam) is kept as the truncation rank.
// blockIdx.y selects the lead, blockIdx.z the matrix.
// Per-matrix state (bytes from st + z * sSt, pchol_state_bytes): d[n] doubles,
// used[n] ints, then three 64-bit pivot keys (key of column j in slot j % 3), the
// stop threshold, tr(Gam), the number of nonzero columns, the truncation rank
// (-1 until reached), and the remaining-trace partial sums of the last two columns
// (slot j % 2 holds those before column j).
constexpr double kPcholTol = 1e-15;
constexpr int kPcholNB = 32;

__device__ inline unsigned long long pchol_key(double d, int i) {
  const double c = d > 0.0 ? d : 0.0;  // rounding below zero counts as zero; NaN too
  return ((unsigned long long)__double_as_longlong(c) & ~0xfffull) | (unsigned long long)(4095 - i);
}

// Sum over a 256-thread workgroup, the same bits on every run; valid in thread 0.
__device__ inline double block_sum256(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __shared__ double red[4];
  __syncthreads();  // red may still be read by an earlier call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

struct PcholState {
  double* d;
  int* used;
  unsigned long long* key;
  double* thr;
  double* tr0;
  int* rank;
  int* rtau;
  double* part;  // [2][kPcholMaxBlocks]
};
__device__ inline PcholState pchol_state(char* st, int n) {
  PcholState p;
  p.d = (double*)st;
  p.used = (int*)(st + 8LL * n);
  p.key = (unsigned long long*)(st + pchol_key_off(n));
  p.thr = (double*)(p.key + 3);
  p.tr0 = p.thr + 1;
  p.rank = (int*)(st + pchol_rank_off(n));
  p.rtau = p.rank + 1;
  p.part = (double*)(st + pchol_rank_off(n) + 8);
  return p;
}

__global__ void k_pchol_clear(PcholArgs a) {
  const PcholState q = pchol_state(a.st[blockIdx.y] + (long long)blockIdx.z * a.sSt, a.n);
  if (threadIdx.x < 3) q.key[threadIdx.x] = 0ull;
  if (threadIdx.x == 3) *q.thr = 0.0;
  if (threadIdx.x == 4) *q.tr0 = 0.0;
  if (threadIdx.x == 5) *q.rank = 0;
  if (threadIdx.x == 6) *q.rtau = -1;
}

__device__ inline unsigned long long wave_umax64(unsigned long long x) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long y = __shfl_xor(x, off);
    x = y > x ? y : x;
  }
  return x;
}

__global__ __launch_bounds__(256) void k_pchol_init(PcholArgs a) {
  const int lead = blockIdx.y;
  const double2* G = a.gam[lead] + (long long)blockIdx.z * a.sG;
  const PcholState q = pchol_state(a.st[lead] + (long long)blockIdx.z * a.sSt, a.n);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long k = 0ull;
  double di = 0.0;
  if (i < a.n) {
    di = G[(long long)i * a.n + i].x;
    q.d[i] = di;
    q.used[i] = 0;
    k = pchol_key(di, i);
  }
  const double part = block_sum256(di > 0.0 ? di : 0.0);
  if (threadIdx.x == 0) q.part[blockIdx.x] = part;  // slot 0: before column 0
  k = wave_umax64(k);
  if ((threadIdx.x & 63) == 0 && k) atomicMax(&q.key[0], k);
}

__global__ __launch_bounds__(256) void k_pchol_col(PcholArgs a, int j, int j0) {
  const int lead = blockIdx.y, n = a.n;
  const double2* G = a.gam[lead] + (long long)blockIdx.z * a.sG;
  double2* W = a.w[lead] + (long long)blockIdx.z * a.sW;
  const PcholState q = pchol_state(a.st[lead] + (long long)blockIdx.z * a.sSt, n);
  const unsigned long long kj = q.key[j % 3];
  // slot (j + 2) % 3 held the key of column j - 1, read by the previous launch; it
  // takes column j + 2's key, written by the next launch
  if (blockIdx.x == 0 && threadIdx.x == 0) q.key[(j + 2) % 3] = 0ull;
  // the trace left out before this column, added over the workgroups in order
  const double* pin = q.part + (j & 1) * kPcholMaxBlocks;
  double rem = 0.0;
  for (int w = 0; w < (int)gridDim.x; ++w) rem += pin[w];
  const double tr0 = j == 0 ? rem : *q.tr0;
  const int p = 4095 - (int)(kj & 0xfffull);
  const double dp = kj ? q.d[p] : 0.0;
  const double t = j == 0 ? dp * kPcholTol : *q.thr;
  if (j == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    *q.thr = t;
    *q.tr0 = tr0;
  }
  const bool stop = kj == 0ull || !(dp > t);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (!stop) *q.rank = j + 1;
    if (*q.rtau < 0 && a.tau > 0.0 && rem <= a.tau * tr0) *q.rtau = j;
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long k = 0ull;
  double left = 0.0;  // this row's remaining diagonal after the column (unused rows)
  if (i < n) {
    double2* Wj = W + (long long)j * n;
    if (stop || (q.used[i] && i != p)) {
      Wj[i] = make_double2(0.0, 0.0);
    } else if (i == p) {
      Wj[i] = make_double2(sqrt(dp), 0.0);
      q.used[i] = 1;
    } else {
      double2 acc = G[(long long)p * n + i];  // Gam minus the earlier panels' part
      for (int l = j0; l < j; ++l) {
        const double2 wi = W[(long long)l * n + i], wp = W[(long long)l * n + p];
        // acc -= wi * conj(wp)
        acc.x -= wi.x * wp.x + wi.y * wp.y;
        acc.y -= wi.y * wp.x - wi.x * wp.y;
      }
      const double s = sqrt(dp);
      const double2 w = make_double2(acc.x / s, acc.y / s);
      Wj[i] = w;
      const double di = q.d[i] - (w.x * w.x + w.y * w.y);
      q.d[i] = di;
      k = pchol_key(di, i);
      left = di > 0.0 ? di : 0.0;
    }
  }
  const double part = block_sum256(left);
  if (threadIdx.x == 0) q.part[((j + 1) & 1) * kPcholMaxBlocks + blockIdx.x] = part;
  k = wave_umax64(k);
  if ((threadIdx.x & 63) == 0 && k) atomicMax(&q.key[(j + 1) % 3], k);
}

// truncation rank = the rank where tau was never reached (or tau <= 0)
__global__ void k_pchol_fin(PcholArgs a) {
  const PcholState q = pchol_state(a.st[blockIdx.y] + (long long)blockIdx.z * a.sSt, a.n);
  if (threadIdx.x == 0) {
    const int r = *q.rank, t = *q.rtau;
    *q.rtau = (t < 0 || t > r) ? r : t;
  }
}

hipError_t launch_pchol(hipStream_t s, const PcholArgs& a, int batch) {
  if (batch <= 0 || a.n <= 0) return hipSuccess;
  if (a.n > 4095) return hipErrorInvalidValue;  // 12-bit row index in the pivot keys
  const dim3 grid((a.n + 255) / 256, 2, batch);
  hipLaunchKernelGGL(k_pchol_clear, dim3(1, 2, batch), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_pchol_init, grid, dim3(256), 0, s, a);
  for (int j0 = 0; j0 < a.n; j0 += kPcholNB) {
    const int j1 = min(a.n, j0 + kPcholNB);
    for (int j = j0; j < j1; ++j) hipLaunchKernelGGL(k_pchol_col, grid, dim3(256), 0, s, a, j, j0);
    if (j1 == a.n) break;
    for (int lead = 0; lead < 2; ++lead) {  // Gam -= W(:, j0:j1) W(:, j0:j1)'
      GemmBatch g = {};
      g.A = a.w[lead] + (long long)j0 * a.n; g.sA = a.sW; g.lda = a.n;
      g.B = a.w[lead] + (long long)j0 * a.n; g.sB = a.sW; g.ldb = a.n;
      g.C = a.gam[lead]; g.sC = a.sG; g.ldc = a.n;
      g.m = a.n; g.n = a.n; g.k = j1 - j0;
      g.m3 = 0;  // 4-product form: componentwise accurate, and timed apart from the 3M products
      const hipError_t e = gemm_batched<double>(s, OP_N, OP_C, g, batch, mk<double>(-1, 0), mk<double>(1, 0));
      if (e != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL(k_pchol_fin, dim3(1, 2, batch), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void k_dos(const cx<T>* Sig, long long sS, int n, double* out,
                                             long long sOut) {
  const int c = blockIdx.x;
  const cx<T>* S = Sig + (long long)c * sS;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += (double)S[(long long)i * n + i].y;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    out[(long long)c * sOut] = -(((red[0] + red[1]) + red[2]) + red[3]) / 3.14159265358979323846;
}

// precision conversion (mixed-precision form)
template <typename Td, typename Ts>
__global__ void k_convert(cx<Td>* Y, long long sY, const cx<Ts>* X, long long sX, long long nn,
                          const int* active) {
  const int b = active ? active[blockIdx.y] : (int)blockIdx.y;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    const cx<Ts> v = X[(long long)b * sX + i];
    Y[(long long)b * sY + i] = mk<Td>((Td)v.x, (Td)v.y);
  }
}

// ---------------------------------------------------------------------------
// Iterative refinement of the block-tridiagonal solve A V = B (A = zI - H - Sigma
// with the slice couplings -X, -X'; Recursive_Greens.F90:6) against complex64
// block-LU factors: the residual of every slice of an R-column iterate in one
// launch.  Element (i, j) of slice x of energy b:
//   T = c64(B_x - z V_x + P_x + X V_{x+1} + X' V_{x-1})
// in FP64, with P_x = (H_x + Sigma_x) V_x from the FP64 GEMM (plus U_x V_x for a
// compact Hamiltonian's site potentials U_x).  B holds the two block columns' right
// sides: Wr (columns j < RB) on the last slice, Wl (columns RB.. as Wl's 0..) on
// the first.  V, P, T share the slice stride sS and have energy strides sV, sP, sT.
// X is site-block-diagonal (NORB x NORB blocks).
template <int NORB>
__global__ void k_lu32_resid(cx<float>* T, long long sT, const double2* V, long long sV,
                             const double2* P, long long sP, long long sS, const double* U,
                             long long sU, const double2* X, const double2* Wr, const double2* Wl,
                             long long sW, const double* energy, double eta, int n, int RB, int L) {
  const int b = blockIdx.z / L, x = blockIdx.z - b * L, j = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long e = (long long)j * n + i;
  const int base = (i / NORB) * NORB;
  const double2* Vx = V + (long long)b * sV + (long long)x * sS;
  const double2 v0 = Vx[e];
  const double2 z = mk<double>(energy[b], eta);
  double2 r = cadd(P[(long long)b * sP + (long long)x * sS + e], cneg(cmul(z, v0)));
  if (U) {
    const double u = U[(long long)x * sU + i / NORB];
    r.x += u * v0.x;
    r.y += u * v0.y;
  }
  if (x == L - 1 && j < RB) r = cadd(r, Wr[(long long)b * sW + e]);
  if (x == 0 && j >= RB) r = cadd(r, Wl[(long long)b * sW + (long long)(j - RB) * n + i]);
  const long long col = (long long)j * n + base;
#pragma unroll
  for (int k = 0; k < NORB; ++k) {
    if (x < L - 1) r = cadd(r, cmul(X[(long long)(base + k) * n + i], Vx[sS + col + k]));
    if (x > 0) r = cadd(r, cmul(cconj(X[(long long)i * n + base + k]), Vx[-sS + col + k]));
  }
  T[(long long)b * sT + (long long)x * sS + e] = mk<float>((float)r.x, (float)r.y);
}

template <int NORB>
static hipError_t lu32_resid_n(hipStream_t s, cx<float>* T, long long sT, const double2* V, long long sV,
                               const double2* P, long long sP, long long sS, const double* U,
                               long long sU, const double2* X, const double2* Wr, const double2* Wl,
                               long long sW, const double* energy, double eta, int n, int R, int RB,
                               int L, int batch) {
  hipLaunchKernelGGL((k_lu32_resid<NORB>), dim3((n + 255) / 256, R, batch * L), dim3(256), 0, s, T, sT, V,
                     sV, P, sP, sS, U, sU, X, Wr, Wl, sW, energy, eta, n, RB, L);
  return hipGetLastError();
}

hipError_t launch_lu32_resid(hipStream_t s, int norb, cx<float>* T, long long sT, const double2* V,
                             long long sV, const double2* P, long long sP, long long sS, const double* U,
                             long long sU, const double2* X, const double2* Wr, const double2* Wl,
                             long long sW, const double* energy, double eta, int n, int R, int RB, int L,
                             int batch) {
  if (batch <= 0 || R <= 0) return hipSuccess;
  if ((long long)batch * L > 65535) return hipErrorInvalidValue;
  if (norb == 1)
    return lu32_resid_n<1>(s, T, sT, V, sV, P, sP, sS, U, sU, X, Wr, Wl, sW, energy, eta, n, R, RB, L, batch);
  if (norb == 2)
    return lu32_resid_n<2>(s, T, sT, V, sV, P, sP, sS, U, sU, X, Wr, Wl, sW, energy, eta, n, R, RB, L, batch);
  if (norb == 4)
    return lu32_resid_n<4>(s, T, sT, V, sV, P, sP, sS, U, sU, X, Wr, Wl, sW, energy, eta, n, R, RB, L, batch);
  return hipErrorInvalidValue;
}

// Y += D (FP64 accumulate of a complex64 correction), nn elements per matrix.
__global__ void k_add_c64(double2* Y, long long sY, const cx<float>* D, long long sD, long long nn) {
  const int b = blockIdx.y;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    const cx<float> d = D[(long long)b * sD + i];
    double2& y = Y[(long long)b * sY + i];
    y.x += (double)d.x;
    y.y += (double)d.y;
  }
}

hipError_t launch_add_c64(hipStream_t s, double2* Y, long long sY, const cx<float>* D, long long sD,
                          long long nn, int batch) {
  if (batch <= 0 || nn <= 0) return hipSuccess;
  long long blocks = (nn + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_add_c64, dim3((unsigned)blocks, batch), dim3(256), 0, s, Y, sY, D, sD, nn);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stage-wise observables on materialised Gn/Gp blocks (Calculate_Observables).

// out[b*n + i] = Re M_b[i][i]
template <typename T>
__global__ void k_diag_re(const cx<T>* M, long long sM, int n, double* out) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[(long long)b * n + i] = (double)M[(long long)b * sM + (long long)i * n + i].x;
}

// out = Re sum_ij P_ij Q_ji  (= Re tr(P Q)), one workgroup
template <typename T>
__global__ __launch_bounds__(1024) void k_trace_prod(const cx<T>* P, const cx<T>* Q, int n,
                                                     double* out) {
  double acc = 0.0;
  const long long nn = (long long)n * n;
  for (long long e = threadIdx.x; e < nn; e += 1024) {
    const int i = (int)(e % n), j = (int)(e / n);
    const cx<T> p = P[e], q = Q[(long long)i * n + j];
    acc += (double)p.x * q.x - (double)p.y * q.y;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  __shared__ double red[16];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += red[w];
    *out = t;
  }
}

// Per (slice x, site s = y + npy*z): spin Re tr(Gn_site sigma_a) (Calculate_Observables.F90:10-30)
// and the bond-current terms of :111-168 before the e^2/hbar and dE/2pi factors:
//   y: Re i [H(a,b)Gn(b,a) - Gn(a,b)H(b,a) - H(a,b)Gp(b,a) + Gp(a,b)H(b,a)], b on site y+1
//   z: the same with b on site z+1;  x: HopX with Gnl/Gnu/Gpl/Gpu of slice x.
// out[(q * L + x) * nsites + s], q = 0..2 spin, 3 Jx, 4 Jy, 5 Jz.
// sH: element stride between the slices of H (n*n for Htotd, 0 when every slice's
// off-diagonal part is one template -- only off-diagonal elements of H are read)
template <typename T, int NORB>
__global__ __launch_bounds__(64) void k_site_obs(const cx<T>* Gnd, const cx<T>* Gpd, const cx<T>* Gnu, const cx<T>* Gnl,
                           const cx<T>* Gpu, const cx<T>* Gpl, const cx<T>* H, long long sH,
                           const cx<T>* X, int n, int L, int npy, int npz, SpinMats sp, double* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = blockIdx.y;
  const int ns = npy * npz;
  if (s >= ns) return;
  const int y = s % npy, z = s / npy;
  const long long nn = (long long)n * n;
  const long long bx = (long long)x * nn;
  const int idx = NORB * s;  // (y-1)*Norb + (z-1)*NN, 0-based
  auto at = [&](const cx<T>* M, long long base, int i, int j) { return M[base + (long long)j * n + i]; };
  double spin[3] = {0, 0, 0};
  for (int ii = 0; ii < NORB; ++ii)
    for (int jj = 0; jj < NORB; ++jj) {
      const cx<T> g = at(Gnd, bx, idx + ii, idx + jj);
      for (int m = 0; m < 3; ++m)
        spin[m] += (double)g.x * sp.re[m][jj][ii] - (double)g.y * sp.im[m][jj][ii];
    }
  // Re(i*c) = -Im(c)
  auto jterm = [&](int b0) {
    double acc = 0.0;
    for (int ii = 0; ii < NORB; ++ii)
      for (int jj = 0; jj < NORB; ++jj) {
        const int a = idx + ii, b = b0 + jj;
        const cx<T> hab = at(H, x * sH, a, b), hba = at(H, x * sH, b, a);
        cx<T> c = cmul(hab, at(Gnd, bx, b, a));
        c = csub(c, cmul(at(Gnd, bx, a, b), hba));
        c = csub(c, cmul(hab, at(Gpd, bx, b, a)));
        c = cadd(c, cmul(at(Gpd, bx, a, b), hba));
        acc += -(double)c.y;
      }
    return acc;
  };
  double jx = 0.0, jy = 0.0, jz = 0.0;
  if (y < npy - 1) jy = jterm(idx + NORB);
  if (z < npz - 1) jz = jterm(idx + NORB * npy);
  if (x < L - 1) {
    for (int ii = 0; ii < NORB; ++ii)
      for (int jj = 0; jj < NORB; ++jj) {
        const int a = idx + ii, b = idx + jj;
        const cx<T> hx = X[(long long)b * n + a];
        cx<T> c = cmul(hx, at(Gnl, bx, b, a));
        c = csub(c, cmul(at(Gnu, bx, a, b), cconj(hx)));
        c = csub(c, cmul(hx, at(Gpl, bx, b, a)));
        c = cadd(c, cmul(at(Gpu, bx, a, b), cconj(hx)));
        jx += -(double)c.y;
      }
  }
  const long long q = (long long)L * ns, o = (long long)x * ns + s;
  out[o] = spin[0];
  out[q + o] = spin[1];
  out[2 * q + o] = spin[2];
  out[3 * q + o] = jx;
  out[4 * q + o] = jy;
  out[5 * q + o] = jz;
}

// Y = a P + b Q (real a, b) over nn elements
template <typename T>
__global__ void k_lincomb(cx<T>* Y, const cx<T>* P, const cx<T>* Q, double a, double b,
                          long long nn) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    const cx<T> p = P[i], q = Q[i];
    Y[i] = mk<T>((T)(a * p.x + b * q.x), (T)(a * p.y + b * q.y));
  }
}

template <typename T>
hipError_t launch_lincomb(hipStream_t s, cx<T>* Y, const cx<T>* P, const cx<T>* Q, double a,
                          double b, long long nn) {
  if (nn <= 0) return hipSuccess;
  const long long blocks = (nn + 255) / 256 < 2048 ? (nn + 255) / 256 : 2048;
  hipLaunchKernelGGL((k_lincomb<T>), dim3((unsigned)blocks), dim3(256), 0, s, Y, P, Q, a, b, nn);
  return hipGetLastError();
}
template hipError_t launch_lincomb<double>(hipStream_t, cx<double>*, const cx<double>*,
                                           const cx<double>*, double, double, long long);

// Y_b += I
template <typename T>
__global__ void k_add_eye(cx<T>* Y, long long sY, int n) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Y[(long long)b * sY + (long long)i * n + i].x += (T)1;
}

template <typename T>
hipError_t launch_diag_re(hipStream_t s, const cx<T>* M, long long sM, int n, int batch,
                          double* out) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_diag_re<T>), dim3((n + 255) / 256, batch), dim3(256), 0, s, M, sM, n, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_trace_prod(hipStream_t s, const cx<T>* P, const cx<T>* Q, int n, double* out) {
  hipLaunchKernelGGL((k_trace_prod<T>), dim3(1), dim3(1024), 0, s, P, Q, n, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_site_obs(hipStream_t s, int norb, const cx<T>* Gnd, const cx<T>* Gpd,
                           const cx<T>* Gnu, const cx<T>* Gnl, const cx<T>* Gpu, const cx<T>* Gpl,
                           const cx<T>* H, long long sH, const cx<T>* X, int n, int L, int npy,
                           int npz, const SpinMats& sp, double* out) {
  dim3 grid((npy * npz + 63) / 64, L);
#define NEGF_SO(NO)                                                                              \
  if (norb == NO) {                                                                              \
    hipLaunchKernelGGL((k_site_obs<T, NO>), grid, dim3(64), 0, s, Gnd, Gpd, Gnu, Gnl, Gpu, Gpl, H, \
                       sH, X, n, L, npy, npz, sp, out);                                          \
    return hipGetLastError();                                                                    \
  }
  NEGF_SO(1) NEGF_SO(2) NEGF_SO(4)
#undef NEGF_SO
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_add_eye(hipStream_t s, cx<T>* Y, long long sY, int n, int batch) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_add_eye<T>), dim3((n + 255) / 256, batch), dim3(256), 0, s, Y, sY, n);
  return hipGetLastError();
}

template hipError_t launch_diag_re<double>(hipStream_t, const cx<double>*, long long, int, int, double*);
template hipError_t launch_trace_prod<double>(hipStream_t, const cx<double>*, const cx<double>*, int,
                                              double*);
template hipError_t launch_site_obs<double>(hipStream_t, int, const cx<double>*, const cx<double>*,
                                            const cx<double>*, const cx<double>*, const cx<double>*,
                                            const cx<double>*, const cx<double>*, long long,
                                            const cx<double>*, int, int, int, int, const SpinMats&,
                                            double*);
template hipError_t launch_add_eye<double>(hipStream_t, cx<double>*, long long, int, int);

// ---------------------------------------------------------------------------
// Host launchers
static inline dim3 grid_nn(long long nn, int batch) {
  long long blocks = (nn + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  return dim3((unsigned)blocks, batch);
}

template <typename T>
hipError_t launch_assemble(hipStream_t s, int norb, int mode, cx<T>* M, long long sM,
                           const cx<T>* H, int add_diag, double hsign, const double* energy,
                           double eta, const cx<T>* S1, long long sS1, const cx<T>* S2,
                           long long sS2, const cx<T>* g, long long sg, const cx<T>* X, int n,
                           int batch, const int* active, const double* V) {
  if (batch <= 0) return hipSuccess;
  dim3 grid((n + 255) / 256, n, batch);
#define NEGF_ASM(NO, MO)                                                                    \
  if (norb == NO && mode == MO) {                                                           \
    hipLaunchKernelGGL((k_assemble<T, NO, MO>), grid, dim3(256), 0, s, M, sM, H, add_diag,   \
                       (T)hsign, energy, eta, S1, sS1, S2, sS2, g, sg, X, n, active, V);     \
    return hipGetLastError();                                                               \
  }
#define NEGF_ASM_N(NO) NEGF_ASM(NO, 0) NEGF_ASM(NO, 1) NEGF_ASM(NO, 2) NEGF_ASM(NO, 3)
  NEGF_ASM_N(1) NEGF_ASM_N(2) NEGF_ASM_N(4)
#undef NEGF_ASM_N
#undef NEGF_ASM
  return hipErrorInvalidValue;
}

// k_assemble_blk for a batch of `batch` slices (g complex128 or, g32 given, complex64).
hipError_t launch_assemble_blk(hipStream_t s, int norb, double2* M, long long sM, const double2* H, int add_diag,
                               double hsign, const double* energy, double eta, const double2* S1, long long sS1,
                               const double2* S2, long long sS2, const double2* g, const float2* g32, long long sg,
                               const double2* X, int n, int batch, const double* V) {
  if (batch <= 0) return hipSuccess;
  if (norb < 1 || norb > 4 || n % norb) return hipErrorInvalidValue;
  const int nsite = n / norb;
  dim3 grid((nsite + 127) / 128, nsite, batch);
#define NEGF_ASB(NO)                                                                                  \
  if (norb == NO) {                                                                                   \
    if (g32)                                                                                          \
      hipLaunchKernelGGL((k_assemble_blk<NO, float2>), grid, dim3(128), 0, s, M, sM, H, add_diag, hsign, \
                         energy, eta, S1, sS1, S2, sS2, g32, sg, X, n, nsite, V);                     \
    else                                                                                              \
      hipLaunchKernelGGL((k_assemble_blk<NO, double2>), grid, dim3(128), 0, s, M, sM, H, add_diag,     \
                         hsign, energy, eta, S1, sS1, S2, sS2, g, sg, X, n, nsite, V);                \
    return hipGetLastError();                                                                         \
  }
  NEGF_ASB(1) NEGF_ASB(2) NEGF_ASB(3) NEGF_ASB(4)
#undef NEGF_ASB
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_bd_mul(hipStream_t s, int norb, int conj, cx<T>* Y, long long sY,
                         const cx<T>* G, long long sG, const cx<T>* X, int n, int cols, int batch,
                         const int* active, int accum) {
  if (batch <= 0 || cols <= 0) return hipSuccess;
  dim3 grid((n + 255) / 256, cols, batch);  // the first `cols` columns of G
#define NEGF_BD(NO, CJ)                                                                       \
  if (norb == NO && conj == CJ) {                                                             \
    hipLaunchKernelGGL((k_bd_mul<T, NO, CJ>), grid, dim3(256), 0, s, Y, sY, G, sG, X, n, active, accum); \
    return hipGetLastError();                                                                 \
  }
  NEGF_BD(1, 0) NEGF_BD(1, 1) NEGF_BD(2, 0) NEGF_BD(2, 1) NEGF_BD(4, 0) NEGF_BD(4, 1)
#undef NEGF_BD
  return hipErrorInvalidValue;
}

// Zero `words` 32-bit words (a kernel rather than hipMemsetAsync, so that the
// clear can sit inside a captured HIP graph).
__global__ void k_zero32(unsigned* p, long long words) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

hipError_t launch_zero(hipStream_t s, void* p, size_t bytes) {
  const long long words = (long long)(bytes / 4);
  if (words <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>(1024, (words + 255) / 256);
  hipLaunchKernelGGL(k_zero32, dim3(blocks), dim3(256), 0, s, (unsigned*)p, words);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bondblocks(hipStream_t s, int norb, const cx<T>* V, long long sV, const cx<T>* Vn,
                             long long sVn, const double2* H, const double2* X, int n, int npy,
                             int npz, double* out, long long sOut, long long qstride, int pairs,
                             const int* kc, long long skc, int batch, const int* active, int kfirst,
                             int klast, int accum) {
  if (batch <= 0) return hipSuccess;
  dim3 grid((n + 15) / 16, 1, batch);
#define NEGF_BB(NO)                                                                                \
  if (norb == NO) {                                                                                \
    hipLaunchKernelGGL((k_bondblocks<T, NO>), grid, dim3(256), 0, s, V, sV, Vn, sVn, H, X, n, npy, npz, \
                       out, sOut, qstride, pairs, kc, skc, active, kfirst, klast, accum);          \
    return hipGetLastError();                                                                      \
  }
  NEGF_BB(1) NEGF_BB(2) NEGF_BB(4)
#undef NEGF_BB
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_siteblocks(hipStream_t s, int norb, const cx<T>* Bm, long long sB,
                             const cx<T>* G, long long sG, int n, double* out, long long sOut,
                             long long qstride, const SpinMats& sp, const int* kc, long long skc,
                             int batch, const int* active, int kfirst, int klast, int accum) {
  if (batch <= 0) return hipSuccess;
  dim3 grid((n + 15) / 16, 1, batch);
#define NEGF_SB(NO)                                                                            \
  if (norb == NO) {                                                                            \
    hipLaunchKernelGGL((k_siteblocks<T, NO>), grid, dim3(256), 0, s, Bm, sB, G, sG, n, out, sOut, \
                       qstride, sp, kc, skc, active, kfirst, klast, accum);                    \
    return hipGetLastError();                                                                  \
  }
  NEGF_SB(1) NEGF_SB(2) NEGF_SB(4)
#undef NEGF_SB
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_inner(hipStream_t s, const cx<T>* P, long long sP, const cx<T>* Q, long long sQ,
                        int ld, int m, int ncols, double* out, long long sOut, int batch,
                        const int* active, int accum) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_inner<T>), dim3(batch), dim3(1024), 0, s, P, sP, Q, sQ, ld, m, ncols, out,
                     sOut, active, accum);
  return hipGetLastError();
}

// Y_b(:, kc_b : cols) = 0 (kc_b = kc[b * skc]): a factor product truncated per matrix.
template <typename T>
__global__ void k_trunc_cols(cx<T>* Y, long long sY, int n, int cols, const int* kc, long long skc) {
  const int b = blockIdx.z;
  const int col = blockIdx.y;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n || col >= cols || col < kc[(long long)b * skc]) return;
  Y[(long long)b * sY + (long long)col * n + a] = mk<T>(0, 0);
}

template <typename T>
hipError_t launch_trunc_cols(hipStream_t s, cx<T>* Y, long long sY, int n, int cols, const int* kc,
                             long long skc, int batch) {
  if (batch <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_trunc_cols<T>), dim3((n + 255) / 256, cols, batch), dim3(256), 0, s, Y, sY, n,
                     cols, kc, skc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_eye(hipStream_t s, cx<T>* Y, long long sY, int n, int batch, const int* active) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_eye<T>), grid_nn((long long)n * n, batch), dim3(256), 0, s, Y, sY, n, active);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_add(hipStream_t s, cx<T>* Y, long long sY, const cx<T>* X, long long sX,
                      const cx<T>* Z, long long sZ, long long nn, int batch, const int* active) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_add<T>), grid_nn(nn, batch), dim3(256), 0, s, Y, sY, X, sX, Z, sZ, nn, active);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy3(hipStream_t s, cx<T>* Y0, const cx<T>* X0, cx<T>* Y1, const cx<T>* X1,
                        cx<T>* Y2, const cx<T>* X2, long long stride, long long nn, int batch,
                        const int* active) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_copy3<T>), grid_nn(nn, batch), dim3(256), 0, s, Y0, X0, Y1, X1, Y2, X2,
                     stride, nn, active);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sr_alpha(hipStream_t s, cx<T>* alpha, cx<T>* work, long long stride,
                           const cx<T>* H0, const cx<T>* H1, const double* energy, int nb,
                           double eta, int extra, int n) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_sr_alpha<T>), grid_nn((long long)n * n, 2 * nb), dim3(256), 0, s, alpha,
                     work, stride, H0, H1, energy, nb, eta, extra, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_maxabs(hipStream_t s, const cx<T>* X, long long sX, long long nn, double* out,
                         int batch, const int* active) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_maxabs_clear, dim3((batch + 63) / 64), dim3(64), 0, s, out, active, batch);
  const long long chunks = std::min<long long>(512, std::max<long long>(1, (nn + 4095) / 4096));
  hipLaunchKernelGGL((k_maxabs<T>), dim3((unsigned)chunks, batch), dim3(256), 0, s, X, sX, nn, out,
                     active);
  return hipGetLastError();
}

hipError_t launch_stencil_check(hipStream_t s, const double2* H, long long sH, int nslices, int n, int npy,
                                int npz, int norb, int* flag) {
  if (nslices <= 0 || n <= 0) return hipSuccess;
  if (npy < 1 || npz < 1 || norb < 1 || npy * npz * norb != n) return hipErrorInvalidValue;
  const int blocks = (int)std::min<long long>(2048, ((long long)n * n + 255) / 256);
  hipLaunchKernelGGL(k_stencil_check, dim3(blocks, nslices), dim3(256), 0, s, H, sH, n, npy, npz, norb, flag);
  return hipGetLastError();
}

hipError_t launch_stencil_mul(hipStream_t s, int norb, const double2* H, long long sH, const double2* V,
                              long long sVb, long long sVx, double2* P, long long sPb, long long sPx, int n, int R,
                              int L, int batch, int npy, int npz) {
  if (batch <= 0 || L <= 0 || R <= 0) return hipSuccess;
  if (npy * npz * norb != n) return hipErrorInvalidValue;
  const dim3 grid((n + 255) / 256, (R + kStencilCols - 1) / kStencilCols, batch * L);
#define NEGF_ST(NO)                                                                                        \
  if (norb == NO) {                                                                                        \
    hipLaunchKernelGGL((k_stencil_mul<NO>), grid, dim3(256), 0, s, H, sH, V, sVb, sVx, P, sPb, sPx, n, R, L, \
                       npy, npz);                                                                          \
    return hipGetLastError();                                                                              \
  }
  NEGF_ST(1) NEGF_ST(2) NEGF_ST(4)
#undef NEGF_ST
  return hipErrorInvalidValue;
}

hipError_t launch_sym_expand(hipStream_t s, const double2* M, long long sM, long long off1, const int2* pair,
                             int d, double2* Y, long long sY, int n, int batch, const SymMap& map) {
  if (batch <= 0) return hipSuccess;
  if (map.axis < 1 || map.axis > 2 || map.norb < 1 || map.norb > kSymMaxOrb || 2 * d != n) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sym_expand, grid_nn((long long)d * d, batch), dim3(256), 0, s, M, sM, off1, pair, d, Y, sY,
                     n, map);
  return hipGetLastError();
}

hipError_t launch_sym_maxabs(hipStream_t s, const double2* M, long long sM, long long off1, const int2* pair,
                             int d, double* out, int batch, const int* active) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_maxabs_clear, dim3((batch + 63) / 64), dim3(64), 0, s, out, active, batch);
  const long long dd = (long long)d * d;
  const long long chunks = std::min<long long>(512, std::max<long long>(1, (dd + 4095) / 4096));
  hipLaunchKernelGGL(k_sym_maxabs, dim3((unsigned)chunks, batch), dim3(256), 0, s, M, sM, off1, pair, d, out,
                     active);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gamma_dos(hipStream_t s, const cx<T>* Sig, long long sS, cx<T>* Gam,
                            long long sG, int n, int batch, double* dos, long long sDos) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_gamma<T>), dim3((n + 255) / 256, n, batch), dim3(256), 0, s, Sig, sS, Gam,
                     sG, n);
  hipLaunchKernelGGL((k_dos<T>), dim3(batch), dim3(256), 0, s, Sig, sS, n, dos, sDos);
  return hipGetLastError();
}

template <typename Td, typename Ts>
hipError_t launch_convert(hipStream_t s, cx<Td>* Y, long long sY, const cx<Ts>* X, long long sX,
                          long long nn, int batch, const int* active) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_convert<Td, Ts>), grid_nn(nn, batch), dim3(256), 0, s, Y, sY, X, sX, nn,
                     active);
  return hipGetLastError();
}

#define NEGF_INST(T)                                                                               \
  template hipError_t launch_assemble<T>(hipStream_t, int, int, cx<T>*, long long, const cx<T>*,   \
                                         int, double, const double*, double, const cx<T>*,         \
                                         long long, const cx<T>*, long long, const cx<T>*,         \
                                         long long, const cx<T>*, int, int, const int*,            \
                                         const double*);                                           \
  template hipError_t launch_bd_mul<T>(hipStream_t, int, int, cx<T>*, long long, const cx<T>*,     \
                                       long long, const cx<T>*, int, int, int, const int*, int);   \
  template hipError_t launch_siteblocks<T>(hipStream_t, int, const cx<T>*, long long,              \
                                           const cx<T>*, long long, int, double*, long long,       \
                                           long long, const SpinMats&, const int*, long long, int,  \
                                           const int*, int, int, int);                             \
  template hipError_t launch_bondblocks<T>(hipStream_t, int, const cx<T>*, long long, const cx<T>*, \
                                           long long, const double2*, const double2*, int, int,    \
                                           int, double*, long long, long long, int, const int*,    \
                                           long long, int, const int*, int, int, int);             \
  template hipError_t launch_inner<T>(hipStream_t, const cx<T>*, long long, const cx<T>*,          \
                                      long long, int, int, int, double*, long long, int,           \
                                      const int*, int);                                            \
  template hipError_t launch_trunc_cols<T>(hipStream_t, cx<T>*, long long, int, int, const int*,   \
                                           long long, int);                                        \
  template hipError_t launch_eye<T>(hipStream_t, cx<T>*, long long, int, int, const int*);         \
  template hipError_t launch_add<T>(hipStream_t, cx<T>*, long long, const cx<T>*, long long,       \
                                    const cx<T>*, long long, long long, int, const int*);          \
  template hipError_t launch_copy3<T>(hipStream_t, cx<T>*, const cx<T>*, cx<T>*, const cx<T>*,     \
                                      cx<T>*, const cx<T>*, long long, long long, int, const int*); \
  template hipError_t launch_sr_alpha<T>(hipStream_t, cx<T>*, cx<T>*, long long, const cx<T>*,     \
                                         const cx<T>*, const double*, int, double, int, int);      \
  template hipError_t launch_maxabs<T>(hipStream_t, const cx<T>*, long long, long long, double*,   \
                                       int, const int*);                                           \
  template hipError_t launch_gamma_dos<T>(hipStream_t, const cx<T>*, long long, cx<T>*, long long, \
                                          int, int, double*, long long);
NEGF_INST(double)
NEGF_INST(float)
#undef NEGF_INST
template hipError_t launch_convert<float, double>(hipStream_t, cx<float>*, long long,
                                                  const cx<double>*, long long, long long, int,
                                                  const int*);
template hipError_t launch_convert<double, float>(hipStream_t, cx<double>*, long long,
                                                  const cx<float>*, long long, long long, int,
                                                  const int*);

}  // namespace negf
