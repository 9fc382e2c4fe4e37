// Blocked triangular solves with the Cholesky factors of the matrix-core kernel
// (chol_mfma.hip).  That kernel factors by panels of 16 columns and leaves, for
// every diagonal block, the strictly lower part of the block's inverse in the
// (otherwise unused) upper triangle of the factor slot: element (J + c, J + r),
// r > c, holds (L_JJ^-1)[r][c]; the inverse's diagonal is 1 / L_rr.  A solve then
// takes three workgroup barriers per block of 16 columns instead of two per
// column, and every loop below has a fixed trip count, so its loads are issued
// together.
#pragma once
#include <hip/hip_runtime.h>

namespace lrb {

constexpr int SOLVE_NB = 16;   // diagonal block size of the factors the blocked solve reads
constexpr int SOLVE_LB = SOLVE_NB * (SOLVE_NB + 1);

// Inverse of the diagonal block at J (w <= 16 columns) into lb[r * 17 + c] (zero
// above the diagonal and outside the block).  NTH >= 256.
template <typename TL, int NTH>
__device__ __forceinline__ void stage_block_inverse(const TL* __restrict__ L, int n, int J, int w, double* lb) {
  const int t = threadIdx.x;
  if (t < SOLVE_NB * SOLVE_NB) {
    const int r = t / SOLVE_NB, c = t % SOLVE_NB;
    double v = 0.0;
    if (r < w && c <= r) {
      const TL e = L[(size_t)(J + r) * n + J + c];   // column J + r: L_rr on the diagonal, inverse entries above it
      v = r == c ? 1.0 / (double)e : (double)e;
    }
    lb[r * (SOLVE_NB + 1) + c] = v;
  }
}

// L L^T u = y in place for nf <= 8 columns y[f * ldy + i], i < n.  NTH threads
// (a multiple of 64, at least 256); zb: shared [16 * 8], lb: shared [16 * 17].
template <typename TL, int NTH>
__device__ void llt_solve_blocked(const TL* __restrict__ L, int n, double* y, int ldy, int nf, double* zb, double* lb) {
  static_assert(NTH >= SOLVE_NB * SOLVE_NB && NTH % 64 == 0, "solve block shape");
  constexpr int NB = SOLVE_NB, LD = SOLVE_NB + 1;
  const int tid = threadIdx.x;
  // forward: z_J = L_JJ^-1 (y_J - updates); rows below -= L[:, J block] z_J
  for (int J = 0; J < n; J += NB) {
    const int w = min(NB, n - J);
    stage_block_inverse<TL, NTH>(L, n, J, w, lb);
    __syncthreads();
    if (tid < NB * nf) {
      const int r = tid % NB, f = tid / NB;
      double z = 0.0;
#pragma unroll
      for (int c = 0; c < NB; c++) z += lb[r * LD + c] * (c < w ? y[(size_t)f * ldy + J + c] : 0.0);
      zb[r * 8 + f] = z;
    }
    __syncthreads();
    if (tid < NB * nf) {
      const int r = tid % NB, f = tid / NB;
      if (r < w) y[(size_t)f * ldy + J + r] = zb[r * 8 + f];
    }
    // rows below exist only under a full block (w = 16); a row's 16 factor entries are
    // loaded once for all fields, two rows per iteration, every load issued before
    // the arithmetic (each field's sum keeps its order)
    int i = J + w + tid;
    for (; i + NTH < n; i += 2 * NTH) {
      TL la[NB], lb2[NB];
#pragma unroll
      for (int c = 0; c < NB; c++) {
        la[c] = L[(size_t)(J + c) * n + i];
        lb2[c] = L[(size_t)(J + c) * n + i + NTH];
      }
#pragma unroll
      for (int f = 0; f < 8; f++) {
        if (f >= nf) break;
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int c = 0; c < NB; c++) {
          a += (double)la[c] * zb[c * 8 + f];
          b += (double)lb2[c] * zb[c * 8 + f];
        }
        y[(size_t)f * ldy + i] -= a;
        y[(size_t)f * ldy + i + NTH] -= b;
      }
    }
    if (i < n) {
      TL la[NB];
#pragma unroll
      for (int c = 0; c < NB; c++) la[c] = L[(size_t)(J + c) * n + i];
#pragma unroll
      for (int f = 0; f < 8; f++) {
        if (f >= nf) break;
        double a = 0.0;
#pragma unroll
        for (int c = 0; c < NB; c++) a += (double)la[c] * zb[c * 8 + f];
        y[(size_t)f * ldy + i] -= a;
      }
    }
    __syncthreads();
  }
  // backward: x_J = L_JJ^-T (z_J - L[below, J block]^T x_below).  The dots: thread
  // (c, g) sums column J + c over the rows J + 16 + g + 16 k, eight rows per
  // iteration with their loads issued together (a patch of thousands of members
  // otherwise waits on one load round trip per two rows), then the 16 lanes of one
  // c reduce.
  const int nblk = (n + NB - 1) / NB;
  const int c = tid / NB, g = tid % NB;
  for (int bi = nblk - 1; bi >= 0; bi--) {
    const int J = bi * NB, w = min(NB, n - J);
    stage_block_inverse<TL, NTH>(L, n, J, w, lb);
    if (c < NB) {
      const TL* Lc = L + (size_t)(J + c) * n;
      // one pass over the column for all fields (each field's sum keeps its order)
      double acc[8];
#pragma unroll
      for (int f = 0; f < 8; f++) acc[f] = 0.0;
      if (c < w) {
        int i = J + w + g;
        for (; i + 7 * NB < n; i += 8 * NB) {
          double l[8];
#pragma unroll
          for (int u = 0; u < 8; u++) l[u] = (double)Lc[i + u * NB];
#pragma unroll
          for (int f = 0; f < 8; f++) {
            if (f >= nf) break;
            const double* yf = y + (size_t)f * ldy + i;
            double s = 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++) s += l[u] * yf[u * NB];
            acc[f] += s;
          }
        }
        for (; i + NB < n; i += 2 * NB) {
          const double l0 = (double)Lc[i], l1 = (double)Lc[i + NB];
#pragma unroll
          for (int f = 0; f < 8; f++) {
            if (f >= nf) break;
            const double* yf = y + (size_t)f * ldy;
            acc[f] += l0 * yf[i] + l1 * yf[i + NB];
          }
        }
        if (i < n) {
          const double l0 = (double)Lc[i];
#pragma unroll
          for (int f = 0; f < 8; f++) {
            if (f >= nf) break;
            acc[f] += l0 * y[(size_t)f * ldy + i];
          }
        }
      }
#pragma unroll
      for (int f = 0; f < 8; f++) {
        if (f >= nf) break;
        double a = acc[f];
#pragma unroll
        for (int o = NB / 2; o > 0; o >>= 1) a += __shfl_xor(a, o);
        if (g == 0) zb[c * 8 + f] = c < w ? y[(size_t)f * ldy + J + c] - a : 0.0;
      }
    }
    __syncthreads();
    if (tid < NB * nf) {
      const int r = tid % NB, f = tid / NB;
      double x = 0.0;
#pragma unroll
      for (int c2 = 0; c2 < NB; c2++) x += lb[c2 * LD + r] * zb[c2 * 8 + f];   // (L_JJ^-T)[r][c2] = (L_JJ^-1)[c2][r]
      if (r < w) y[(size_t)f * ldy + J + r] = x;
    }
    __syncthreads();
  }
}

}  // namespace lrb
