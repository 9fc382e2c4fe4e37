// Lean blocked triangular solves L L^T y = b for one workgroup of LEAN_NT threads,
// with the factors of the matrix-core Cholesky (chol_mfma.hip), which leave the
// strictly lower part of each 16 x 16 diagonal block's inverse in the block's upper
// triangle: element (J + c, J + r), r > c, holds (L_JJ^-1)[r][c].  NF right-hand
// sides fixed at compile time, y[f * S + i] in LDS, arithmetic in AT; factor
// dimension n <= 512 (two rows per thread in the forward update).
#pragma once
#include <hip/hip_runtime.h>

namespace lrb {

constexpr int LEAN_NT = 256;

// lb[r * 17 + c] = (L_JJ^-1)[r][c] for the w <= 16 columns at J (zero elsewhere).
template <typename TL, typename AT>
__device__ __forceinline__ void llt_stage_inverse(const TL* __restrict__ L, int n, int J, int w, AT* lb) {
  const int t = threadIdx.x;
  const int r = t >> 4, c = t & 15;
  AT v = AT(0);
  if (r < w && c <= r) {
    const AT e = (AT)L[(size_t)(J + r) * n + J + c];
    v = r == c ? AT(1) / e : e;
  }
  lb[r * 17 + c] = v;
}

// y <- L^-1 y, then y <- L^-T y, for the NF columns y[f * S + i], i < n, in the
// arithmetic of AT.
template <typename TL, typename AT, int NF>
__device__ __forceinline__ void llt_solve_lean(const TL* __restrict__ L, int n, AT* y, int S, AT* zb, AT* lb) {
  const int tid = threadIdx.x;
  for (int J = 0; J < n; J += 16) {
    const int w = min(16, n - J);
    llt_stage_inverse<TL, AT>(L, n, J, w, lb);
    __syncthreads();
    if (tid < 16 * NF) {
      const int r = tid & 15, f = tid >> 4;
      AT z = AT(0);
#pragma unroll
      for (int c = 0; c < 16; c++) z += lb[r * 17 + c] * (c < w ? y[f * S + J + c] : AT(0));
      zb[r * NF + f] = z;
    }
    __syncthreads();
    if (tid < 16 * NF) {
      const int r = tid & 15, f = tid >> 4;
      if (r < w) y[f * S + J + r] = zb[r * NF + f];
    }
    // rows below the block (only under a full block): y_i -= L[i, J:J+16] z_J for the
    // rows J + 16 + tid and J + 272 + tid, which cover them all (n <= 512); written
    // without a loop, so that no z_J read is loop-invariant and held in registers
    {
      const int i0 = J + 16 + tid, i1 = i0 + LEAN_NT;
      if (i0 < n) {
        const bool h1 = i1 < n;
        TL l0[16], l1[16];
#pragma unroll
        for (int c = 0; c < 16; c++) {
          l0[c] = L[(size_t)(J + c) * n + i0];
          l1[c] = h1 ? L[(size_t)(J + c) * n + i1] : TL(0);
        }
#pragma unroll 1
        for (int f = 0; f < NF; f++) {
          AT a0 = AT(0), a1 = AT(0);
#pragma unroll
          for (int c = 0; c < 16; c++) {
            const AT z = zb[c * NF + f];
            a0 += (AT)l0[c] * z;
            a1 += (AT)l1[c] * z;
          }
          y[f * S + i0] -= a0;
          if (h1) y[f * S + i1] -= a1;
        }
      }
    }
    __syncthreads();
  }
  // backward: x_J = L_JJ^-T (z_J - L[below, J block]^T x_below); thread (c, g) sums
  // column J + c over the rows J + 16 + g + 16 k, the 16 lanes of one c reduce
  const int c = tid >> 4, gl = tid & 15;
  for (int J = ((n - 1) / 16) * 16; J >= 0; J -= 16) {
    const int w = min(16, n - J);
    llt_stage_inverse<TL, AT>(L, n, J, w, lb);
    AT acc[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) acc[f] = AT(0);
    if (c < w) {
      const TL* Lc = L + (size_t)(J + c) * n;
      int i = J + 16 + gl;
#pragma unroll 1
      for (; i + 48 < n; i += 64) {
        const AT l0 = (AT)Lc[i], l1 = (AT)Lc[i + 16], l2 = (AT)Lc[i + 32], l3 = (AT)Lc[i + 48];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          const AT* yf = y + f * S + i;
          acc[f] += l0 * yf[0] + l1 * yf[16] + l2 * yf[32] + l3 * yf[48];
        }
      }
#pragma unroll 1
      for (; i < n; i += 16) {
        const AT l0 = (AT)Lc[i];
#pragma unroll
        for (int f = 0; f < NF; f++) acc[f] += l0 * y[f * S + i];
      }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      AT a = acc[f];
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) a += __shfl_xor(a, o);
      if (gl == 0) zb[c * NF + f] = c < w ? y[f * S + J + c] - a : AT(0);
    }
    __syncthreads();
    if (tid < 16 * NF) {
      const int r = tid & 15, f = tid >> 4;
      AT x = AT(0);
#pragma unroll
      for (int c2 = 0; c2 < 16; c2++) x += lb[c2 * 17 + r] * zb[c2 * NF + f];   // (L_JJ^-T)[r][c2]
      if (r < w) y[f * S + J + r] = x;
    }
    __syncthreads();
  }
}

}  // namespace lrb
