// Solve of the patch systems L L^T X = B for patches of up to 512 members whose
// factor comes from the matrix-core kernel (chol_mfma.hip), i.e. with the inverse
// of every 16 x 16 diagonal block stored in its upper triangle (element (J + c,
// J + r), r > c, holds (L_JJ^-1)[r][c]).  One workgroup of 256 threads per patch,
// the number of fields NF fixed at compile time (the launch's nf rounded up to 1,
// 2, 3, 4, 6 or 8; padded fields are zero and never stored).
//
// The register footprint is what bounds this kernel: a solve is a chain of
// dependent global loads (2 x n / 16 blocks, each waiting on its panel), so its
// speed is the number of patches in flight per compute unit.  LDS holds only the
// right-hand side / solution y [NF][S] (S = the launch's largest patch); the
// forward update takes one row per thread and iteration; the backward dot
// products are unrolled by four rows.  With an FP32 factor the FP64 solution is
// refined against the FP64 matrix recomputed from the member coordinates
// (iterative refinement, PartitionSolve of partition_factor.c:92-110 solves once
// in FP64; the refinement converges to that solution): the residual is computed
// in FP64 and rounded into y, and the correction solves run in FP32 arithmetic,
// as mixed-precision iterative refinement has them (the FP32 factor's rounding,
// not the solve's, bounds each correction's accuracy); the refined FP64 solution
// stays in LDS beside the FP64 coordinates.
#include "kernels.hpp"
#include "lean_solve.hpp"

#include <algorithm>

namespace lrb {

static constexpr int LNT = LEAN_NT;   // threads per patch
static constexpr int LNMAX = 512; // largest patch

__device__ __forceinline__ double llt_centre(const Grid& g, int p, int d) {
  const int idx = d == 0 ? p % g.n[0] : (d == 1 ? (p / g.n[0]) % g.n[1] : p / (g.n[0] * g.n[1]));
  return g.O[d] + idx * g.h[d];
}

template <typename TL, int KT, int NF>
__global__ __launch_bounds__(LNT) void k_solve_llt(Grid g, const int32_t* __restrict__ list, const int32_t* __restrict__ pn,
                                                   const int64_t* __restrict__ mem_off, const int64_t* __restrict__ fac_off,
                                                   const int32_t* __restrict__ mem_pool, const double* __restrict__ xd,
                                                   const TL* __restrict__ fac, const double* __restrict__ values, int nf,
                                                   double* __restrict__ coef, int max_iter, double tol,
                                                   int32_t* __restrict__ status, int32_t* __restrict__ iters, int S,
                                                   const double* __restrict__ peps) {
  constexpr bool REF = sizeof(TL) == sizeof(float);
  using AT = TL;   // arithmetic of the triangular solves: FP64, or FP32 for the corrections
  extern __shared__ __align__(16) double lsm[];
  double* xl = lsm;                                 // REF: [NF][S] refined solution
  double* xr = xl + (REF ? (size_t)NF * S : 0);     // REF: [3][S] member coordinates
  AT* y = reinterpret_cast<AT*>(xr + (REF ? (size_t)3 * S : 0));   // [NF][S]
  __shared__ AT zb[16 * NF], lb[16 * 17];
  __shared__ double red[LNT / 64][2];
  const int p = list[blockIdx.x];
  const int n = pn[p];
  const int tid = threadIdx.x;
  if (n == 0) return;
  if (n > S) {   // not sized for this patch: reported as not converged
    if (tid == 0) { if (status) status[p] = 1; if (iters) iters[p] = 0; }
    return;
  }
  const int32_t* mem = mem_pool + mem_off[p];
  const TL* L = fac + fac_off[p];
  double* X = coef + mem_off[p] * nf;   // row-major [n][nf]
  for (int i = tid; i < n; i += LNT) {
    const int node = mem[i];
#pragma unroll
    for (int f = 0; f < NF; f++) y[f * S + i] = (AT)(f < nf ? values[(size_t)node * nf + f] : 0.0);
    if constexpr (REF) {
#pragma unroll
      for (int f = 0; f < NF; f++) xl[f * S + i] = 0.0;
      const double* x = xd + (size_t)node * g.dim;
      for (int d = 0; d < 3; d++) xr[d * S + i] = d < g.dim ? x[d] - llt_centre(g, p, d) : 0.0;
    }
  }
  __syncthreads();
  if constexpr (!REF) {
    llt_solve_lean<TL, AT, NF>(L, n, y, S, zb, lb);
    for (int i = tid; i < n; i += LNT)
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (f < nf) X[(size_t)i * nf + f] = y[f * S + i];
    if (tid == 0) { if (status) status[p] = 0; if (iters) iters[p] = 0; }
  } else {
    const double eps_p = peps ? peps[p] : g.eps_r;
    const int lane = tid & 63, wid = tid >> 6;
    int it = 0, converged = 0;
    double dprev = 1e300;
    for (; it <= max_iter; it++) {
      if (it > 0) {
        // residual b - A x in FP64 into y, A recomputed from the member coordinates
        for (int i = tid; i < n; i += LNT) {
          double s[NF];
          const int node = mem[i];
#pragma unroll
          for (int f = 0; f < NF; f++) s[f] = f < nf ? values[(size_t)node * nf + f] : 0.0;
          const double xi0 = xr[i], xi1 = xr[S + i], xi2 = xr[2 * S + i];
#pragma unroll 2
          for (int j = 0; j < n; j++) {
            const double d0 = xi0 - xr[j], d1 = xi1 - xr[S + j], d2 = xi2 - xr[2 * S + j];
            const double a = rbf<KT, double>(sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2), eps_p);
#pragma unroll
            for (int f = 0; f < NF; f++) s[f] -= a * xl[f * S + j];
          }
#pragma unroll
          for (int f = 0; f < NF; f++) y[f * S + i] = (AT)s[f];
        }
        __syncthreads();
      }
      llt_solve_lean<TL, AT, NF>(L, n, y, S, zb, lb);
      // x += d; converged when max |d| <= tol max |x|, or when the correction has
      // stopped shrinking (FP64 rounding floor of the residual) at or below 100 tol |x|
      double dmax = 0.0, xmax = 0.0;
      for (int i = tid; i < n; i += LNT)
#pragma unroll
        for (int f = 0; f < NF; f++) {
          const double d = (double)y[f * S + i];
          const double xn = xl[f * S + i] + d;
          xl[f * S + i] = xn;
          dmax = fmax(dmax, fabs(d));
          xmax = fmax(xmax, fabs(xn));
        }
      for (int o = 32; o > 0; o >>= 1) { dmax = fmax(dmax, __shfl_xor(dmax, o)); xmax = fmax(xmax, __shfl_xor(xmax, o)); }
      if (lane == 0) { red[wid][0] = dmax; red[wid][1] = xmax; }
      __syncthreads();
      dmax = red[0][0]; xmax = red[0][1];
      for (int w2 = 1; w2 < LNT / 64; w2++) { dmax = fmax(dmax, red[w2][0]); xmax = fmax(xmax, red[w2][1]); }
      __syncthreads();
      if (it > 0 && dmax <= tol * xmax) { converged = 1; break; }
      if (it > 1 && dmax > 0.5 * dprev && dmax <= 100.0 * tol * xmax) { converged = 1; break; }
      dprev = dmax;
    }
    for (int i = tid; i < n; i += LNT)
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (f < nf) X[(size_t)i * nf + f] = xl[f * S + i];
    if (tid == 0) { if (status) status[p] = converged ? 0 : 1; if (iters) iters[p] = it; }
  }
}

static size_t llt_smem(bool ref, int nf_pad, int S) {
  // REF: FP64 solution and coordinates, FP32 right-hand side; else the FP64 one
  return ref ? (size_t)S * (sizeof(double) * (nf_pad + 3) + sizeof(float) * nf_pad)
             : sizeof(double) * (size_t)S * nf_pad;
}

static int llt_nf_pad(int nf) { return nf <= 4 ? nf : (nf <= 6 ? 6 : 8); }

template <typename TL, int KT, int NF>
static void solve_llt_nf(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                         const int64_t* fac_off, const int32_t* mem_pool, const double* xd, const TL* fac,
                         const double* values, int nf, double* coef, int max_iter, double tol, int32_t* status,
                         int32_t* iters, int S, const double* peps, hipStream_t s) {
  constexpr bool REF = sizeof(TL) == sizeof(float);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_solve_llt<TL, KT, NF>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)llt_smem(REF, NF, LNMAX));
    attr = true;
  }
  hipLaunchKernelGGL((k_solve_llt<TL, KT, NF>), dim3(nlist), dim3(LNT), llt_smem(REF, NF, S), s, g, list, pn, mem_off,
                     fac_off, mem_pool, xd, fac, values, nf, coef, max_iter, tol, status, iters, S, peps);
}

template <typename TL, int KT>
static void solve_llt_kt(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                         const int64_t* fac_off, const int32_t* mem_pool, const double* xd, const TL* fac,
                         const double* values, int nf, double* coef, int max_iter, double tol, int32_t* status,
                         int32_t* iters, int S, const double* peps, hipStream_t s) {
#define LRB_NF(K) solve_llt_nf<TL, KT, K>(g, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, values, nf, coef, max_iter, tol, status, iters, S, peps, s)
  switch (llt_nf_pad(nf)) {
    case 1: LRB_NF(1); break;
    case 2: LRB_NF(2); break;
    case 3: LRB_NF(3); break;
    case 4: LRB_NF(4); break;
    case 6: LRB_NF(6); break;
    default: LRB_NF(8); break;
  }
#undef LRB_NF
}

int solve_llt_max_n() { return LNMAX; }

// The matrix is needed only by the FP32 factor's refinement, so the FP64 solves
// take one instantiation per field count whatever the kernel.
void launch_solve_llt(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                      const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                      const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                      int32_t* status, int32_t* iters, int nmax, const double* peps, hipStream_t s) {
  if (nlist <= 0 || nf <= 0) return;
  const int S = std::max(1, std::min(nmax, LNMAX));
  if (!fp32) {
    solve_llt_kt<double, K_WC2>(g, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (const double*)fac, values, nf, coef,
                                max_iter, tol, status, iters, S, peps, s);
    return;
  }
  const float* f = (const float*)fac;
#define LRB_K(K) solve_llt_kt<float, K>(g, list, nlist, pn, mem_off, fac_off, mem_pool, xd, f, values, nf, coef, max_iter, tol, status, iters, S, peps, s)
  switch (ktype) {
    case K_WC0: LRB_K(K_WC0); break;
    case K_WC2: LRB_K(K_WC2); break;
    case K_IMQ: LRB_K(K_IMQ); break;
    default: LRB_K(K_GAUSS); break;
  }
#undef LRB_K
}

}  // namespace lrb
