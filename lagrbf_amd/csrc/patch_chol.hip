// Batched Cholesky factorisation and solve of the patch RBF systems for the
// positive definite kernels (Wendland C0/C2, Gaussian, inverse multiquadric).
//
// One workgroup per patch.  Left-looking, blocked by NB columns: the current
// panel A[J:n, J:J+NB] is assembled on the fly (phi of member distances, never
// stored in HBM), updated with the already factored columns L[J:n, 0:J] read
// from the patch's factor slot (L2/MALL resident), factored in LDS and written
// back.  T = double gives the FP64 path, T = float the FP32 factor of the
// mixed path, whose solve refines the FP64 coefficients against the FP64
// matrix recomputed on the fly (iterative refinement, residual in FP64).
//
// Coordinates are shifted to the patch's PU node in FP64 before any rounding to
// T, so the FP32 matrix carries relative, not absolute, rounding.
#include "kernels.hpp"
#include "blocked_solve.hpp"

#include <algorithm>
#include <type_traits>

#include <cmath>

namespace lrb {

static constexpr int CW = 64;          // wavefront
static constexpr int CNT = 256;        // threads per patch
static constexpr int CNB = 32;         // panel width
static constexpr int CTT = 64;         // k-chunk of the panel update
static constexpr int CNMAX = 512;      // largest patch handled here

// loop over the nf <= 8 data columns with the register arrays kept in VGPRs
#define FOR_F _Pragma("unroll") for (int f = 0; f < 8; f++) if (f < nf)

__device__ __forceinline__ double patch_centre(const Grid& g, int p, int d) {
  int idx = d == 0 ? p % g.n[0] : (d == 1 ? (p / g.n[0]) % g.n[1] : p / (g.n[0] * g.n[1]));
  return g.O[d] + idx * g.h[d];
}

// Saddle-point systems (MM > 0, extension): the factored matrix is B22 of
// Q^T A Q = A - X V^T - V X^T (V, X from k_aug_prep), restricted to members m..n-1;
// its entries are assembled in FP64 (phi minus the rank-2m correction) before any
// rounding to T.  AT is the assembly type.
template <int MM, typename T>
using assembly_t = typename std::conditional<(MM > 0), double, T>::type;
template <int MM, typename T>
__host__ __device__ constexpr int chol_kchunk() { return (MM > 4 && sizeof(T) == 8) ? 32 : CTT; }

template <typename T, int KT, int MM>
__global__ __launch_bounds__(CNT) void k_factor_chol(Grid g, const int32_t* __restrict__ list,
                                                     const int32_t* __restrict__ pn, const int64_t* __restrict__ mem_off,
                                                     const int64_t* __restrict__ fac_off,
                                                     const int32_t* __restrict__ mem_pool,
                                                     const double* __restrict__ xd, T* __restrict__ fac,
                                                     int32_t* __restrict__ status, AugView aug) {
  using AT = assembly_t<MM, T>;
  constexpr int TT = chol_kchunk<MM, T>();
  constexpr int MR = MM > 0 ? MM : 1;
  extern __shared__ __align__(16) unsigned char smem[];
  const int p = list[blockIdx.x];
  const int n0 = pn[p];
  const int m = MM > 0 ? aug.m[p] : 0;
  const int n = n0 - m;                               // factor dimension
  if (n <= 0) { if (threadIdx.x == 0) status[p] = 0; return; }
  AT* xr = reinterpret_cast<AT*>(smem);               // [3][CNMAX] member coords relative to the patch node
  T* P = reinterpret_cast<T*>(xr + 3 * CNMAX);        // panel [rows][CNB] column-major, ld = rows
  T* Lr = P + (size_t)CNMAX * CNB;                    // [CNB][TT] rows J..J+w of L, columns t0..t0+TT
  double* cv = reinterpret_cast<double*>(Lr + CNB * TT);   // [2][MM][CNB] V and X rows of the panel columns
  __shared__ int s_fail;
  const int tid = threadIdx.x;
  const int dim = g.dim;
  const int32_t* mem = mem_pool + mem_off[p] + m;     // factor row i = member m + i
  T* L = fac + fac_off[p];
  const double* V = MM > 0 ? aug.aux + aug.aux_off[p] + 3 * (size_t)n0 + m : nullptr;   // V[q * n0 + i]
  const double* X = MM > 0 ? V + (size_t)m * n0 : nullptr;
  if (tid == 0) s_fail = 0;
  for (int i = tid; i < n; i += CNT) {
    const double* x = xd + (size_t)mem[i] * dim;
    for (int d = 0; d < 3; d++) xr[d * CNMAX + i] = d < dim ? (AT)(x[d] - patch_centre(g, p, d)) : AT(0);
  }
  __syncthreads();
  const AT eps = (AT)(aug.eps ? aug.eps[p] : g.eps_r);
  for (int J = 0; J < n; J += CNB) {
    const int w = min(CNB, n - J), rows = n - J;
    if constexpr (MM > 0) {
      __syncthreads();
      for (int e = tid; e < 2 * MM * CNB; e += CNT) {
        const int which = e / (MM * CNB), q = (e / CNB) % MM, c = e % CNB;
        cv[e] = (q < m && c < w) ? (which ? X : V)[(size_t)q * n0 + J + c] : 0.0;
      }
      __syncthreads();
    }
    for (int r0 = 0; r0 < rows; r0 += CNT) {
      const int r = r0 + tid;
      const bool live = r < rows;
      const int gi = J + (live ? r : 0);
      T acc[CNB];
      const AT xi0 = xr[gi], xi1 = xr[CNMAX + gi], xi2 = xr[2 * CNMAX + gi];
      double rv[MR], rx[MR];
      if constexpr (MM > 0) {
#pragma unroll
        for (int q = 0; q < MM; q++) {
          rv[q] = q < m ? V[(size_t)q * n0 + gi] : 0.0;
          rx[q] = q < m ? X[(size_t)q * n0 + gi] : 0.0;
        }
      }
#pragma unroll
      for (int c = 0; c < CNB; c++) {
        const int gj = J + (c < w ? c : 0);
        const AT d0 = xi0 - xr[gj], d1 = xi1 - xr[CNMAX + gj], d2 = xi2 - xr[2 * CNMAX + gj];
        AT a = rbf<KT, AT>(sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2), eps);
        if constexpr (MM > 0) {
#pragma unroll
          for (int q = 0; q < MM; q++) a -= rx[q] * cv[q * CNB + c] + rv[q] * cv[(MM + q) * CNB + c];
        }
        acc[c] = (T)a;
      }
      for (int t0 = 0; t0 < J; t0 += TT) {
        const int tw = min(TT, J - t0);
        __syncthreads();
        for (int e = tid; e < CNB * TT; e += CNT) {
          const int c = e / TT, tt = e % TT;
          Lr[e] = (c < w && tt < tw) ? L[(size_t)(t0 + tt) * n + J + c] : T(0);
        }
        __syncthreads();
        if (live) {
          for (int tt = 0; tt < tw; tt++) {
            const T lit = L[(size_t)(t0 + tt) * n + gi];
#pragma unroll
            for (int c = 0; c < CNB; c++) acc[c] -= lit * Lr[c * TT + tt];
          }
        }
      }
      if (live)
#pragma unroll
        for (int c = 0; c < CNB; c++) if (c < w) P[(size_t)c * rows + r] = acc[c];
    }
    __syncthreads();
    // unblocked right-looking factorisation of the panel (rows J..n-1, w columns)
    for (int c = 0; c < w; c++) {
      const T pcc = P[(size_t)c * rows + c];
      if (!(pcc > T(0))) { if (tid == 0) s_fail = J + c + 1; }
      const T dd = sqrt(pcc > T(0) ? pcc : T(1));
      const T inv = T(1) / dd;
      for (int r = c + 1 + tid; r < rows; r += CNT) P[(size_t)c * rows + r] *= inv;
      __syncthreads();
      if (tid == 0) P[(size_t)c * rows + c] = dd;
      for (int r = c + 1 + tid; r < rows; r += CNT) {
        const T lrc = P[(size_t)c * rows + r];
        const int cmax = min(w - 1, r);
        for (int c2 = c + 1; c2 <= cmax; c2++) P[(size_t)c2 * rows + r] -= lrc * P[(size_t)c * rows + c2];
      }
      __syncthreads();
    }
    for (int c = 0; c < w; c++)
      for (int r = c + tid; r < rows; r += CNT) L[(size_t)(J + c) * n + J + r] = P[(size_t)c * rows + r];
    __syncthreads();
  }
  if (tid == 0) status[p] = s_fail;
}

// Solve L L^T X = B for the nf data columns of a patch, B gathered from the node
// values (PartitionMatSetRHS, partition_factor.c:57-90).  With a float factor the
// FP64 solution is refined: r = B - A X with A recomputed in FP64, X += solve(r),
// until the correction is below tol * |X| (or max_iter); status 1 = not converged.
template <typename TL, int KT, int NT>
__global__ __launch_bounds__(NT, 3) void k_solve_chol(Grid g, const int32_t* __restrict__ list, const int32_t* __restrict__ pn,
                                                    const int64_t* __restrict__ mem_off, const int64_t* __restrict__ fac_off,
                                                    const int32_t* __restrict__ mem_pool, const double* __restrict__ xd,
                                                    const TL* __restrict__ fac, const double* __restrict__ values, int nf,
                                                    double* __restrict__ coef, int max_iter, double tol,
                                                    int32_t* __restrict__ status, int32_t* __restrict__ iters, int S,
                                                    const double* __restrict__ peps, const uint8_t* __restrict__ fnb) {
  // S: row stride of the LDS arrays (the largest patch of the launch)
  extern __shared__ __align__(16) unsigned char smem[];
  double* xr = reinterpret_cast<double*>(smem);      // [3][S] relative coords (FP64)
  double* b = xr + (size_t)3 * S;                    // [nf][S] right-hand side / residual
  double* y = b + (size_t)nf * S;                    // [nf][S] work vector
  double* xl = y + (size_t)nf * S;                   // [nf][S] the solution, kept here until the end
  __shared__ double red[NT / CW][8];
  __shared__ double zb[SOLVE_NB * 8], lb[SOLVE_LB];
  const int p = list[blockIdx.x];
  const int n = pn[p];
  const int tid = threadIdx.x, lane = tid & (CW - 1), wid = tid / CW;
  if (n == 0 || nf == 0) return;
  const int nbp = fnb ? fnb[p] : 0;   // matrix-core factor (16-column blocks): blocked solves
  if (n > S) {   // not sized for this patch: reported as not converged (refactored in FP64)
    if (tid == 0) {
      if (status) status[p] = 1;
      if (iters) iters[p] = 0;
    }
    return;
  }
  const int dim = g.dim;
  const int32_t* mem = mem_pool + mem_off[p];
  const TL* L = fac + fac_off[p];
  const double eps_p = peps ? peps[p] : g.eps_r;
  double* X = coef + mem_off[p] * nf;                // row-major [n][nf]
  for (int i = tid; i < n; i += NT) {
    const double* x = xd + (size_t)mem[i] * dim;
    for (int d = 0; d < 3; d++) xr[d * S + i] = d < dim ? x[d] - patch_centre(g, p, d) : 0.0;
    for (int f = 0; f < nf; f++) { const double v = values[(size_t)mem[i] * nf + f]; b[f * S + i] = v; xl[f * S + i] = 0.0; }
  }
  __syncthreads();
  const bool refine = sizeof(TL) == sizeof(float);
  int it = 0;
  int converged = 0;
  double dprev = 1e300;
  for (; it <= (refine ? max_iter : 0); it++) {
    if (it > 0) {
      // residual b - A X in FP64, A recomputed from the member coordinates
      for (int i = tid; i < n; i += NT) {
        double s[8];
        FOR_F s[f] = values[(size_t)mem[i] * nf + f];
        const double xi0 = xr[i], xi1 = xr[S + i], xi2 = xr[2 * S + i];
        for (int j = 0; j < n; j++) {
          const double d0 = xi0 - xr[j], d1 = xi1 - xr[S + j], d2 = xi2 - xr[2 * S + j];
          const double a = rbf<KT, double>(sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2), eps_p);
          FOR_F s[f] -= a * xl[f * S + j];
        }
        FOR_F b[f * S + i] = s[f];
      }
      __syncthreads();
    }
    for (int f = 0; f < nf; f++)
      for (int i = tid; i < n; i += NT) y[f * S + i] = b[f * S + i];
    __syncthreads();
    if (nbp == SOLVE_NB) {
      llt_solve_blocked<TL, NT>(L, n, y, S, nf, zb, lb);
    } else {
    // forward: L z = b (column oriented).  A thread holds rows k+1+tid and
    // k+1+tid+NT of column k (n <= 2 NT); the next column's entries are loaded one
    // step ahead, so the global loads overlap the current step instead of
    // stalling every one of the n sequential steps.
    static_assert(2 * NT >= CNMAX, "two rows per thread cover a column");
    {
      double dk = (double)L[0];
      double l0 = 1 + tid < n ? (double)L[1 + tid] : 0.0;
      double l1 = 1 + tid + NT < n ? (double)L[1 + tid + NT] : 0.0;
      for (int k = 0; k < n; k++) {
        double dn = 0.0, n0 = 0.0, n1 = 0.0;
        if (k + 1 < n) {
          const TL* c1 = L + (size_t)(k + 1) * n;
          dn = (double)c1[k + 1];
          if (k + 2 + tid < n) n0 = (double)c1[k + 2 + tid];
          if (k + 2 + tid + NT < n) n1 = (double)c1[k + 2 + tid + NT];
        }
        double zk[8];
        FOR_F zk[f] = y[f * S + k] / dk;
        const int i0 = k + 1 + tid, i1 = i0 + NT;
        if (i0 < n) {
          const double lik = l0;
          FOR_F y[f * S + i0] -= lik * zk[f];
        }
        if (i1 < n) {
          const double lik = l1;
          FOR_F y[f * S + i1] -= lik * zk[f];
        }
        __syncthreads();
        if (tid == 0) FOR_F y[f * S + k] = zk[f];
        dk = dn; l0 = n0; l1 = n1;
      }
    }
    __syncthreads();
    // backward: L^T d = z (dot products down each column), the previous column's
    // entries loaded one step ahead
    double b0 = 0.0, b1 = 0.0;                          // column n-1 has no rows below
    double dkb = (double)L[(size_t)(n - 1) * n + (n - 1)];
    for (int k = n - 1; k >= 0; k--) {
      double pn0 = 0.0, pn1 = 0.0, pd = 0.0;
      if (k >= 1) {
        const TL* cp = L + (size_t)(k - 1) * n;
        pd = (double)cp[k - 1];
        if (k + tid < n) pn0 = (double)cp[k + tid];
        if (k + tid + NT < n) pn1 = (double)cp[k + tid + NT];
      }
      double s[8];
      FOR_F s[f] = 0.0;
      {
        const int i0 = k + 1 + tid, i1 = i0 + NT;
        if (i0 < n) {
          const double lik = b0;
          FOR_F s[f] += lik * y[f * S + i0];
        }
        if (i1 < n) {
          const double lik = b1;
          FOR_F s[f] += lik * y[f * S + i1];
        }
      }
      b0 = pn0; b1 = pn1;
      FOR_F {
        double v = s[f];
        for (int o = CW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wid][f] = v;
      }
      __syncthreads();
      if (tid < nf) {
        double t = red[0][tid];
        for (int w2 = 1; w2 < NT / CW; w2++) t += red[w2][tid];
        y[tid * S + k] = (y[tid * S + k] - t) / dkb;
      }
      dkb = pd;
      __syncthreads();
    }
    }
    // X += d; convergence test on max |d| / max |X|
    double dmax = 0.0, xmax = 0.0;
    for (int i = tid; i < n; i += NT)
      for (int f = 0; f < nf; f++) {
        const double d = y[f * S + i];
        const double xn = xl[f * S + i] + d;
        xl[f * S + i] = xn;
        dmax = fmax(dmax, fabs(d));
        xmax = fmax(xmax, fabs(xn));
      }
    for (int o = CW / 2; o > 0; o >>= 1) { dmax = fmax(dmax, __shfl_xor(dmax, o)); xmax = fmax(xmax, __shfl_xor(xmax, o)); }
    if (lane == 0) { red[wid][0] = dmax; red[wid][1] = xmax; }
    __syncthreads();
    dmax = red[0][0]; xmax = red[0][1];
    for (int w2 = 1; w2 < NT / CW; w2++) { dmax = fmax(dmax, red[w2][0]); xmax = fmax(xmax, red[w2][1]); }
    __syncthreads();
    if (!refine) { converged = 1; break; }
    // done when the correction is below tol * |X|, or when it has stopped shrinking
    // (FP64 rounding floor of the residual, ~cond(A) * 1e-16) at or below 100 tol * |X|
    if (it > 0 && dmax <= tol * xmax) { converged = 1; break; }
    if (it > 1 && dmax > 0.5 * dprev && dmax <= 100.0 * tol * xmax) { converged = 1; break; }
    dprev = dmax;
  }
  for (int i = tid; i < n; i += NT)
    for (int f = 0; f < nf; f++) X[(size_t)i * nf + f] = xl[f * S + i];
  if (tid == 0) {
    if (status) status[p] = converged ? 0 : 1;
    if (iters) iters[p] = it;
  }
}

// ---------------------------------------------------------------------------
// Patches larger than CNMAX: the same left-looking blocked Cholesky without any
// n-sized LDS.  Per panel J: the w x w diagonal block is assembled, updated with
// the factored columns and factored in LDS; every row below it (in chunks of CNT,
// one row per thread) is assembled, updated and solved against the diagonal block
// (l_r = a_r L_JJ^-T, forward substitution).  Member coordinates are gathered
// from the node array and shifted to the patch node in FP64.
template <typename T, int KT, int MM>
__global__ __launch_bounds__(CNT) void k_factor_chol_big(Grid g, const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ pn,
                                                         const int64_t* __restrict__ mem_off,
                                                         const int64_t* __restrict__ fac_off,
                                                         const int32_t* __restrict__ mem_pool,
                                                         const double* __restrict__ xd, T* __restrict__ fac,
                                                         int32_t* __restrict__ status, AugView aug) {
  using AT = assembly_t<MM, T>;
  constexpr int MR = MM > 0 ? MM : 1;
  __shared__ T Lr[CNB * CTT];
  __shared__ T D[CNB * CNB];
  __shared__ AT colx[3][CNB];
  __shared__ double cv[2 * MR * CNB];
  __shared__ int s_fail;
  const int p = list[blockIdx.x];
  const int n0 = pn[p];
  const int m = MM > 0 ? aug.m[p] : 0;
  const int n = n0 - m;                               // factor dimension
  const int tid = threadIdx.x, dim = g.dim;
  if (n <= 0) { if (tid == 0) status[p] = 0; return; }
  const int32_t* mem = mem_pool + mem_off[p] + m;     // factor row i = member m + i
  T* L = fac + fac_off[p];
  const double* V = MM > 0 ? aug.aux + aug.aux_off[p] + 3 * (size_t)n0 + m : nullptr;   // V[q * n0 + i]
  const double* X = MM > 0 ? V + (size_t)m * n0 : nullptr;
  const double cx = patch_centre(g, p, 0), cy = patch_centre(g, p, 1), cz = dim == 3 ? patch_centre(g, p, 2) : 0.0;
  if (tid == 0) s_fail = 0;
  const AT eps = (AT)(aug.eps ? aug.eps[p] : g.eps_r);
  for (int J = 0; J < n; J += CNB) {
    const int w = min(CNB, n - J), rows = n - J;
    __syncthreads();
    if (tid < CNB) {
      const int gj = J + (tid < w ? tid : 0);
      const double* x = xd + (size_t)mem[gj] * dim;
      colx[0][tid] = (AT)(x[0] - cx); colx[1][tid] = (AT)(x[1] - cy); colx[2][tid] = dim == 3 ? (AT)(x[2] - cz) : AT(0);
    }
    if constexpr (MM > 0) {
      for (int e = tid; e < 2 * MM * CNB; e += CNT) {
        const int which = e / (MM * CNB), q = (e / CNB) % MM, c = e % CNB;
        cv[e] = (q < m && c < w) ? (which ? X : V)[(size_t)q * n0 + J + c] : 0.0;
      }
    }
    __syncthreads();
    for (int r0 = 0; r0 < rows; r0 += CNT) {
      const int r = r0 + tid;
      const bool live = r < rows;
      const int gi = J + (live ? r : 0);
      const double* x = xd + (size_t)mem[gi] * dim;
      const AT xi0 = (AT)(x[0] - cx), xi1 = (AT)(x[1] - cy), xi2 = dim == 3 ? (AT)(x[2] - cz) : AT(0);
      double rv[MR], rx[MR];
      if constexpr (MM > 0) {
#pragma unroll
        for (int q = 0; q < MM; q++) {
          rv[q] = q < m ? V[(size_t)q * n0 + gi] : 0.0;
          rx[q] = q < m ? X[(size_t)q * n0 + gi] : 0.0;
        }
      }
      T acc[CNB];
#pragma unroll
      for (int c = 0; c < CNB; c++) {
        const AT d0 = xi0 - colx[0][c], d1 = xi1 - colx[1][c], d2 = xi2 - colx[2][c];
        AT a = rbf<KT, AT>(sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2), eps);
        if constexpr (MM > 0) {
#pragma unroll
          for (int q = 0; q < MM; q++) a -= rx[q] * cv[q * CNB + c] + rv[q] * cv[(MM + q) * CNB + c];
        }
        acc[c] = (T)a;
      }
      for (int t0 = 0; t0 < J; t0 += CTT) {
        const int tw = min(CTT, J - t0);
        __syncthreads();
        for (int e = tid; e < CNB * CTT; e += CNT) {
          const int c = e / CTT, tt = e % CTT;
          Lr[e] = (c < w && tt < tw) ? L[(size_t)(t0 + tt) * n + J + c] : T(0);
        }
        __syncthreads();
        if (live) {
          for (int tt = 0; tt < tw; tt++) {
            const T lit = L[(size_t)(t0 + tt) * n + gi];
#pragma unroll
            for (int c = 0; c < CNB; c++) acc[c] -= lit * Lr[c * CTT + tt];
          }
        }
      }
      if (r0 == 0) {
        // diagonal block: rows 0..w-1 of the first chunk, factored in LDS
        if (live && r < w)
#pragma unroll
          for (int c = 0; c < CNB; c++) if (c < w) D[c * CNB + r] = acc[c];
        __syncthreads();
        for (int c = 0; c < w; c++) {
          const T pcc = D[c * CNB + c];
          if (!(pcc > T(0)) && tid == 0 && s_fail == 0) s_fail = J + c + 1;
          const T dd = sqrt(pcc > T(0) ? pcc : T(1));
          __syncthreads();
          if (tid > c && tid < w) D[c * CNB + tid] /= dd;
          __syncthreads();
          if (tid == 0) D[c * CNB + c] = dd;
          if (tid > c && tid < w) {
            const T lrc = D[c * CNB + tid];
            for (int c2 = c + 1; c2 <= tid; c2++) D[c2 * CNB + tid] -= lrc * D[c * CNB + c2];
          }
          __syncthreads();
        }
        for (int e = tid; e < w * w; e += CNT) {
          const int c = e / w, rr = e % w;
          if (rr >= c) L[(size_t)(J + c) * n + J + rr] = D[c * CNB + rr];
        }
      }
      __syncthreads();
      // rows below the diagonal block: l_r = a_r L_JJ^-T
      if (live && r >= w) {
#pragma unroll
        for (int c = 0; c < CNB; c++) {
          if (c < w) {
            T v = acc[c];
            for (int c2 = 0; c2 < c; c2++) v -= acc[c2] * D[c2 * CNB + c];
            acc[c] = v / D[c * CNB + c];
            L[(size_t)(J + c) * n + gi] = acc[c];
          }
        }
      }
    }
  }
  __syncthreads();
  if (tid == 0) status[p] = s_fail;
}

// Solve and FP64 refinement for patches larger than CNMAX; right-hand side and
// work vectors live in a global workspace (2 nf doubles per member, indexed by
// the patch's member offset).
template <typename TL, int KT>
__global__ __launch_bounds__(CNT) void k_solve_chol_big(Grid g, const int32_t* __restrict__ list,
                                                        const int32_t* __restrict__ pn,
                                                        const int64_t* __restrict__ mem_off,
                                                        const int64_t* __restrict__ fac_off,
                                                        const int32_t* __restrict__ mem_pool,
                                                        const double* __restrict__ xd, const TL* __restrict__ fac,
                                                        const double* __restrict__ values, int nf, double* __restrict__ coef,
                                                        int max_iter, double tol, double* __restrict__ ws,
                                                        int32_t* __restrict__ status, int32_t* __restrict__ iters,
                                                        const double* __restrict__ peps,
                                                        const uint8_t* __restrict__ fnb) {
  __shared__ double red[CNT / CW][8];
  __shared__ double zk_s[8];
  __shared__ double zb[SOLVE_NB * 8], lb[SOLVE_LB];
  const int p = list[blockIdx.x];
  const int n = pn[p];
  const int tid = threadIdx.x, lane = tid & (CW - 1), wid = tid / CW;
  if (n == 0 || nf == 0) return;
  const int dim = g.dim;
  const int32_t* mem = mem_pool + mem_off[p];
  const TL* L = fac + fac_off[p];
  double* X = coef + mem_off[p] * nf;
  double* b = ws + (size_t)mem_off[p] * 2 * nf;     // [nf][n]
  double* y = b + (size_t)nf * n;                   // [nf][n]
  const double cx = patch_centre(g, p, 0), cy = patch_centre(g, p, 1), cz = dim == 3 ? patch_centre(g, p, 2) : 0.0;
  for (int i = tid; i < n; i += CNT)
    for (int f = 0; f < nf; f++) { b[(size_t)f * n + i] = values[(size_t)mem[i] * nf + f]; X[(size_t)i * nf + f] = 0.0; }
  __syncthreads();
  const bool refine = sizeof(TL) == sizeof(float);
  int it = 0, converged = 0;
  double dprev = 1e300;
  for (; it <= (refine ? max_iter : 0); it++) {
    if (it > 0) {
      for (int i = tid; i < n; i += CNT) {
        double sres[8];
        FOR_F sres[f] = values[(size_t)mem[i] * nf + f];
        const double* xi = xd + (size_t)mem[i] * dim;
        const double xi0 = xi[0] - cx, xi1 = xi[1] - cy, xi2 = dim == 3 ? xi[2] - cz : 0.0;
        for (int j = 0; j < n; j++) {
          const double* xj = xd + (size_t)mem[j] * dim;
          const double d0 = xi0 - (xj[0] - cx), d1 = xi1 - (xj[1] - cy), d2 = dim == 3 ? xi2 - (xj[2] - cz) : 0.0;
          const double a = rbf<KT, double>(sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2), peps ? peps[p] : g.eps_r);
          FOR_F sres[f] -= a * X[(size_t)j * nf + f];
        }
        FOR_F b[(size_t)f * n + i] = sres[f];
      }
      __syncthreads();
    }
    for (int f = 0; f < nf; f++)
      for (int i = tid; i < n; i += CNT) y[(size_t)f * n + i] = b[(size_t)f * n + i];
    __syncthreads();
    if (fnb && fnb[p] == SOLVE_NB) {
      // factor of the tiled matrix-core kernel: blocked solves on its inverted diagonal blocks
      llt_solve_blocked<TL, CNT>(L, n, y, n, nf, zb, lb);
    } else {
    for (int k = 0; k < n; k++) {
      if (tid == 0) {
        const double lkk = (double)L[(size_t)k * n + k];
        FOR_F { zk_s[f] = y[(size_t)f * n + k] / lkk; y[(size_t)f * n + k] = zk_s[f]; }
      }
      __syncthreads();
      double zk[8];
      FOR_F zk[f] = zk_s[f];
      for (int i = k + 1 + tid; i < n; i += CNT) {
        const double lik = (double)L[(size_t)k * n + i];
        FOR_F y[(size_t)f * n + i] -= lik * zk[f];
      }
      __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {
      double sacc[8];
      FOR_F sacc[f] = 0.0;
      for (int i = k + 1 + tid; i < n; i += CNT) {
        const double lik = (double)L[(size_t)k * n + i];
        FOR_F sacc[f] += lik * y[(size_t)f * n + i];
      }
      FOR_F {
        double v = sacc[f];
        for (int o = CW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wid][f] = v;
      }
      __syncthreads();
      if (tid < nf) {
        double t = red[0][tid];
        for (int w2 = 1; w2 < CNT / CW; w2++) t += red[w2][tid];
        y[(size_t)tid * n + k] = (y[(size_t)tid * n + k] - t) / (double)L[(size_t)k * n + k];
      }
      __syncthreads();
    }
    }
    double dmax = 0.0, xmax = 0.0;
    for (int i = tid; i < n; i += CNT)
      for (int f = 0; f < nf; f++) {
        const double d = y[(size_t)f * n + i];
        const double xn = X[(size_t)i * nf + f] + d;
        X[(size_t)i * nf + f] = xn;
        dmax = fmax(dmax, fabs(d));
        xmax = fmax(xmax, fabs(xn));
      }
    for (int o = CW / 2; o > 0; o >>= 1) { dmax = fmax(dmax, __shfl_xor(dmax, o)); xmax = fmax(xmax, __shfl_xor(xmax, o)); }
    if (lane == 0) { red[wid][0] = dmax; red[wid][1] = xmax; }
    __syncthreads();
    dmax = red[0][0]; xmax = red[0][1];
    for (int w2 = 1; w2 < CNT / CW; w2++) { dmax = fmax(dmax, red[w2][0]); xmax = fmax(xmax, red[w2][1]); }
    __syncthreads();
    if (!refine) { converged = 1; break; }
    if (it > 0 && dmax <= tol * xmax) { converged = 1; break; }
    if (it > 1 && dmax > 0.5 * dprev && dmax <= 100.0 * tol * xmax) { converged = 1; break; }
    dprev = dmax;
  }
  if (tid == 0) {
    if (status) status[p] = converged ? 0 : 1;
    if (iters) iters[p] = it;
  }
}

static size_t factor_smem(size_t esz, size_t asz, int mm, int tt) {
  return asz * 3 * CNMAX + esz * ((size_t)CNMAX * CNB + (size_t)CNB * tt) + sizeof(double) * 2 * (size_t)mm * CNB;
}
// LDS of the solve: coordinates [3][S]; right-hand side, work vector, solution [nf][S]
static size_t solve_smem(int S, int nf) { return sizeof(double) * (size_t)S * (3 + 3 * (size_t)nf); }
static constexpr int SNT = 256;  // threads per patch in the solve (measured against one wave:
                                 // the residual's parallel work outweighs the cheaper barriers)

int chol_max_n() { return CNMAX; }

template <typename T, int KT, int MM>
static void factor_chol_k(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                          const int64_t* fac_off, const int32_t* mem_pool, const double* xd, T* fac, int32_t* status,
                          bool big, const AugView& a, hipStream_t s) {
  if (big) {
    hipLaunchKernelGGL((k_factor_chol_big<T, KT, MM>), dim3(nlist), dim3(CNT), 0, s, g, list, pn, mem_off, fac_off,
                       mem_pool, xd, fac, status, a);
    return;
  }
  static bool attr = false;
  const size_t sm = factor_smem(sizeof(T), sizeof(assembly_t<MM, T>), MM, chol_kchunk<MM, T>());
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_factor_chol<T, KT, MM>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    attr = true;
  }
  hipLaunchKernelGGL((k_factor_chol<T, KT, MM>), dim3(nlist), dim3(CNT), sm, s, g, list, pn, mem_off, fac_off, mem_pool,
                     xd, fac, status, a);
}

template <typename TL, int KT>
static void solve_chol_k(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                         const int64_t* fac_off, const int32_t* mem_pool, const double* xd, const TL* fac,
                         const double* values, int nf, double* coef, int max_iter, double tol, int32_t* status,
                         int32_t* iters, double* ws, int nmax, const double* peps, const uint8_t* fnb, hipStream_t s) {
  if (ws) {
    hipLaunchKernelGGL((k_solve_chol_big<TL, KT>), dim3(nlist), dim3(CNT), 0, s, g, list, pn, mem_off, fac_off, mem_pool,
                       xd, fac, values, nf, coef, max_iter, tol, ws, status, iters, peps, fnb);
    return;
  }
  static bool attr = false;
  const int S = std::max(1, std::min(nmax, CNMAX));
  const size_t sm = solve_smem(S, nf);
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_solve_chol<TL, KT, SNT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)solve_smem(CNMAX, 8));
    attr = true;
  }
  hipLaunchKernelGGL((k_solve_chol<TL, KT, SNT>), dim3(nlist), dim3(SNT), sm, s, g, list, pn, mem_off, fac_off, mem_pool,
                     xd, fac, values, nf, coef, max_iter, tol, status, iters, S, peps, fnb);
}

// positive definite kernels; r^3 only on the null space of a polynomial block
bool chol_kernel_supported(int ktype) { return ktype == K_WC0 || ktype == K_WC2 || ktype == K_GAUSS || ktype == K_IMQ; }

template <typename T, int MM>
static void factor_chol_mm(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                           const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                           T* fac, int32_t* status, bool big, const AugView& a, hipStream_t s) {
#define LRB_FC(K) factor_chol_k<T, K, MM>(g, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, big, a, s)
  switch (ktype) {
    case K_WC0: LRB_FC(K_WC0); break;
    case K_WC2: LRB_FC(K_WC2); break;
    case K_IMQ: LRB_FC(K_IMQ); break;
    case K_PHS3:
      if constexpr (MM > 0) LRB_FC(K_PHS3);
      break;
    default: LRB_FC(K_GAUSS); break;
  }
#undef LRB_FC
}

template <typename T>
static void factor_chol_dispatch(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                                 const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                                 T* fac, int32_t* status, bool big, const AugView& a, hipStream_t s) {
  if (!a.m) factor_chol_mm<T, 0>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, big, a, s);
  else if (a.mmax <= 4) factor_chol_mm<T, 4>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, big, a, s);
  else factor_chol_mm<T, POLY_MAX>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, big, a, s);
}

void launch_factor_chol(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                        const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                        void* fac, int32_t* status, bool big, const AugView& a, int nb, int nmax, hipStream_t s) {
  if (nlist <= 0) return;
  if (!big && nb > 0) {
    launch_factor_chol_mfma(g, ktype, fp32, nmax, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s);
    return;
  }
  if (fp32) factor_chol_dispatch<float>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (float*)fac, status, big, a, s);
  else factor_chol_dispatch<double>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (double*)fac, status, big, a, s);
}

template <typename TL>
static void solve_chol_dispatch(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                                const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                                const TL* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                                int32_t* status, int32_t* iters, double* ws, int nmax, const double* peps,
                                const uint8_t* fnb, hipStream_t s) {
#define LRB_SC(K) solve_chol_k<TL, K>(g, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, peps, fnb, s)
  switch (ktype) {
    case K_WC0: LRB_SC(K_WC0); break;
    case K_WC2: LRB_SC(K_WC2); break;
    case K_IMQ: LRB_SC(K_IMQ); break;
    default: LRB_SC(K_GAUSS); break;
  }
#undef LRB_SC
}

void launch_solve_chol(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                       const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                       const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                       int32_t* status, int32_t* iters, double* ws, int nmax, const AugView& a, hipStream_t s) {
  if (nlist <= 0) return;
  if (fp32) solve_chol_dispatch<float>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (const float*)fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, a.eps, a.fnb, s);
  else solve_chol_dispatch<double>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (const double*)fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, a.eps, a.fnb, s);
}

}  // namespace lrb
