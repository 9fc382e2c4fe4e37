// Lean one-workgroup solve of the saddle-point patch systems [A P; P^T 0] [c; beta]
// = [f; 0] whose B22 factor comes from the matrix-core Cholesky (inverse diagonal
// blocks stored, lean_solve.hpp), for patches whose vectors fit in LDS (factor
// dimension n - m <= 512).  The same null-space method and refinement as
// k_solve_aug (aug.hip, which keeps the large patches and the other factors):
//   h = Q^T r, u1 = R^-T g, u2 = B22^-1 (h2 - B21 u1),
//   beta = R^-1 (h1 - B11 u1 - B12 u2), c = Q [u1; u2],
// with the residual (r, g) = (f - A c - P beta, -P^T c) of the full system in FP64.
//
// What makes it lean is the register and LDS footprint, which sets how many patches
// share a compute unit (the solve is a chain of dependent steps, each short):
//   - the number of fields NF is fixed at compile time (the launch's nf rounded up
//     to 1, 2, 3, 4, 6 or 8; padded fields are zero and never stored);
//   - LDS holds the coordinates, the solution and one work vector, (3 + 2 NF) S
//     doubles: the residual r and h = Q^T r share the work vector (h_i needs r_i
//     only);
//   - the m x NF dot products of a step (V^T r, P^T c, B1^T y, V^T u) take one
//     block reduction (wave shuffles, one barrier pair) instead of one per term.
// With an FP32 factor (the mixed path) the corrections are refined until they fall
// below tol |x| (or stagnate at the FP64 floor of the residual), as in k_solve_aug;
// their B22 solves run in FP32 arithmetic on an FP32 copy of the right-hand side.
#include "kernels.hpp"
#include "lean_solve.hpp"

#include <algorithm>

namespace lrb {

static constexpr int GNT = LEAN_NT;   // threads per patch
static constexpr int GW = 64;         // wavefront
static constexpr int GMAX = 512 + POLY_MAX;   // largest patch (members) handled here
static constexpr int GRT = (GMAX + GNT - 1) / GNT;   // rows per thread (row tid + r GNT)

// q-th polynomial term at x (the order of poly_eval: 1, x_a, then x_a x_b, a <= b)
__device__ __forceinline__ double poly_term(int dim, int q, const double* x) {
  if (q == 0) return 1.0;
  if (q <= dim) return x[q - 1];
  int k = q - 1 - dim;
  for (int a = 0; a < dim; a++) {
    const int cnt = dim - a;
    if (k < cnt) return x[a] * x[a + k];
    k -= cnt;
  }
  return 0.0;
}

// out[q * NF + f] = sum over rows i0 <= i < n of w(q, i) y[f * S + i], q < m: each
// thread sums its rows, the waves reduce by shuffles, their partials meet in LDS.
// red: shared [GNT / GW][MM * NF].  On return out is visible to every thread.
template <int NF, int MM, typename WF>
__device__ __forceinline__ void block_dots(int m, int i0, int n, const double* y, int S, WF wf, double* red,
                                           double* out) {
  constexpr int MF = MM * NF;
  const int tid = threadIdx.x, lane = tid & (GW - 1), wv = tid / GW;
#pragma unroll 1
  for (int q = 0; q < m; q++) {
    double part[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) part[f] = 0.0;
    for (int i = i0 + tid; i < n; i += GNT) {
      const double w = wf(q, i);
#pragma unroll
      for (int f = 0; f < NF; f++) part[f] += w * y[f * S + i];
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      double v = part[f];
#pragma unroll
      for (int o = GW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) red[wv * MF + q * NF + f] = v;
    }
  }
  __syncthreads();
  for (int e = tid; e < m * NF; e += GNT) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < GNT / GW; w++) t += red[w * MF + e];
    out[e] = t;
  }
  __syncthreads();
}

template <typename TL, int KT, int MM, int NF>
__global__ __launch_bounds__(GNT) void k_solve_aug_lean(Grid g, const int32_t* __restrict__ list,
                                                        const int32_t* __restrict__ pn,
                                                        const int64_t* __restrict__ mem_off,
                                                        const int64_t* __restrict__ fac_off,
                                                        const int32_t* __restrict__ mem_pool,
                                                        const TL* __restrict__ fac, const double* __restrict__ values,
                                                        int nf, double* __restrict__ coef, int max_iter, double tol,
                                                        int32_t* __restrict__ status, int32_t* __restrict__ iters,
                                                        int S, AugView a) {
  constexpr bool REF = sizeof(TL) == sizeof(float);
  constexpr int MF = MM * NF;
  extern __shared__ __align__(16) double gsm[];
  __shared__ double red[(GNT / GW) * MF];
  __shared__ double sv[MF], sb[MF], su1[MF], sh1[MF], sdb[MF], st2[MF];
  __shared__ double sR[MM * MM], sT[MM * MM], sB11[MM * MM];
  __shared__ double zb[16 * NF], lb[16 * 17];
  __shared__ float zb32[16 * NF], lb32[16 * 17];
  __shared__ double smx[GNT / GW][2];
  const int p = list[blockIdx.x];
  const int n = pn[p];
  const int m = a.m[p];
  const int tid = threadIdx.x;
  if (n == 0) return;
  if (n > S) {   // not sized for this patch: reported as not converged
    if (tid == 0) { if (status) status[p] = 1; if (iters) iters[p] = 0; }
    return;
  }
  const int nn = n - m, deg = a.deg[p], dim = g.dim;
  const double* xsg = a.aux + a.aux_off[p];   // [3][n] member coordinates about the PU node
  const double* V = xsg + 3 * (size_t)n;      // [m][n] Householder vectors
  const double* B1 = V + 2 * (size_t)m * n;   // [m][n] first m columns of Q^T A Q
  const double* R = B1 + (size_t)m * n;
  const double* Tm = R + (size_t)m * m;
  const int32_t* mem = mem_pool + mem_off[p];
  const TL* L = fac + fac_off[p];
  const double eps = a.eps ? a.eps[p] : g.eps_r;
  double* xs = gsm;                           // [3][S]
  double* c = xs + 3 * (size_t)S;             // [NF][S] solution
  double* y = c + (size_t)NF * S;             // [NF][S] residual, then h, then u
  float* y32 = reinterpret_cast<float*>(y + (size_t)NF * S);   // FP32 factor: [NF][S] B22 correction
  for (int i = tid; i < n; i += GNT) {
#pragma unroll
    for (int d = 0; d < 3; d++) xs[d * S + i] = xsg[(size_t)d * n + i];
#pragma unroll
    for (int f = 0; f < NF; f++) c[f * S + i] = 0.0;
  }
  for (int e = tid; e < m * m; e += GNT) {
    const int k = e % m, j = e / m;
    sR[j * MM + k] = R[(size_t)j * m + k];
    sT[j * MM + k] = Tm[(size_t)j * m + k];
    sB11[j * MM + k] = B1[(size_t)j * n + k];   // B(k, j), j < m
  }
  for (int e = tid; e < MF; e += GNT) sb[e] = 0.0;
  __syncthreads();
  auto vw = [&](int q, int i) { return V[(size_t)q * n + i]; };
  auto bw = [&](int q, int i) { return B1[(size_t)q * n + i]; };
  int it = 0, converged = 0;
  double dprev = 1e300;
  for (; it <= (REF ? max_iter : 0); it++) {
    // residual r = f - A c - P beta (into y) and g = -P^T c (into su1).  The per-row
    // passes keep the thread's rows tid + r GNT (r < GRT) in registers and loop over
    // the columns (or polynomial terms) outside them, so that only one column's or
    // term's values are live at a time.
    {
      double s[GRT][NF];
#pragma unroll
      for (int r = 0; r < GRT; r++) {
        const int i = tid + r * GNT;
        const int node = i < n ? mem[i] : 0;
#pragma unroll
        for (int f = 0; f < NF; f++) s[r][f] = (i < n && f < nf) ? values[(size_t)node * nf + f] : 0.0;
      }
      if (REF && it > 0) {
        double xr[GRT][3];
#pragma unroll
        for (int r = 0; r < GRT; r++) {
          const int i = min(tid + r * GNT, n - 1);
          xr[r][0] = xs[i]; xr[r][1] = xs[S + i]; xr[r][2] = xs[2 * S + i];
        }
#pragma unroll 1
        for (int j = 0; j < n; j++) {
          const double xj0 = xs[j], xj1 = xs[S + j], xj2 = xs[2 * S + j];
          double cj[NF];
#pragma unroll
          for (int f = 0; f < NF; f++) cj[f] = c[f * S + j];
#pragma unroll
          for (int r = 0; r < GRT; r++) {
            if (tid + r * GNT >= n) continue;
            const double d0 = xr[r][0] - xj0, d1 = xr[r][1] - xj1, d2 = xr[r][2] - xj2;
            const double phi = rbf<KT, double>(sqrt(d0 * d0 + d1 * d1 + d2 * d2), eps);
#pragma unroll
            for (int f = 0; f < NF; f++) s[r][f] -= phi * cj[f];
          }
        }
#pragma unroll 1
        for (int q = 0; q < m; q++) {
          double bq[NF];
#pragma unroll
          for (int f = 0; f < NF; f++) bq[f] = sb[q * NF + f];
#pragma unroll
          for (int r = 0; r < GRT; r++) {
            const double xi[3] = {xr[r][0] * a.ps, xr[r][1] * a.ps, xr[r][2] * a.ps};
            const double pq = poly_term(dim, q, xi);
#pragma unroll
            for (int f = 0; f < NF; f++) s[r][f] -= pq * bq[f];
          }
        }
      }
#pragma unroll
      for (int r = 0; r < GRT; r++) {
        const int i = tid + r * GNT;
        if (i < n)
#pragma unroll
          for (int f = 0; f < NF; f++) y[f * S + i] = s[r][f];
      }
    }
    if (REF && it > 0) {
      auto pw = [&](int q, int i) {
        const double xi[3] = {xs[i] * a.ps, xs[S + i] * a.ps, xs[2 * S + i] * a.ps};
        return -poly_term(dim, q, xi);
      };
      block_dots<NF, MM>(m, 0, n, c, S, pw, red, su1);   // its barriers also publish y
    } else {
      for (int e = tid; e < MF; e += GNT) su1[e] = 0.0;
      __syncthreads();
    }
    // t = V^T r; st2 = T^T t
    block_dots<NF, MM>(m, 0, n, y, S, vw, red, sv);
    for (int e = tid; e < m * NF; e += GNT) {
      const int q = e / NF, f = e % NF;
      double t = 0.0;
      for (int s2 = 0; s2 <= q; s2++) t += sT[q * MM + s2] * sv[s2 * NF + f];   // T(s2, q), upper triangular
      st2[q * NF + f] = t;
    }
    // u1 = R^-T g (forward substitution with R^T)
    if (tid < NF) {
      const int f = tid;
      for (int k = 0; k < m; k++) {
        double t = su1[k * NF + f];
        for (int j = 0; j < k; j++) t -= sR[k * MM + j] * su1[j * NF + f];   // R(j, k) u1_j
        su1[k * NF + f] = t / sR[k * MM + k];
      }
    }
    __syncthreads();
    // h = r - V st2: rows < m are h1, rows >= m become h2 - B21 u1
    {
      double h[GRT][NF];
#pragma unroll
      for (int r = 0; r < GRT; r++) {
        const int i = min(tid + r * GNT, n - 1);
#pragma unroll
        for (int f = 0; f < NF; f++) h[r][f] = y[f * S + i];
      }
#pragma unroll 1
      for (int q = 0; q < m; q++) {
        double tq[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) tq[f] = st2[q * NF + f];
#pragma unroll
        for (int r = 0; r < GRT; r++) {
          const double v = V[(size_t)q * n + min(tid + r * GNT, n - 1)];
#pragma unroll
          for (int f = 0; f < NF; f++) h[r][f] -= v * tq[f];
        }
      }
#pragma unroll
      for (int r = 0; r < GRT; r++) {
        const int i = tid + r * GNT;
        if (i < m)
#pragma unroll
          for (int f = 0; f < NF; f++) sh1[i * NF + f] = h[r][f];
      }
#pragma unroll 1
      for (int q = 0; q < m; q++) {
        double uq[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) uq[f] = su1[q * NF + f];
#pragma unroll
        for (int r = 0; r < GRT; r++) {
          const int i = tid + r * GNT;
          const double b = i >= m ? B1[(size_t)q * n + min(i, n - 1)] : 0.0;
#pragma unroll
          for (int f = 0; f < NF; f++) h[r][f] -= b * uq[f];
        }
      }
#pragma unroll
      for (int r = 0; r < GRT; r++) {
        const int i = tid + r * GNT;
        if (i < n)
#pragma unroll
          for (int f = 0; f < NF; f++) y[f * S + i] = h[r][f];
      }
    }
    __syncthreads();
    // u2 = B22^-1 (h2 - B21 u1); with an FP32 factor in FP32 arithmetic, as the
    // refinement's corrections need (the factor's rounding bounds their accuracy)
    if constexpr (REF) {
      for (int f = 0; f < NF; f++)
        for (int i = tid; i < nn; i += GNT) y32[f * S + i] = (float)y[f * S + m + i];
      __syncthreads();
      llt_solve_lean<TL, float, NF>(L, nn, y32, S, zb32, lb32);
      for (int f = 0; f < NF; f++)
        for (int i = tid; i < nn; i += GNT) y[f * S + m + i] = (double)y32[f * S + i];
      __syncthreads();
    } else {
      llt_solve_lean<TL, double, NF>(L, nn, y + m, S, zb, lb);
    }
    // beta correction: R^-1 (h1 - B11 u1 - B12 u2)
    block_dots<NF, MM>(m, m, n, y, S, bw, red, sv);
    if (tid < NF) {
      const int f = tid;
      for (int q = 0; q < m; q++) {
        double t = sh1[q * NF + f] - sv[q * NF + f];
        for (int k = 0; k < m; k++) t -= sB11[k * MM + q] * su1[k * NF + f];   // B(q, k)
        sdb[q * NF + f] = t;
      }
      for (int k = m - 1; k >= 0; k--) {
        double t = sdb[k * NF + f];
        for (int j = k + 1; j < m; j++) t -= sR[j * MM + k] * sdb[j * NF + f];   // R(k, j)
        sdb[k * NF + f] = t / sR[k * MM + k];
      }
    }
    for (int e = tid; e < m * NF; e += GNT) y[(e % NF) * S + e / NF] = su1[(e / NF) * NF + e % NF];
    __syncthreads();
    // dc = Q u = u - V T (V^T u)
    block_dots<NF, MM>(m, 0, n, y, S, vw, red, sv);
    for (int e = tid; e < m * NF; e += GNT) {
      const int q = e / NF, f = e % NF;
      double t = 0.0;
      for (int s2 = q; s2 < m; s2++) t += sT[s2 * MM + q] * sv[s2 * NF + f];   // T(q, s2)
      st2[q * NF + f] = t;
    }
    __syncthreads();
    double dmax = 0.0, xmax = 0.0;
    {
      double d[GRT][NF];
#pragma unroll
      for (int r = 0; r < GRT; r++) {
        const int i = min(tid + r * GNT, n - 1);
#pragma unroll
        for (int f = 0; f < NF; f++) d[r][f] = y[f * S + i];
      }
#pragma unroll 1
      for (int q = 0; q < m; q++) {
        double tq[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) tq[f] = st2[q * NF + f];
#pragma unroll
        for (int r = 0; r < GRT; r++) {
          const double v = V[(size_t)q * n + min(tid + r * GNT, n - 1)];
#pragma unroll
          for (int f = 0; f < NF; f++) d[r][f] -= v * tq[f];
        }
      }
#pragma unroll
      for (int r = 0; r < GRT; r++) {
        const int i = tid + r * GNT;
        if (i < n)
#pragma unroll
          for (int f = 0; f < NF; f++) {
            const double xn = c[f * S + i] + d[r][f];
            c[f * S + i] = xn;
            dmax = fmax(dmax, fabs(d[r][f]));
            xmax = fmax(xmax, fabs(xn));
          }
      }
    }
    if (tid < NF)
      for (int q = 0; q < m; q++) {
        const double d = sdb[q * NF + tid];
        const double xn = sb[q * NF + tid] + d;
        sb[q * NF + tid] = xn;
        dmax = fmax(dmax, fabs(d));
        xmax = fmax(xmax, fabs(xn));
      }
    for (int o = GW / 2; o > 0; o >>= 1) {
      dmax = fmax(dmax, __shfl_xor(dmax, o));
      xmax = fmax(xmax, __shfl_xor(xmax, o));
    }
    if ((tid & (GW - 1)) == 0) { smx[tid / GW][0] = dmax; smx[tid / GW][1] = xmax; }
    __syncthreads();
    dmax = 0.0; xmax = 0.0;
#pragma unroll
    for (int w = 0; w < GNT / GW; w++) { dmax = fmax(dmax, smx[w][0]); xmax = fmax(xmax, smx[w][1]); }
    __syncthreads();
    if (!REF) { converged = 1; break; }
    if (it > 0 && dmax <= tol * xmax) { converged = 1; break; }
    // stagnated at the FP64 rounding floor of the residual: accepted below 1e-8 |x|; a
    // factor too inaccurate to refine never shrinks its corrections that far and
    // falls back to FP64
    if (it > 1 && dmax > 0.5 * dprev && dmax <= fmax(100.0 * tol, 1e-8) * xmax) { converged = 1; break; }
    dprev = dmax;
  }
  double* X = coef + mem_off[p] * nf;
  for (int i = tid; i < n; i += GNT)
#pragma unroll
    for (int f = 0; f < NF; f++)
      if (f < nf) X[(size_t)i * nf + f] = c[f * S + i];
  double* bt = a.beta + (size_t)p * a.bstride;
  for (int e = tid; e < m * nf; e += GNT) bt[e] = sb[(e / nf) * NF + e % nf];
  if (tid == 0) {
    if (status) status[p] = converged ? 0 : 1;
    if (iters) iters[p] = it;
  }
}

// coordinates, solution and work vector in FP64; an FP32 factor adds the FP32
// correction vector
static size_t aug_lean_smem(bool ref, int nf_pad, int S) {
  return sizeof(double) * (size_t)S * (3 + 2 * (size_t)nf_pad) + (ref ? sizeof(float) * (size_t)S * nf_pad : 0);
}
static int aug_lean_nf_pad(int nf) { return nf <= 4 ? nf : (nf <= 6 ? 6 : 8); }

template <typename TL, int KT, int MM, int NF>
static void aug_lean_nf(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                        const int64_t* fac_off, const int32_t* mem_pool, const TL* fac, const double* values, int nf,
                        double* coef, int max_iter, double tol, int32_t* status, int32_t* iters, int S,
                        const AugView& a, hipStream_t s) {
  constexpr bool REF = sizeof(TL) == sizeof(float);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_solve_aug_lean<TL, KT, MM, NF>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)aug_lean_smem(REF, NF, GMAX));
    attr = true;
  }
  hipLaunchKernelGGL((k_solve_aug_lean<TL, KT, MM, NF>), dim3(nlist), dim3(GNT), aug_lean_smem(REF, NF, S), s, g, list, pn,
                     mem_off, fac_off, mem_pool, fac, values, nf, coef, max_iter, tol, status, iters, S, a);
}

template <typename TL, int KT, int MM>
static void aug_lean_mm(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                        const int64_t* fac_off, const int32_t* mem_pool, const TL* fac, const double* values, int nf,
                        double* coef, int max_iter, double tol, int32_t* status, int32_t* iters, int S,
                        const AugView& a, hipStream_t s) {
#define LRB_NF(K) aug_lean_nf<TL, KT, MM, K>(g, list, nlist, pn, mem_off, fac_off, mem_pool, fac, values, nf, coef, max_iter, tol, status, iters, S, a, s)
  switch (aug_lean_nf_pad(nf)) {
    case 1: LRB_NF(1); break;
    case 2: LRB_NF(2); break;
    case 3: LRB_NF(3); break;
    case 4: LRB_NF(4); break;
    case 6: LRB_NF(6); break;
    default: LRB_NF(8); break;
  }
#undef LRB_NF
}

template <typename TL, int KT>
static void aug_lean_kt(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                        const int64_t* fac_off, const int32_t* mem_pool, const TL* fac, const double* values, int nf,
                        double* coef, int max_iter, double tol, int32_t* status, int32_t* iters, int S,
                        const AugView& a, hipStream_t s) {
  if (a.mmax <= 4)
    aug_lean_mm<TL, KT, 4>(g, list, nlist, pn, mem_off, fac_off, mem_pool, fac, values, nf, coef, max_iter, tol, status,
                           iters, S, a, s);
  else
    aug_lean_mm<TL, KT, POLY_MAX>(g, list, nlist, pn, mem_off, fac_off, mem_pool, fac, values, nf, coef, max_iter, tol,
                                  status, iters, S, a, s);
}

int solve_aug_lean_max_n() { return GMAX; }

bool solve_aug_lean_supported(int ktype, bool fp32) { return !fp32 || ktype == K_PHS3 || ktype == K_GAUSS; }

// The matrix is needed only by the FP32 factor's refinement, so the FP64 solves
// take one instantiation per polynomial block and field count whatever the kernel.
void launch_solve_aug_lean(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                           const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const void* fac,
                           const double* values, int nf, double* coef, int max_iter, double tol, int32_t* status,
                           int32_t* iters, int nmax, const AugView& a, hipStream_t s) {
  if (nlist <= 0 || nf <= 0) return;
  const int S = std::max(1, std::min(nmax, GMAX));
  if (!fp32)
    aug_lean_kt<double, K_PHS3>(g, list, nlist, pn, mem_off, fac_off, mem_pool, (const double*)fac, values, nf, coef,
                                max_iter, tol, status, iters, S, a, s);
  else if (ktype == K_PHS3)
    aug_lean_kt<float, K_PHS3>(g, list, nlist, pn, mem_off, fac_off, mem_pool, (const float*)fac, values, nf, coef,
                               max_iter, tol, status, iters, S, a, s);
  else
    aug_lean_kt<float, K_GAUSS>(g, list, nlist, pn, mem_off, fac_off, mem_pool, (const float*)fac, values, nf, coef,
                                max_iter, tol, status, iters, S, a, s);
}

}  // namespace lrb
