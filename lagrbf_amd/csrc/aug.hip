// Saddle-point patch systems [A P; P^T 0] [c; beta] = [f; 0] (extension with no
// reference counterpart: the reference solves pure RBF systems,
// python_api/lagrbfpy/interpolation.py:43-45; the factor it would extend is the
// dsytrf of partition_factor.c:37-55).
//
// Null-space method, one workgroup per patch, everything in FP64 except the
// Cholesky factor of the mixed path:
//   P = Q [R; 0]  Householder QR of the n x m polynomial block (k_aug_prep), Q =
//                 I - V T V^T in compact WY form;
//   Q^T A Q = A - X V^T - V X^T with U = A V, S = T^T (V^T U) T, X = U T - V S / 2;
//   B22 = (Q^T A Q)[m:, m:] is positive definite for the kernels used here (r^3 is
//   conditionally positive definite of order 2, so linear or quadratic blocks make
//   it definite on the null space of P^T); patch_chol.hip factors it with its
//   entries assembled on the fly from phi, X and V;
//   solve (k_solve_aug): for the right-hand side (r, g) of the saddle-point
//   system, h = Q^T r, u1 = R^-T g, u2 = B22^-1 (h2 - B21 u1),
//   beta = R^-1 (h1 - B11 u1 - B12 u2), c = Q [u1; u2].
// The mixed path factors B22 in FP32 and refines [c; beta] against the FP64
// residual of the full saddle-point system (f - A c - P beta, -P^T c), A recomputed
// from the member coordinates, until the correction is below tol.
#include "kernels.hpp"
#include "blocked_solve.hpp"

#include <algorithm>

namespace lrb {

static constexpr int ANT = 256;      // threads per patch
static constexpr int AW = 64;        // wavefront
static constexpr int AMAXF = 8;      // fields handled by the solve

// Sum of v[0..k) over the block; the result in out[0..k) (shared), visible to every
// thread on return.  red: shared [ANT / AW][K].
template <int K>
__device__ __forceinline__ void block_sum_vec(const double (&v)[K], int k, double* red, double* out) {
  const int lane = threadIdx.x & (AW - 1), wid = threadIdx.x / AW;
#pragma unroll
  for (int e = 0; e < K; e++) {
    if (e < k) {
      double x = v[e];
      for (int o = AW / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
      if (lane == 0) red[wid * K + e] = x;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < k; e += ANT) {
    double t = 0.0;
    for (int w = 0; w < ANT / AW; w++) t += red[w * K + e];
    out[e] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ double centre_of(const Grid& g, int p, int d) {
  const int idx = d == 0 ? p % g.n[0] : (d == 1 ? (p / g.n[0]) % g.n[1] : p / (g.n[0] * g.n[1]));
  return g.O[d] + idx * g.h[d];
}

// ---------------------------------------------------------------------------
// Per patch: member coordinates shifted to the PU node, Householder QR of P,
// T, X and the first m columns B1 of Q^T A Q.
template <int KT, int MM>
__global__ __launch_bounds__(ANT) void k_aug_prep(Grid g, const int32_t* __restrict__ list, const int32_t* __restrict__ pn,
                                                  const int64_t* __restrict__ mem_off,
                                                  const int32_t* __restrict__ mem_pool, const double* __restrict__ xd,
                                                  AugView a, int stage) {
  // stage 0: everything; 1: steps 1-4 (QR, T); 2: steps 6-8, after U = A V came from
  // k_aug_av_big (large patches: the O(n^2) product over many workgroups)
  __shared__ double red[(ANT / AW) * MM];
  __shared__ double sv[MM];
  __shared__ double sG[MM * MM], sT[MM * MM], sM[MM * MM], sS[MM * MM];
  __shared__ double stau[MM], sbeta[MM];
  const int p = list[blockIdx.x];
  const int n = pn[p];
  const int m = a.m[p];
  if (n == 0 || m == 0) return;
  const int tid = threadIdx.x, dim = g.dim, deg = a.deg[p];
  double* xs = a.aux + a.aux_off[p];   // [3][n]
  double* V = xs + 3 * (size_t)n;      // [m][n]
  double* X = V + (size_t)m * n;       // [m][n] (U = A V first)
  double* B1 = X + (size_t)m * n;      // [m][n]
  double* R = B1 + (size_t)m * n;      // R(k, j) at R[j * m + k]
  double* Tm = R + (size_t)m * m;      // T(k, j) at Tm[j * m + k]
  const int32_t* mem = mem_pool + mem_off[p];
  const double eps = a.eps ? a.eps[p] : g.eps_r;
  const double c0 = centre_of(g, p, 0), c1 = centre_of(g, p, 1), c2 = dim == 3 ? centre_of(g, p, 2) : 0.0;
  if (stage == 2) {
    for (int e = tid; e < MM * MM; e += ANT) {
      const int q = e / MM, r = e % MM;
      sT[e] = (q < m && r < m) ? Tm[(size_t)q * m + r] : 0.0;
    }
    __syncthreads();
  } else {
  // 1. coordinates and the polynomial block
  for (int i = tid; i < n; i += ANT) {
    const double* x = xd + (size_t)mem[i] * dim;
    double xi[3], pv[POLY_MAX];
    xs[i] = x[0] - c0;
    xs[n + i] = x[1] - c1;
    xs[2 * n + i] = dim == 3 ? x[2] - c2 : 0.0;
    for (int d = 0; d < 3; d++) xi[d] = xs[d * n + i] * a.ps;
    poly_eval(dim, deg, xi, pv);
    for (int q = 0; q < m; q++) V[(size_t)q * n + i] = pv[q];
  }
  __syncthreads();
  // 2. Householder QR (LAPACK dgeqr2 / dlarfg conventions: H = I - tau v v^T, v_k = 1)
  for (int k = 0; k < m; k++) {
    double part[1] = {0.0};
    for (int i = k + 1 + tid; i < n; i += ANT) part[0] += V[(size_t)k * n + i] * V[(size_t)k * n + i];
    block_sum_vec<1>(part, 1, red, sv);
    const double xn2 = sv[0];
    const double alpha = V[(size_t)k * n + k];
    double beta = alpha, tau = 0.0, scl = 0.0;
    if (xn2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tau = (beta - alpha) / beta;
      scl = 1.0 / (alpha - beta);
    }
    __syncthreads();
    if (tau != 0.0)
      for (int i = k + 1 + tid; i < n; i += ANT) V[(size_t)k * n + i] *= scl;
    if (tid == 0) { V[(size_t)k * n + k] = 1.0; stau[k] = tau; sbeta[k] = beta; }
    __syncthreads();
    double wv[MM];
#pragma unroll
    for (int j = 0; j < MM; j++) wv[j] = 0.0;
    for (int i = k + tid; i < n; i += ANT) {
      const double vi = V[(size_t)k * n + i];
#pragma unroll
      for (int j = 0; j < MM; j++)
        if (j > k && j < m) wv[j] += vi * V[(size_t)j * n + i];
    }
    block_sum_vec<MM>(wv, m, red, sv);
    for (int i = k + tid; i < n; i += ANT) {
      const double vi = V[(size_t)k * n + i];
      for (int j = k + 1; j < m; j++) V[(size_t)j * n + i] -= tau * sv[j] * vi;
    }
    __syncthreads();
  }
  // 3. R from the upper triangle, then V below its unit diagonal only
  for (int e = tid; e < m * m; e += ANT) {
    const int k = e % m, j = e / m;
    R[(size_t)j * m + k] = k < j ? V[(size_t)j * n + k] : (k == j ? sbeta[j] : 0.0);
  }
  __syncthreads();
  for (int e = tid; e < m * m; e += ANT) {
    const int k = e % m, j = e / m;
    if (k < j) V[(size_t)j * n + k] = 0.0;
  }
  __syncthreads();
  // 4. Gram matrix of the reflectors, then T (dlarft, forward, columnwise)
  for (int q = 0; q < m; q++) {
    double wv[MM];
#pragma unroll
    for (int r = 0; r < MM; r++) wv[r] = 0.0;
    for (int i = tid; i < n; i += ANT) {
      const double vq = V[(size_t)q * n + i];
#pragma unroll
      for (int r = 0; r < MM; r++)
        if (r < m) wv[r] += vq * V[(size_t)r * n + i];
    }
    block_sum_vec<MM>(wv, m, red, sG + q * MM);
  }
  if (tid == 0) {
    for (int k = 0; k < m; k++) {
      for (int i = 0; i < k; i++) {
        double t = 0.0;
        for (int j = i; j < k; j++) t += sT[j * MM + i] * (-stau[k] * sG[j * MM + k]);
        sT[k * MM + i] = t;
      }
      sT[k * MM + k] = stau[k];
      for (int i = k + 1; i < m; i++) sT[k * MM + i] = 0.0;
    }
  }
  __syncthreads();
  for (int e = tid; e < m * m; e += ANT) Tm[(size_t)(e / m) * m + e % m] = sT[(e / m) * MM + e % m];
  if (stage == 1) return;
  // 5. U = A V (into X)
  for (int i = tid; i < n; i += ANT) {
    double u[MM];
#pragma unroll
    for (int q = 0; q < MM; q++) u[q] = 0.0;
    const double xi0 = xs[i], xi1 = xs[n + i], xi2 = xs[2 * n + i];
    for (int j = 0; j < n; j++) {
      const double d0 = xi0 - xs[j], d1 = xi1 - xs[n + j], d2 = xi2 - xs[2 * n + j];
      const double phi = rbf<KT, double>(sqrt(d0 * d0 + d1 * d1 + d2 * d2), eps);
#pragma unroll
      for (int q = 0; q < MM; q++)
        if (q < m) u[q] += phi * V[(size_t)q * n + j];
    }
    for (int q = 0; q < m; q++) X[(size_t)q * n + i] = u[q];
  }
  }
  __syncthreads();
  // 6. M = V^T U, S = T^T M T (symmetrised)
  for (int q = 0; q < m; q++) {
    double wv[MM];
#pragma unroll
    for (int r = 0; r < MM; r++) wv[r] = 0.0;
    for (int i = tid; i < n; i += ANT) {
      const double vq = V[(size_t)q * n + i];
#pragma unroll
      for (int r = 0; r < MM; r++)
        if (r < m) wv[r] += vq * X[(size_t)r * n + i];
    }
    block_sum_vec<MM>(wv, m, red, sM + q * MM);   // sM[q * MM + r] = M(q, r)
  }
  if (tid == 0) {
    double W[MM * MM];
    for (int q = 0; q < m; q++)
      for (int r = 0; r < m; r++) {
        double t = 0.0;
        for (int s2 = 0; s2 < m; s2++) t += sM[q * MM + s2] * sT[r * MM + s2];   // (M T)(q, r)
        W[q * MM + r] = t;
      }
    for (int q = 0; q < m; q++)
      for (int r = 0; r < m; r++) {
        double t = 0.0;
        for (int s2 = 0; s2 < m; s2++) t += sT[q * MM + s2] * W[s2 * MM + r];     // (T^T W)(q, r)
        sS[q * MM + r] = t;
      }
    for (int q = 0; q < m; q++)
      for (int r = 0; r < q; r++) {
        const double t = 0.5 * (sS[q * MM + r] + sS[r * MM + q]);
        sS[q * MM + r] = t;
        sS[r * MM + q] = t;
      }
  }
  __syncthreads();
  // 7. X = U T - V S / 2
  for (int i = tid; i < n; i += ANT) {
    double u[MM], v[MM];
#pragma unroll
    for (int q = 0; q < MM; q++) {
      u[q] = q < m ? X[(size_t)q * n + i] : 0.0;
      v[q] = q < m ? V[(size_t)q * n + i] : 0.0;
    }
    for (int q = 0; q < m; q++) {
      double t = 0.0;
      for (int r = 0; r < m; r++) t += u[r] * sT[q * MM + r] - 0.5 * v[r] * sS[r * MM + q];
      X[(size_t)q * n + i] = t;
    }
  }
  __syncthreads();
  // 8. B1 = first m columns of Q^T A Q
  for (int e = tid; e < m * n; e += ANT) {
    const int k = e / n, i = e % n;
    const double d0 = xs[i] - xs[k], d1 = xs[n + i] - xs[n + k], d2 = xs[2 * n + i] - xs[2 * n + k];
    double b = rbf<KT, double>(sqrt(d0 * d0 + d1 * d1 + d2 * d2), eps);
    for (int q = 0; q < m; q++)
      b -= X[(size_t)q * n + i] * V[(size_t)q * n + k] + V[(size_t)q * n + i] * X[(size_t)q * n + k];
    B1[(size_t)k * n + i] = b;
  }
}

// U = A V of the large patches (step 5 of k_aug_prep) over many workgroups: one
// block of ANT rows per task (patch, row block), the columns through LDS in chunks;
// each row sums in ascending column order with the same operations, so the values
// are those of the one-workgroup loop.
template <int KT, int MM>
__global__ __launch_bounds__(ANT) void k_aug_av_big(Grid g, const int2* __restrict__ tasks,
                                                    const int32_t* __restrict__ pn, AugView a) {
  __shared__ double sx[3][ANT];
  __shared__ double sv[MM][ANT];
  const int2 tk = tasks[blockIdx.x];
  const int p = tk.x;
  const int n = pn[p], m = a.m[p];
  if (n == 0 || m == 0) return;
  const int tid = threadIdx.x;
  const int i = tk.y * ANT + tid;
  const bool live = i < n;
  const int ii = live ? i : n - 1;
  const double* xs = a.aux + a.aux_off[p];
  const double* V = xs + 3 * (size_t)n;
  double* X = const_cast<double*>(V) + (size_t)m * n;
  const double eps = a.eps ? a.eps[p] : g.eps_r;
  double u[MM];
#pragma unroll
  for (int q = 0; q < MM; q++) u[q] = 0.0;
  const double xi0 = xs[ii], xi1 = xs[n + ii], xi2 = xs[2 * n + ii];
  for (int j0 = 0; j0 < n; j0 += ANT) {
    const int jn = min(ANT, n - j0);
    __syncthreads();
    if (tid < jn) {
      sx[0][tid] = xs[j0 + tid]; sx[1][tid] = xs[n + j0 + tid]; sx[2][tid] = xs[2 * n + j0 + tid];
#pragma unroll
      for (int q = 0; q < MM; q++)
        if (q < m) sv[q][tid] = V[(size_t)q * n + j0 + tid];
    }
    __syncthreads();
    for (int jj = 0; jj < jn; jj++) {
      const double d0 = xi0 - sx[0][jj], d1 = xi1 - sx[1][jj], d2 = xi2 - sx[2][jj];
      const double phi = rbf<KT, double>(sqrt(d0 * d0 + d1 * d1 + d2 * d2), eps);
#pragma unroll
      for (int q = 0; q < MM; q++)
        if (q < m) u[q] += phi * sv[q][jj];
    }
  }
  if (live)
    for (int q = 0; q < m; q++) X[(size_t)q * n + i] = u[q];
}

// ---------------------------------------------------------------------------
// L L^T u = y in place for nf columns (y[f * ldy + i], i < nn), L column-major with
// leading dimension nn.  Column-oriented forward sweep, dot-product backward sweep.
template <typename TL>
__device__ void llt_solve(const TL* __restrict__ L, int nn, double* y, int ldy, int nf, double* red /* [ANT/AW][8] */,
                          double* zk /* [8] */) {
  const int tid = threadIdx.x, lane = tid & (AW - 1), wid = tid / AW;
  for (int k = 0; k < nn; k++) {
    const TL* ck = L + (size_t)k * nn;
    if (tid < nf) {
      const double z = y[(size_t)tid * ldy + k] / (double)ck[k];
      y[(size_t)tid * ldy + k] = z;
      zk[tid] = z;
    }
    __syncthreads();
    for (int i = k + 1 + tid; i < nn; i += ANT) {
      const double lik = (double)ck[i];
      for (int f = 0; f < nf; f++) y[(size_t)f * ldy + i] -= lik * zk[f];
    }
    __syncthreads();
  }
  for (int k = nn - 1; k >= 0; k--) {
    const TL* ck = L + (size_t)k * nn;
    double s[AMAXF];
#pragma unroll
    for (int f = 0; f < AMAXF; f++) s[f] = 0.0;
    for (int i = k + 1 + tid; i < nn; i += ANT) {
      const double lik = (double)ck[i];
#pragma unroll
      for (int f = 0; f < AMAXF; f++)
        if (f < nf) s[f] += lik * y[(size_t)f * ldy + i];
    }
#pragma unroll
    for (int f = 0; f < AMAXF; f++) {
      if (f < nf) {
        double v = s[f];
        for (int o = AW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wid * AMAXF + f] = v;
      }
    }
    __syncthreads();
    if (tid < nf) {
      double t = 0.0;
      for (int w = 0; w < ANT / AW; w++) t += red[w * AMAXF + tid];
      y[(size_t)tid * ldy + k] = (y[(size_t)tid * ldy + k] - t) / (double)ck[k];
    }
    __syncthreads();
  }
}

// out[q * AMAXF + f] = sum over rows i0 <= i < n of W[q * n + i] * y[f * ldy + i],
// q < m, f < nf (block reductions, one per column q).
__device__ void col_dots(const double* __restrict__ W, int n, int i0, int m, const double* y, int ldy, int nf,
                         double* red, double* out) {
  for (int q = 0; q < m; q++) {
    double tp[AMAXF];
#pragma unroll
    for (int f = 0; f < AMAXF; f++) tp[f] = 0.0;
    for (int i = i0 + threadIdx.x; i < n; i += ANT) {
      const double wq = W[(size_t)q * n + i];
#pragma unroll
      for (int f = 0; f < AMAXF; f++)
        if (f < nf) tp[f] += wq * y[(size_t)f * ldy + i];
    }
    block_sum_vec<AMAXF>(tp, nf, red, out + q * AMAXF);
  }
}

// Solve of the saddle-point systems (see the file comment).  Vectors live in LDS
// (stride S) when the patch fits, else in ws at the patch's member offset.
template <typename TL, int KT, int MM>
__global__ __launch_bounds__(ANT, 3) void k_solve_aug(Grid g, const int32_t* __restrict__ list,
                                                   const int32_t* __restrict__ pn, const int64_t* __restrict__ mem_off,
                                                   const int64_t* __restrict__ fac_off,
                                                   const int32_t* __restrict__ mem_pool, const double* __restrict__ xd,
                                                   const TL* __restrict__ fac, const double* __restrict__ values, int nf,
                                                   double* __restrict__ coef, int max_iter, double tol,
                                                   int32_t* __restrict__ status, int32_t* __restrict__ iters,
                                                   double* __restrict__ ws, int S, AugView a, int step_it,
                                                   double* __restrict__ dstate, int phase,
                                                   double* __restrict__ hstate) {
  extern __shared__ __align__(16) double sm[];
  constexpr int MF = MM * AMAXF;
  __shared__ double red[(ANT / AW) * MF];
  __shared__ double sv[MF], sb[MF], su1[MF], sh1[MF], sdb[MF], st2[MF], zk[AMAXF];
  __shared__ double sR[MM * MM], sT[MM * MM], sB11[MM * MM];
  __shared__ double zb[SOLVE_NB * 8], lb[SOLVE_LB];
  const int p = list[blockIdx.x];
  const int n = pn[p];
  const int m = a.m[p];
  const int tid = threadIdx.x;
  if (n == 0 || nf == 0) return;
  const int nn = n - m, deg = a.deg[p], dim = g.dim;
  const double* xsg = a.aux + a.aux_off[p];
  const double* V = xsg + 3 * (size_t)n;
  const double* B1 = V + 2 * (size_t)m * n;
  const double* R = B1 + (size_t)m * n;
  const double* Tm = R + (size_t)m * m;
  const int32_t* mem = mem_pool + mem_off[p];
  const TL* L = fac + fac_off[p];
  const double eps = a.eps ? a.eps[p] : g.eps_r;
  // step_it >= 0: one refinement iteration per launch (the large patches of the FP32
  // factor), the state kept in the workspace (c), in the polynomial coefficients
  // (beta), in status (2 while running), iters and dstate (the last correction);
  // the residual of iterations >= 1 comes from k_aug_resid_big.  step_it < 0: the
  // whole solve in this launch.
  const bool stepping = step_it >= 0;
  if (stepping && step_it > 0 && status[p] != 2) return;
  const bool lds = n <= S && !stepping && phase == 0;   // a two-phase solve keeps its vectors in ws
  if (!lds && !ws) {   // not sized for this patch: reported as not converged (refactored in FP64)
    if (tid == 0) {
      if (status) status[p] = 1;
      if (iters) iters[p] = 0;
    }
    return;
  }
  const int ldv = lds ? S : n;
  double* xs = lds ? sm : ws + (size_t)mem_off[p] * (3 + 3 * (size_t)nf);
  double* c = xs + 3 * (size_t)ldv;     // [nf][ldv] solution
  double* y = c + (size_t)nf * ldv;     // [nf][ldv] work vector
  double* r = y + (size_t)nf * ldv;     // [nf][ldv] residual
  for (int i = tid; i < n; i += ANT) {
    for (int d = 0; d < 3; d++) xs[(size_t)d * ldv + i] = xsg[(size_t)d * n + i];
    if (!(stepping && step_it > 0))
      for (int f = 0; f < nf; f++) c[(size_t)f * ldv + i] = 0.0;
  }
  for (int e = tid; e < m * m; e += ANT) {
    const int k = e % m, j = e / m;
    sR[j * MM + k] = R[(size_t)j * m + k];
    sT[j * MM + k] = Tm[(size_t)j * m + k];
    sB11[j * MM + k] = B1[(size_t)j * n + k];   // B(k, j), j < m
  }
  double* bt = a.beta + (size_t)p * a.bstride;
  for (int e = tid; e < MF; e += ANT) {
    const int q = e / AMAXF, f = e % AMAXF;
    sb[e] = (stepping && step_it > 0 && q < m && f < nf) ? bt[q * nf + f] : 0.0;
  }
  __syncthreads();
  const bool refine = sizeof(TL) == sizeof(float);
  int it = stepping ? step_it : 0, converged = 0;
  const int it_end = stepping ? step_it : (refine ? max_iter : 0);
  double dprev = (stepping && step_it > 0) ? dstate[p] : 1e300;
  // phase 1: up to h and u1, which go to hstate (at the block's list position) with
  // the B22 right-hand side left in the workspace for the multi-workgroup triangular
  // solves (trsv_tiles.hip); phase 2: from u2 = that solution on.  Phase 0: all.
  for (; it <= it_end; it++) {
    if (phase == 2) {
      const double* hs = hstate + (size_t)blockIdx.x * 2 * MF;
      for (int e = tid; e < MF; e += ANT) { su1[e] = hs[e]; sh1[e] = hs[MF + e]; }
      __syncthreads();
      goto after_b22;
    }
    // residual of the saddle-point system: r = f - A c - P beta, g = -P^T c  (g in su1)
    {
      // (iterations >= 1 of a stepped solve: r is already in the workspace)
      if (!(stepping && it > 0))
      for (int i = tid; i < n; i += ANT) {
        double s[AMAXF];
#pragma unroll
        for (int f = 0; f < AMAXF; f++) s[f] = f < nf ? values[(size_t)mem[i] * nf + f] : 0.0;
        if (it > 0) {
          const double xi0 = xs[i], xi1 = xs[ldv + i], xi2 = xs[2 * (size_t)ldv + i];
          for (int j = 0; j < n; j++) {
            const double d0 = xi0 - xs[j], d1 = xi1 - xs[ldv + j], d2 = xi2 - xs[2 * (size_t)ldv + j];
            const double phi = rbf<KT, double>(sqrt(d0 * d0 + d1 * d1 + d2 * d2), eps);
#pragma unroll
            for (int f = 0; f < AMAXF; f++)
              if (f < nf) s[f] -= phi * c[(size_t)f * ldv + j];
          }
          double xi[3] = {xi0 * a.ps, xi1 * a.ps, xi2 * a.ps}, pv[POLY_MAX];
          poly_eval(dim, deg, xi, pv);
#pragma unroll
          for (int q = 0; q < MM; q++)
            if (q < m)
#pragma unroll
              for (int f = 0; f < AMAXF; f++)
                if (f < nf) s[f] -= pv[q] * sb[q * AMAXF + f];
        }
#pragma unroll
        for (int f = 0; f < AMAXF; f++)
          if (f < nf) r[(size_t)f * ldv + i] = s[f];
      }
      for (int q = 0; q < m; q++) {
        double gp[AMAXF];
#pragma unroll
        for (int f = 0; f < AMAXF; f++) gp[f] = 0.0;
        if (it > 0)
          for (int i = tid; i < n; i += ANT) {
            double xi[3] = {xs[i] * a.ps, xs[ldv + i] * a.ps, xs[2 * (size_t)ldv + i] * a.ps}, pv[POLY_MAX];
            poly_eval(dim, deg, xi, pv);
#pragma unroll
            for (int f = 0; f < AMAXF; f++)
              if (f < nf) gp[f] -= pv[q] * c[(size_t)f * ldv + i];
          }
        block_sum_vec<AMAXF>(gp, nf, red, su1 + q * AMAXF);
      }
    }
    // t = V^T r, then h = r - V T^T t (into y)
    col_dots(V, n, 0, m, r, ldv, nf, red, sv);
    for (int e = tid; e < m * nf; e += ANT) {
      const int q = e / nf, f = e % nf;
      double t = 0.0;
      for (int s2 = 0; s2 <= q; s2++) t += sT[q * MM + s2] * sv[s2 * AMAXF + f];   // T(s2, q), upper triangular
      st2[q * AMAXF + f] = t;
    }
    // u1 = R^-T g (forward substitution with R^T, lower triangular)
    if (tid < nf) {
      const int f = tid;
      for (int k = 0; k < m; k++) {
        double t = su1[k * AMAXF + f];
        for (int j = 0; j < k; j++) t -= sR[k * MM + j] * su1[j * AMAXF + f];   // R(j, k) u1_j
        su1[k * AMAXF + f] = t / sR[k * MM + k];
      }
    }
    __syncthreads();
    for (int i = tid; i < n; i += ANT) {
      for (int f = 0; f < nf; f++) {
        double h = r[(size_t)f * ldv + i];
        for (int q = 0; q < m; q++) h -= V[(size_t)q * n + i] * st2[q * AMAXF + f];
        if (i < m) {
          sh1[i * AMAXF + f] = h;
        } else {
          for (int k = 0; k < m; k++) h -= B1[(size_t)k * n + i] * su1[k * AMAXF + f];
        }
        y[(size_t)f * ldv + i] = h;
      }
    }
    __syncthreads();
    // u2 = B22^-1 (h2 - B21 u1)
    if (phase == 1) {
      double* hs = hstate + (size_t)blockIdx.x * 2 * MF;
      for (int e = tid; e < MF; e += ANT) { hs[e] = su1[e]; hs[MF + e] = sh1[e]; }
      return;
    }
    {
      const int nbp = a.fnb ? a.fnb[p] : 0;
      if (nbp == SOLVE_NB) llt_solve_blocked<TL, ANT>(L, nn, y + m, ldv, nf, zb, lb);
      else llt_solve<TL>(L, nn, y + m, ldv, nf, red, zk);
    }
  after_b22:
    // beta correction: R^-1 (h1 - B11 u1 - B12 u2)
    col_dots(B1, n, m, m, y, ldv, nf, red, sv);
    if (tid < nf) {
      const int f = tid;
      for (int q = 0; q < m; q++) {
        double t = sh1[q * AMAXF + f] - sv[q * AMAXF + f];
        for (int k = 0; k < m; k++) t -= sB11[k * MM + q] * su1[k * AMAXF + f];   // B(q, k)
        sdb[q * AMAXF + f] = t;
      }
      for (int k = m - 1; k >= 0; k--) {
        double t = sdb[k * AMAXF + f];
        for (int j = k + 1; j < m; j++) t -= sR[j * MM + k] * sdb[j * AMAXF + f];   // R(k, j)
        sdb[k * AMAXF + f] = t / sR[k * MM + k];
      }
    }
    for (int e = tid; e < m * nf; e += ANT) y[(size_t)(e % nf) * ldv + e / nf] = su1[(e / nf) * AMAXF + e % nf];
    __syncthreads();
    // dc = Q u = u - V T (V^T u)
    col_dots(V, n, 0, m, y, ldv, nf, red, sv);
    for (int e = tid; e < m * nf; e += ANT) {
      const int q = e / nf, f = e % nf;
      double t = 0.0;
      for (int s2 = q; s2 < m; s2++) t += sT[s2 * MM + q] * sv[s2 * AMAXF + f];   // T(q, s2)
      st2[q * AMAXF + f] = t;
    }
    __syncthreads();
    double dmax = 0.0, xmax = 0.0;
    for (int i = tid; i < n; i += ANT)
      for (int f = 0; f < nf; f++) {
        double d = y[(size_t)f * ldv + i];
        for (int q = 0; q < m; q++) d -= V[(size_t)q * n + i] * st2[q * AMAXF + f];
        const double xn = c[(size_t)f * ldv + i] + d;
        c[(size_t)f * ldv + i] = xn;
        dmax = fmax(dmax, fabs(d));
        xmax = fmax(xmax, fabs(xn));
      }
    if (tid < nf)
      for (int q = 0; q < m; q++) {
        const double d = sdb[q * AMAXF + tid];
        const double xn = sb[q * AMAXF + tid] + d;
        sb[q * AMAXF + tid] = xn;
        dmax = fmax(dmax, fabs(d));
        xmax = fmax(xmax, fabs(xn));
      }
    {
      double mx[2] = {dmax, xmax};
      for (int o = AW / 2; o > 0; o >>= 1) {
        mx[0] = fmax(mx[0], __shfl_xor(mx[0], o));
        mx[1] = fmax(mx[1], __shfl_xor(mx[1], o));
      }
      if ((tid & (AW - 1)) == 0) { red[(tid / AW) * 2] = mx[0]; red[(tid / AW) * 2 + 1] = mx[1]; }
      __syncthreads();
      dmax = 0.0; xmax = 0.0;
      for (int w = 0; w < ANT / AW; w++) { dmax = fmax(dmax, red[w * 2]); xmax = fmax(xmax, red[w * 2 + 1]); }
      __syncthreads();
    }
    if (!refine) { converged = 1; break; }
    if (it > 0 && dmax <= tol * xmax) { converged = 1; break; }
    // stagnated at the FP64 rounding floor of the residual (about cond * 1e-16 of
    // |x|, which the FP64 factor's solution shares): accepted below 1e-8 |x|; a
    // factor too inaccurate to refine (cond beyond 1 / u32) never shrinks its
    // corrections that far and falls back to FP64
    if (it > 1 && dmax > 0.5 * dprev && dmax <= fmax(100.0 * tol, 1e-8) * xmax) { converged = 1; break; }
    dprev = dmax;
  }
  double* X = coef + mem_off[p] * nf;
  for (int i = tid; i < n; i += ANT)
    for (int f = 0; f < nf; f++) X[(size_t)i * nf + f] = c[(size_t)f * ldv + i];
  for (int e = tid; e < m * nf; e += ANT) bt[e] = sb[(e / nf) * AMAXF + e % nf];
  if (tid == 0) {
    if (stepping && !converged && step_it < max_iter) {
      status[p] = 2;   // another iteration follows
      dstate[p] = dprev;
    } else {
      if (status) status[p] = converged ? 0 : 1;
      if (iters) iters[p] = it;
    }
  }
}

// Residual r = f - A c - P beta of the saddle-point systems of the large patches
// during a stepped FP32 refinement (k_solve_aug, step_it >= 1): one block of ANT
// rows per task (patch, row block), so the O(n^2) matrix-free product of a patch of
// thousands of members spreads over many compute units.  The columns come through
// LDS in chunks of ANT; each row sums them in ascending column order with the same
// operations as the in-kernel residual, so the values are the same.
template <int KT, int MM>
__global__ __launch_bounds__(ANT) void k_aug_resid_big(Grid g, const int2* __restrict__ tasks,
                                                       const int32_t* __restrict__ pn,
                                                       const int64_t* __restrict__ mem_off,
                                                       const int32_t* __restrict__ mem_pool,
                                                       const double* __restrict__ values, int nf,
                                                       const int32_t* __restrict__ status, double* __restrict__ ws,
                                                       AugView a) {
  __shared__ double sx[3][ANT];
  __shared__ double sc[AMAXF][ANT];
  const int2 tk = tasks[blockIdx.x];
  const int p = tk.x;
  if (status[p] != 2) return;
  const int n = pn[p], m = a.m[p], deg = a.deg[p], dim = g.dim;
  const int tid = threadIdx.x;
  const int i = tk.y * ANT + tid;
  const bool live = i < n;
  const int ii = live ? i : n - 1;
  const size_t ldv = (size_t)n;
  double* xs = ws + (size_t)mem_off[p] * (3 + 3 * (size_t)nf);
  const double* c = xs + 3 * ldv;
  double* r = xs + (3 + 2 * (size_t)nf) * ldv;
  const int32_t* mem = mem_pool + mem_off[p];
  const double eps = a.eps ? a.eps[p] : g.eps_r;
  const double* bt = a.beta + (size_t)p * a.bstride;
  double s[AMAXF];
#pragma unroll
  for (int f = 0; f < AMAXF; f++) s[f] = f < nf ? values[(size_t)mem[ii] * nf + f] : 0.0;
  const double xi0 = xs[ii], xi1 = xs[ldv + ii], xi2 = xs[2 * ldv + ii];
  for (int j0 = 0; j0 < n; j0 += ANT) {
    const int jn = min(ANT, n - j0);
    __syncthreads();
    if (tid < jn) {
      sx[0][tid] = xs[j0 + tid]; sx[1][tid] = xs[ldv + j0 + tid]; sx[2][tid] = xs[2 * ldv + j0 + tid];
#pragma unroll
      for (int f = 0; f < AMAXF; f++)
        if (f < nf) sc[f][tid] = c[(size_t)f * ldv + j0 + tid];
    }
    __syncthreads();
    for (int jj = 0; jj < jn; jj++) {
      const double d0 = xi0 - sx[0][jj], d1 = xi1 - sx[1][jj], d2 = xi2 - sx[2][jj];
      const double phi = rbf<KT, double>(sqrt(d0 * d0 + d1 * d1 + d2 * d2), eps);
#pragma unroll
      for (int f = 0; f < AMAXF; f++)
        if (f < nf) s[f] -= phi * sc[f][jj];
    }
  }
  double xi[3] = {xi0 * a.ps, xi1 * a.ps, xi2 * a.ps}, pv[POLY_MAX];
  poly_eval(dim, deg, xi, pv);
#pragma unroll
  for (int q = 0; q < MM; q++)
    if (q < m)
#pragma unroll
      for (int f = 0; f < AMAXF; f++)
        if (f < nf) s[f] -= pv[q] * bt[q * nf + f];
  if (live) {
#pragma unroll
    for (int f = 0; f < AMAXF; f++)
      if (f < nf) r[(size_t)f * ldv + i] = s[f];
  }
}

__global__ void k_build_ptab(int P, const uint8_t* __restrict__ active, AugView a, int mm, int nft, int nf, int f0,
                             double* __restrict__ ptab) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)P * mm * nft) return;
  const int p = (int)(e / (mm * nft)), q = (int)((e / nft) % mm), f = (int)(e % nft);
  const int m = a.m ? a.m[p] : 0;
  ptab[e] = (active[p] && q < m) ? a.beta[(size_t)p * a.bstride + (size_t)q * nf + f0 + f] : 0.0;
}

// ---------------------------------------------------------------------------
template <int KT>
static void prep_k(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                   const int32_t* mem_pool, const double* xd, const AugView& a, const int2* tasks, int ntasks,
                   int stage, hipStream_t s) {
  if (a.mmax <= 4) hipLaunchKernelGGL((k_aug_prep<KT, 4>), dim3(nlist), dim3(ANT), 0, s, g, list, pn, mem_off, mem_pool, xd, a, stage);
  else hipLaunchKernelGGL((k_aug_prep<KT, POLY_MAX>), dim3(nlist), dim3(ANT), 0, s, g, list, pn, mem_off, mem_pool, xd, a, stage);
  if (stage == 1 && ntasks > 0) {
    if (a.mmax <= 4) hipLaunchKernelGGL((k_aug_av_big<KT, 4>), dim3(ntasks), dim3(ANT), 0, s, g, tasks, pn, a);
    else hipLaunchKernelGGL((k_aug_av_big<KT, POLY_MAX>), dim3(ntasks), dim3(ANT), 0, s, g, tasks, pn, a);
    if (a.mmax <= 4) hipLaunchKernelGGL((k_aug_prep<KT, 4>), dim3(nlist), dim3(ANT), 0, s, g, list, pn, mem_off, mem_pool, xd, a, 2);
    else hipLaunchKernelGGL((k_aug_prep<KT, POLY_MAX>), dim3(nlist), dim3(ANT), 0, s, g, list, pn, mem_off, mem_pool, xd, a, 2);
  }
}

void launch_aug_prep(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                     const int64_t* mem_off, const int32_t* mem_pool, const double* xd, const AugView& a, hipStream_t s) {
  launch_aug_prep_tasks(g, ktype, list, nlist, pn, mem_off, mem_pool, xd, a, nullptr, 0, s);
}

void launch_aug_prep_tasks(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                           const int64_t* mem_off, const int32_t* mem_pool, const double* xd, const AugView& a,
                           const int2* tasks, int ntasks, hipStream_t s) {
  if (nlist <= 0) return;
  const int stage = tasks ? 1 : 0;
#define LRB_PK(K) prep_k<K>(g, list, nlist, pn, mem_off, mem_pool, xd, a, tasks, ntasks, stage, s)
  switch (ktype) {
    case K_WC0: LRB_PK(K_WC0); break;
    case K_WC2: LRB_PK(K_WC2); break;
    case K_IMQ: LRB_PK(K_IMQ); break;
    case K_PHS3: LRB_PK(K_PHS3); break;
    default: LRB_PK(K_GAUSS); break;
  }
#undef LRB_PK
}

// largest patch whose vectors sit in LDS: every patch of the small-patch methods
// (factor dimension n - m <= 512)
static constexpr int AUG_LDS_MAX = 512 + POLY_MAX;
static size_t aug_solve_smem(int S, int nf) { return sizeof(double) * (size_t)S * (3 + 3 * (size_t)nf); }
size_t aug_solve_ws_per_member(int nf) { return 3 + 3 * (size_t)nf; }

// a stepped solve keeps every vector in the workspace
static size_t stepping_smem(int S, int nf, int step_it) { return step_it >= 0 ? 0 : aug_solve_smem(S, nf); }

template <typename TL, int KT, int MM>
static void solve_aug_k(const Grid& g, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                        const int64_t* fac_off, const int32_t* mem_pool, const double* xd, const TL* fac,
                        const double* values, int nf, double* coef, int max_iter, double tol, int32_t* status,
                        int32_t* iters, double* ws, int nmax, const AugView& a, int step_it, double* dstate,
                        int phase, double* hstate, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_solve_aug<TL, KT, MM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)aug_solve_smem(AUG_LDS_MAX, AMAXF));
    attr = true;
  }
  const int S = std::max(1, std::min(nmax, AUG_LDS_MAX));
  hipLaunchKernelGGL((k_solve_aug<TL, KT, MM>), dim3(nlist), dim3(ANT), stepping_smem(S, nf, step_it), s, g, list, pn,
                     mem_off, fac_off, mem_pool, xd, fac, values, nf, coef, max_iter, tol, status, iters, ws, S, a,
                     step_it, dstate, phase, hstate);
}

template <typename TL, int MM>
static void solve_aug_mm(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                         const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                         const TL* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                         int32_t* status, int32_t* iters, double* ws, int nmax, const AugView& a, int step_it,
                         double* dstate, int phase, double* hstate, hipStream_t s) {
#define LRB_SA(K) solve_aug_k<TL, K, MM>(g, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, a, step_it, dstate, phase, hstate, s)
  switch (ktype) {
    case K_WC0: LRB_SA(K_WC0); break;
    case K_WC2: LRB_SA(K_WC2); break;
    case K_IMQ: LRB_SA(K_IMQ); break;
    case K_PHS3: LRB_SA(K_PHS3); break;
    default: LRB_SA(K_GAUSS); break;
  }
#undef LRB_SA
}

void launch_solve_aug(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                      const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                      const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                      int32_t* status, int32_t* iters, double* ws, int nmax, const AugView& a, hipStream_t s) {
  if (nlist <= 0 || nf <= 0) return;
  launch_solve_aug_step(g, ktype, fp32, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, values, nf, coef, max_iter,
                        tol, status, iters, ws, nmax, a, -1, nullptr, 0, nullptr, s);
}

void launch_solve_aug_step(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                           const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                           const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                           int32_t* status, int32_t* iters, double* ws, int nmax, const AugView& a, int step_it,
                           double* dstate, int phase, double* hstate, hipStream_t s) {
  if (nlist <= 0 || nf <= 0) return;
  if (fp32) {
    if (a.mmax <= 4) solve_aug_mm<float, 4>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (const float*)fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, a, step_it, dstate, phase, hstate, s);
    else solve_aug_mm<float, POLY_MAX>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (const float*)fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, a, step_it, dstate, phase, hstate, s);
  } else {
    if (a.mmax <= 4) solve_aug_mm<double, 4>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (const double*)fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, a, step_it, dstate, phase, hstate, s);
    else solve_aug_mm<double, POLY_MAX>(g, ktype, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (const double*)fac, values, nf, coef, max_iter, tol, status, iters, ws, nmax, a, step_it, dstate, phase, hstate, s);
  }
}

template <int KT, int MM>
static void resid_big_k(const Grid& g, const int2* tasks, int ntasks, const int32_t* pn, const int64_t* mem_off,
                        const int32_t* mem_pool, const double* values, int nf, const int32_t* status, double* ws,
                        const AugView& a, hipStream_t s) {
  hipLaunchKernelGGL((k_aug_resid_big<KT, MM>), dim3(ntasks), dim3(ANT), 0, s, g, tasks, pn, mem_off, mem_pool, values,
                     nf, status, ws, a);
}

template <int MM>
static void resid_big_mm(const Grid& g, int ktype, const int2* tasks, int ntasks, const int32_t* pn,
                         const int64_t* mem_off, const int32_t* mem_pool, const double* values, int nf,
                         const int32_t* status, double* ws, const AugView& a, hipStream_t s) {
#define LRB_RB(K) resid_big_k<K, MM>(g, tasks, ntasks, pn, mem_off, mem_pool, values, nf, status, ws, a, s)
  switch (ktype) {
    case K_WC0: LRB_RB(K_WC0); break;
    case K_WC2: LRB_RB(K_WC2); break;
    case K_IMQ: LRB_RB(K_IMQ); break;
    case K_PHS3: LRB_RB(K_PHS3); break;
    default: LRB_RB(K_GAUSS); break;
  }
#undef LRB_RB
}

int aug_resid_rows() { return ANT; }

void launch_aug_resid_big(const Grid& g, int ktype, const int2* tasks, int ntasks, const int32_t* pn,
                          const int64_t* mem_off, const int32_t* mem_pool, const double* values, int nf,
                          const int32_t* status, double* ws, const AugView& a, hipStream_t s) {
  if (ntasks <= 0 || nf <= 0) return;
  if (a.mmax <= 4) resid_big_mm<4>(g, ktype, tasks, ntasks, pn, mem_off, mem_pool, values, nf, status, ws, a, s);
  else resid_big_mm<POLY_MAX>(g, ktype, tasks, ntasks, pn, mem_off, mem_pool, values, nf, status, ws, a, s);
}

void launch_build_ptab(const Grid& g, const uint8_t* active, const AugView& a, int mm, int nft, int nf, int f0,
                       double* ptab, hipStream_t s) {
  const int64_t tot = (int64_t)g.P * mm * nft;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_build_ptab, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g.P, active, a, mm, nft, nf, f0,
                     ptab);
}

}  // namespace lrb
