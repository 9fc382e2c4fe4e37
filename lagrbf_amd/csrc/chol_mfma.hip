// Batched Cholesky of the patch systems with the panel products on the matrix
// cores (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  One workgroup (4
// waves) per patch, factor dimension up to 512, left-looking by panels of 16
// columns, the panel held in the MFMA accumulators (not in LDS, so that a
// workgroup needs only its coordinates and two 16 x 16 blocks of LDS and several
// patches share a compute unit):
//
//   1. each wave accumulates the transposed tiles of its 16-row strips s = wave,
//      wave + 4, ... of the panel,
//        D[c][r] = sum_k L[J + c][k] L[r][k]  (+ sum_q V[J+c][q] X[r][q] + X[J+c][q] V[r][q]),
//      the saddle-point correction X V^T + V X^T of a polynomial block as 2m more
//      columns of the same product, and subtracts them from the matrix entries
//      assembled on the fly (phi of the member distances);
//   2. wave 0 factors the 16 x 16 diagonal block and inverts the factor in
//      registers (lane broadcasts, no LDS round trip on the serial path);
//   3. a lane holds column r of its tile, which is the B operand layout of
//      L_below^T = L_JJ^-1 P_below^T: the triangular solve of each strip is 4 MFMAs
//      on the accumulators as they are, the result written to the factor slot.
//
// The factor slot holds L column-major with leading dimension n (the lower
// triangle), as the vector-ALU kernels of patch_chol.hip write it, so the solve
// kernels read either; the upper triangle of each diagonal block holds the
// strictly lower part of its inverse for the blocked solves (blocked_solve.hpp).
// A non-positive pivot sets status[p] to its column + 1.
#include "kernels.hpp"
#include "mfma_chol.hpp"

#include <type_traits>

namespace lrb {

static constexpr int MNT = 256;      // threads per patch
static constexpr int MNMAX = 512;    // largest factor dimension

__device__ __forceinline__ double centre_coord(const Grid& g, int p, int d) {
  const int idx = d == 0 ? p % g.n[0] : (d == 1 ? (p / g.n[0]) % g.n[1] : p / (g.n[0] * g.n[1]));
  return g.O[d] + idx * g.h[d];
}

// Panel accumulation and assembly for the first NA strips of a wave (NA known at
// compile time, so every MFMA and accumulator index is static).
template <typename T, typename AT, int KT, int MM, int SPW, int NA>
__device__ __forceinline__ void rp_panel(typename Mfma<T>::V (&acc)[SPW], const T* __restrict__ L, int n, int n0,
                                         int J, int wv, int lane, const AT* __restrict__ xr, AT eps,
                                         const double* __restrict__ Vg, const double* __restrict__ Xg, int m) {
  using M = Mfma<T>;
  using V = typename M::V;
  constexpr int NMAX = 64 * SPW;
  const int cl = lane & 15, kl = lane >> 4;
  const int ca = J + cl < n ? J + cl : n - 1;   // panel column of this lane's A operand (clamped)
  int rs[NA];
#pragma unroll
  for (int t = 0; t < NA; t++) {
    rs[t] = J + 16 * (wv + 4 * t) + cl;
    acc[t] = V{};
  }
  // L[J:J+16, 0:J] L[rows, 0:J]^T, 4 KU columns of k per iteration with every load of
  // the iteration issued before its MFMAs (J is a multiple of 16); fewer for many
  // FP64 strips, whose operands would otherwise cost the second wave per SIMD
  constexpr int KU = (sizeof(T) == 8 && NA > 4) ? 2 : 4;
  for (int k = 0; k < J; k += 4 * KU) {
    T a[KU], b[KU][NA];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const T* Lk = L + (size_t)(k + 4 * u + kl) * n;
      a[u] = Lk[ca];
#pragma unroll
      for (int t = 0; t < NA; t++) b[u][t] = rs[t] < n ? Lk[rs[t]] : T(0);
    }
#pragma unroll
    for (int u = 0; u < KU; u++)
#pragma unroll
      for (int t = 0; t < NA; t++) acc[t] = M::mma(a[u], b[u][t], acc[t]);
  }
  if constexpr (MM > 0) {
    // X V^T + V X^T: columns q < m pair V of the panel with X of the rows, columns
    // m <= q < 2m X with V
#pragma unroll
    for (int k = 0; k < 2 * MM; k += 4) {
      const int q = k + kl;
      const bool qa = q < 2 * m;
      const int qq = q < m ? q : q - m;
      const double* A = q < m ? Vg : Xg;
      const double* B = q < m ? Xg : Vg;
      const T a = qa ? (T)A[(size_t)qq * n0 + ca] : T(0);
#pragma unroll
      for (int t = 0; t < NA; t++) {
        const T b = (qa && rs[t] < n) ? (T)B[(size_t)qq * n0 + rs[t]] : T(0);
        acc[t] = M::mma(a, b, acc[t]);
      }
    }
  }
  // panel = assembled matrix - accumulated products; identity beyond n so that the
  // last diagonal block factors as a 16 x 16 one
  AT xc[4][3];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int c = min(J + M::row(lane, v), n - 1);
#pragma unroll
    for (int d = 0; d < 3; d++) xc[v][d] = xr[d * NMAX + c];
  }
#pragma unroll
  for (int t = 0; t < NA; t++) {
    const int r = rs[t], rr = min(r, n - 1);
    const AT x0 = xr[rr], x1 = xr[NMAX + rr], x2 = xr[2 * NMAX + rr];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int c = J + M::row(lane, v);
      const AT d0 = x0 - xc[v][0], d1 = x1 - xc[v][1], d2 = x2 - xc[v][2];
      const T av = (T)rbf<KT, AT>(sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2), eps);
      acc[t][v] = (r < n && c < n) ? av - acc[t][v] : (r == c ? T(1) : T(0));
    }
  }
}

// L_below^T = L_JJ^-1 P_below^T for the first NA strips of a wave (strip 0, the
// diagonal block, skipped), written column-major.
template <typename T, int SPW, int NA>
__device__ __forceinline__ void rp_trsm(const typename Mfma<T>::V (&acc)[SPW], T* __restrict__ L, int n, int J,
                                        int wv, int lane, const T* __restrict__ Li) {
  using M = Mfma<T>;
  using V = typename M::V;
  constexpr int PLD = 17;
  const int cl = lane & 15;
  T lia[4];
#pragma unroll
  for (int u = 0; u < 4; u++) lia[u] = Li[cl * PLD + M::row(lane, u)];
#pragma unroll
  for (int t = 0; t < NA; t++) {
    if (t == 0 && wv == 0) continue;
    V o = V{};
#pragma unroll
    for (int u = 0; u < 4; u++) o = M::mma(lia[u], acc[t][u], o);
    const int r = J + 16 * (wv + 4 * t) + cl;
    if (r < n) {
#pragma unroll
      for (int v = 0; v < 4; v++) L[(size_t)(J + M::row(lane, v)) * n + r] = o[v];
    }
  }
}

template <typename T, int KT, int MM, int SPW>
__global__ __launch_bounds__(MNT) void k_chol_rp(Grid g, const int32_t* __restrict__ list,
                                                 const int32_t* __restrict__ pn, const int64_t* __restrict__ mem_off,
                                                 const int64_t* __restrict__ fac_off,
                                                 const int32_t* __restrict__ mem_pool, const double* __restrict__ xd,
                                                 T* __restrict__ fac, int32_t* __restrict__ status, AugView aug) {
  using M = Mfma<T>;
  using V = typename M::V;
  using AT = typename std::conditional<(MM > 0), double, T>::type;
  constexpr int NMAX = 64 * SPW, PLD = 17;
  __shared__ AT xr[3 * NMAX];
  __shared__ T Pd[16 * PLD];   // diagonal block, row-major, factored in place
  __shared__ T Li[16 * PLD];   // inverse of its factor, row-major
  __shared__ int s_fail;
  const int p = list[blockIdx.x];
  const int n0 = pn[p];
  const int m = MM > 0 ? aug.m[p] : 0;
  const int n = n0 - m;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (n <= 0) { if (tid == 0) status[p] = 0; return; }
  const int dim = g.dim;
  const int32_t* mem = mem_pool + mem_off[p] + m;
  T* L = fac + fac_off[p];
  const double* Vg = MM > 0 ? aug.aux + aug.aux_off[p] + 3 * (size_t)n0 + m : nullptr;   // V[q * n0 + i]
  const double* Xg = MM > 0 ? Vg + (size_t)m * n0 : nullptr;
  if (tid == 0) s_fail = 0;
  for (int i = tid; i < n; i += MNT) {
    const double* x = xd + (size_t)mem[i] * dim;
#pragma unroll
    for (int d = 0; d < 3; d++) xr[d * NMAX + i] = d < dim ? (AT)(x[d] - centre_coord(g, p, d)) : AT(0);
  }
  __syncthreads();
  const AT eps = (AT)(aug.eps ? aug.eps[p] : g.eps_r);
  const int cl = lane & 15;
  for (int J = 0; J < n; J += 16) {
    const int w = min(16, n - J);
    const int nstrip = (n - J + 15) / 16;
    // strips of this wave in this panel: wv, wv + 4, ... below nstrip
    const int na = nstrip > wv ? min(SPW, (nstrip - wv + 3) / 4) : 0;
    V acc[SPW];
#define LRB_RPP(NA) rp_panel<T, AT, KT, MM, SPW, NA>(acc, L, n, n0, J, wv, lane, xr, eps, Vg, Xg, m)
    switch (na) {
      case 1: LRB_RPP(1); break;
      case 2: if constexpr (SPW >= 2) LRB_RPP(2); break;
      case 3: if constexpr (SPW >= 3) LRB_RPP(3); break;
      case 4: if constexpr (SPW >= 4) LRB_RPP(4); break;
      case 5: if constexpr (SPW >= 5) LRB_RPP(5); break;
      case 6: if constexpr (SPW >= 6) LRB_RPP(6); break;
      case 7: if constexpr (SPW >= 7) LRB_RPP(7); break;
      case 8: if constexpr (SPW >= 8) LRB_RPP(8); break;
      default: break;
    }
#undef LRB_RPP
    // wave 0 holds strip 0, the diagonal block: factor it and invert the factor
    if (wv == 0) diag_block_rp<T>(acc[0], L, n, J, w, lane, Pd, Li, &s_fail);
    __syncthreads();
#define LRB_RPT(NA) rp_trsm<T, SPW, NA>(acc, L, n, J, wv, lane, Li)
    switch (na) {
      case 1: LRB_RPT(1); break;
      case 2: if constexpr (SPW >= 2) LRB_RPT(2); break;
      case 3: if constexpr (SPW >= 3) LRB_RPT(3); break;
      case 4: if constexpr (SPW >= 4) LRB_RPT(4); break;
      case 5: if constexpr (SPW >= 5) LRB_RPT(5); break;
      case 6: if constexpr (SPW >= 6) LRB_RPT(6); break;
      case 7: if constexpr (SPW >= 7) LRB_RPT(7); break;
      case 8: if constexpr (SPW >= 8) LRB_RPT(8); break;
      default: break;
    }
#undef LRB_RPT
    __syncthreads();
  }
  if (tid == 0) status[p] = s_fail;
}

template <typename T, int KT, int MM>
static void chol_rp_k(const Grid& g, int nmax, const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                      const int64_t* fac_off, const int32_t* mem_pool, const double* xd, T* fac, int32_t* status,
                      const AugView& a, hipStream_t s) {
#define LRB_RP(SPW) hipLaunchKernelGGL((k_chol_rp<T, KT, MM, SPW>), dim3(nlist), dim3(MNT), 0, s, g, list, pn, mem_off, \
                                       fac_off, mem_pool, xd, fac, status, a)
  if (nmax <= 128) LRB_RP(2);
  else if (nmax <= 256) LRB_RP(4);
  else if (nmax <= 320) LRB_RP(5);
  else if (nmax <= 384) LRB_RP(6);
  else LRB_RP(8);
#undef LRB_RP
}

template <typename T, int MM>
static void chol_rp_kt(const Grid& g, int ktype, int nmax, const int32_t* list, int nlist, const int32_t* pn,
                       const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                       T* fac, int32_t* status, const AugView& a, hipStream_t s) {
#define LRB_CR(K) chol_rp_k<T, K, MM>(g, nmax, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s)
  switch (ktype) {
    case K_WC0: LRB_CR(K_WC0); break;
    case K_WC2: LRB_CR(K_WC2); break;
    case K_IMQ: LRB_CR(K_IMQ); break;
    case K_PHS3:
      if constexpr (MM > 0) LRB_CR(K_PHS3);
      break;
    default: LRB_CR(K_GAUSS); break;
  }
#undef LRB_CR
}

template <typename T>
static void chol_rp_mm(const Grid& g, int ktype, int nmax, const int32_t* list, int nlist, const int32_t* pn,
                       const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                       T* fac, int32_t* status, const AugView& a, hipStream_t s) {
  if (!a.m) chol_rp_kt<T, 0>(g, ktype, nmax, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s);
  else if (a.mmax <= 4) chol_rp_kt<T, 4>(g, ktype, nmax, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s);
  else chol_rp_kt<T, POLY_MAX>(g, ktype, nmax, list, nlist, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s);
}

int chol_mfma_max_n() { return MNMAX; }

void launch_factor_chol_mfma(const Grid& g, int ktype, bool fp32, int nmax, const int32_t* list, int nlist,
                             const int32_t* pn, const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool,
                             const double* xd, void* fac, int32_t* status, const AugView& a, hipStream_t s) {
  if (nlist <= 0) return;
  if (fp32) chol_rp_mm<float>(g, ktype, nmax, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (float*)fac, status, a, s);
  else chol_rp_mm<double>(g, ktype, nmax, list, nlist, pn, mem_off, fac_off, mem_pool, xd, (double*)fac, status, a, s);
}

}  // namespace lrb
