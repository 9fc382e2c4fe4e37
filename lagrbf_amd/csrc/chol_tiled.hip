// Cholesky of the patch systems too large for one workgroup (factor dimension
// above the 512 of chol_mfma.hip): clustered nodes put thousands of members into
// one PU patch, and a single compute unit would stream n^3 / 6 elements of the
// factor through its own cache.  Here each patch is cut into tiles of 64 columns
// spread over many workgroups, the products on the matrix cores
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  Left-looking by tile columns:
// tile column k (columns J = 64 k .. J + 63) is two launches over every listed
// patch with more than k tiles, the (patch, row tile) pairs of the step listed by
// the host:
//
//   k_tile_update: one workgroup per row tile s >= k (rows R = 64 s ..).  Its four
//     waves take one 16-row strip each and accumulate the transposed products
//       D[c][r] = sum_{t < J} L[c][t] L[r][t]  (+ X V^T + V X^T of a polynomial block)
//     of the tile's 64 columns, subtract them from the matrix entries assembled on
//     the fly (phi of the member distances, FP64 before rounding with a polynomial
//     block), and write the updated tile to the factor slot.  The workgroup of the
//     diagonal tile (s == k) factors its tile instead: each 16-column block is
//     factored and inverted in one wave's registers (diag_block_rp), the strips
//     below it solved (L^-1 P^T, four MFMAs on the accumulators as they are) and
//     the later blocks updated, the solved blocks passed through LDS.
//   k_tile_solve: one workgroup per row tile s > k, L_sk = P_sk L_kk^-T by 16-column
//     blocks with the inverted diagonal blocks the update left in the factor slot;
//     no barriers (every operand is final).
//
// The factor slot layout is that of the single-workgroup kernels: L column-major
// with leading dimension n (lower triangle), the upper triangle of every 16 x 16
// diagonal block holding the strictly lower part of its inverse, so the blocked
// solves (blocked_solve.hpp) read it unchanged.  Rows and columns past n are the
// identity (the last tile factors as a full one); nothing past n is written.  A
// non-positive pivot sets status[p] to its column + 1 (the first one).
#include "kernels.hpp"
#include "mfma_chol.hpp"

#include <type_traits>

namespace lrb {

static constexpr int TNT = 256;   // threads per tile: four waves, one 16-row strip each

__device__ __forceinline__ double tile_centre(const Grid& g, int p, int d) {
  const int idx = d == 0 ? p % g.n[0] : (d == 1 ? (p / g.n[0]) % g.n[1] : p / (g.n[0] * g.n[1]));
  return g.O[d] + idx * g.h[d];
}

template <typename T, int KT, int MM>
__global__ __launch_bounds__(TNT) void k_tile_update(Grid g, const int2* __restrict__ tasks, int k,
                                                     const int32_t* __restrict__ pn,
                                                     const int64_t* __restrict__ mem_off,
                                                     const int64_t* __restrict__ fac_off,
                                                     const int32_t* __restrict__ mem_pool,
                                                     const double* __restrict__ xd, T* __restrict__ fac,
                                                     int32_t* __restrict__ status, AugView aug) {
  using M = Mfma<T>;
  using V = typename M::V;
  using AT = typename std::conditional<(MM > 0), double, T>::type;
  constexpr int PLD = 17;
  __shared__ AT cx[3][TILE_NB], rx[3][TILE_NB];   // coordinates of the tile's columns / rows
  __shared__ T Pd[16 * PLD], Li[16 * PLD];        // diagonal block and its inverse (diag_block_rp)
  __shared__ T Lx[4][16 * PLD];                   // solved blocks of the diagonal tile's strips
  __shared__ int s_fail;
  const int2 tk = tasks[blockIdx.x];
  const int p = tk.x, s = tk.y;
  const int n0 = pn[p];
  const int m = MM > 0 ? aug.m[p] : 0;
  const int n = n0 - m;
  const int J = TILE_NB * k, R = TILE_NB * s;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, cl = lane & 15, kl = lane >> 4;
  T* L = fac + fac_off[p];
  const int32_t* mem = mem_pool + mem_off[p] + m;   // factor row i = member m + i
  const int dim = g.dim;
  if (tid < 2 * TILE_NB) {
    const bool col = tid < TILE_NB;
    const int i = tid & (TILE_NB - 1);
    const int gi = min((col ? J : R) + i, n - 1);
    const double* x = xd + (size_t)mem[gi] * dim;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const AT v = d < dim ? (AT)(x[d] - tile_centre(g, p, d)) : AT(0);
      if (col) cx[d][i] = v;
      else rx[d][i] = v;
    }
  }
  if (tid == 0) s_fail = 0;
  __syncthreads();

  const int r = R + 16 * wv + cl;   // row of L of this lane's results
  const bool rok = r < n;
  int ca[4];
#pragma unroll
  for (int cb = 0; cb < 4; cb++) ca[cb] = min(J + 16 * cb + cl, n - 1);
  V acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; cb++) acc[cb] = V{};
  // L[J:J+64, 0:J] L[R:R+64, 0:J]^T: every load of an iteration issued before its MFMAs
  constexpr int KU = 2;
  for (int t = 0; t < J; t += 4 * KU) {
    T a[KU][4], b[KU];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const T* Lt = L + (size_t)(t + 4 * u + kl) * n;
#pragma unroll
      for (int cb = 0; cb < 4; cb++) a[u][cb] = Lt[ca[cb]];
      b[u] = rok ? Lt[r] : T(0);
    }
#pragma unroll
    for (int u = 0; u < KU; u++)
#pragma unroll
      for (int cb = 0; cb < 4; cb++) acc[cb] = M::mma(a[u][cb], b[u], acc[cb]);
  }
  if constexpr (MM > 0) {
    // X V^T + V X^T: columns q < m pair V of the tile's columns with X of its rows,
    // columns m <= q < 2m X with V
    const double* Vg = aug.aux + aug.aux_off[p] + 3 * (size_t)n0 + m;   // V[q * n0 + i]
    const double* Xg = Vg + (size_t)m * n0;
#pragma unroll
    for (int kq = 0; kq < 2 * MM; kq += 4) {
      const int q = kq + kl;
      const bool qa = q < 2 * m;
      const int qq = q < m ? q : q - m;
      const double* A = q < m ? Vg : Xg;
      const double* B = q < m ? Xg : Vg;
      const T bq = (qa && rok) ? (T)B[(size_t)qq * n0 + r] : T(0);
#pragma unroll
      for (int cb = 0; cb < 4; cb++) {
        const T aq = qa ? (T)A[(size_t)qq * n0 + ca[cb]] : T(0);
        acc[cb] = M::mma(aq, bq, acc[cb]);
      }
    }
  }
  // tile = assembled matrix - accumulated products; the identity past n
  const AT eps = (AT)(aug.eps ? aug.eps[p] : g.eps_r);
  {
    const int ri = 16 * wv + cl;
    const AT x0 = rx[0][ri], x1 = rx[1][ri], x2 = rx[2][ri];
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int ci = 16 * cb + M::row(lane, v);
        const int c = J + ci;
        const AT d0 = x0 - cx[0][ci], d1 = x1 - cx[1][ci], d2 = x2 - cx[2][ci];
        const T av = (T)rbf<KT, AT>(sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2), eps);
        acc[cb][v] = (rok && c < n) ? av - acc[cb][v] : (r == c ? T(1) : T(0));
      }
  }
  if (s != k) {
    // a tile below the diagonal: its columns are solved by k_tile_solve
    if (rok) {
#pragma unroll
      for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = J + 16 * cb + M::row(lane, v);
          if (c < n) L[(size_t)c * n + r] = acc[cb][v];
        }
    }
    return;
  }
  // the diagonal tile: factor it by blocks of 16 columns
#pragma unroll
  for (int cb = 0; cb < 4; cb++) {
    const int Jb = J + 16 * cb;
    if (wv == cb) diag_block_rp<T>(acc[cb], L, n, Jb, max(0, min(16, n - Jb)), lane, Pd, Li, &s_fail);
    __syncthreads();
    if (wv > cb) {
      // X^T = L_bb^-1 P^T: the strip's block cb of L
      V o = V{};
#pragma unroll
      for (int u = 0; u < 4; u++) o = M::mma(Li[cl * PLD + M::row(lane, u)], acc[cb][u], o);
      acc[cb] = o;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = Jb + M::row(lane, v);
        if (rok && c < n) L[(size_t)c * n + r] = o[v];
        Lx[wv][cl * PLD + M::row(lane, v)] = o[v];   // L[J + 16 wv + cl][Jb + row]
      }
    }
    __syncthreads();
    if (wv > cb) {
      // later blocks b (cb < b <= wv): D_b[c][r] -= sum_t L[c][Jb + t] L[r][Jb + t]
#pragma unroll
      for (int b = cb + 1; b < 4; b++) {
        if (b <= wv) {
#pragma unroll
          for (int u = 0; u < 4; u++) acc[b] = M::mma(-Lx[b][cl * PLD + M::row(lane, u)], acc[cb][u], acc[b]);
        }
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (k == 0) status[p] = s_fail;
    else if (s_fail != 0 && status[p] == 0) status[p] = s_fail;
  }
}

template <typename T>
__global__ __launch_bounds__(TNT) void k_tile_solve(const int2* __restrict__ tasks, int k,
                                                    const int32_t* __restrict__ pn, const int32_t* __restrict__ pm,
                                                    const int64_t* __restrict__ fac_off, T* __restrict__ fac) {
  using M = Mfma<T>;
  using V = typename M::V;
  const int2 tk = tasks[blockIdx.x];
  const int p = tk.x, s = tk.y;
  if (s == k) return;
  const int n = pn[p] - (pm ? pm[p] : 0);
  const int J = TILE_NB * k, R = TILE_NB * s;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, cl = lane & 15;
  T* L = fac + fac_off[p];
  const int r = R + 16 * wv + cl;
  const bool rok = r < n;
  V acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; cb++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int c = J + 16 * cb + M::row(lane, v);
      acc[cb][v] = (rok && c < n) ? L[(size_t)c * n + r] : T(0);
    }
#pragma unroll
  for (int cb = 0; cb < 4; cb++) {
    const int Jb = J + 16 * cb;
    // (L_bb^-1)[cl][kk], kk = row(lane, u): strictly lower part in the block's upper
    // triangle, 1 / L_ii on the diagonal, the identity past n
    T lia[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int kk = M::row(lane, u);
      const int i = Jb + cl, j = Jb + kk;
      T v = T(0);
      if (i < n && j < n) {
        if (cl > kk) v = L[(size_t)i * n + j];
        else if (cl == kk) v = T(1) / L[(size_t)i * n + i];
      } else if (cl == kk) {
        v = T(1);
      }
      lia[u] = v;
    }
    V o = V{};
#pragma unroll
    for (int u = 0; u < 4; u++) o = M::mma(lia[u], acc[cb][u], o);
    if (rok) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = Jb + M::row(lane, v);
        if (c < n) L[(size_t)c * n + r] = o[v];
      }
    }
    // later blocks: D_b[c][r] -= sum_t L[J + 16 b + c][Jb + t] L[r][Jb + t]
#pragma unroll
    for (int b = cb + 1; b < 4; b++) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int i = J + 16 * b + cl, j = Jb + M::row(lane, u);
        const T ld = (i < n && j < n) ? L[(size_t)j * n + i] : T(0);
        acc[b] = M::mma(-ld, o[u], acc[b]);
      }
    }
  }
  (void)tid;
}

template <typename T, int KT, int MM>
static void tiled_k(const Grid& g, const int2* tasks, const std::vector<int64_t>& off, const int32_t* pn,
                    const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd, T* fac,
                    int32_t* status, const AugView& a, hipStream_t s) {
  const int nsteps = (int)off.size() - 1;
  for (int k = 0; k < nsteps; k++) {
    const int64_t cnt = off[k + 1] - off[k];
    if (cnt <= 0) continue;
    hipLaunchKernelGGL((k_tile_update<T, KT, MM>), dim3((unsigned)cnt), dim3(TNT), 0, s, g, tasks + off[k], k, pn,
                       mem_off, fac_off, mem_pool, xd, fac, status, a);
    hipLaunchKernelGGL((k_tile_solve<T>), dim3((unsigned)cnt), dim3(TNT), 0, s, tasks + off[k], k, pn,
                       MM > 0 ? a.m : nullptr, fac_off, fac);
  }
}

template <typename T, int MM>
static void tiled_kt(const Grid& g, int ktype, const int2* tasks, const std::vector<int64_t>& off, const int32_t* pn,
                     const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd, T* fac,
                     int32_t* status, const AugView& a, hipStream_t s) {
#define LRB_TK(K) tiled_k<T, K, MM>(g, tasks, off, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s)
  switch (ktype) {
    case K_WC0: LRB_TK(K_WC0); break;
    case K_WC2: LRB_TK(K_WC2); break;
    case K_IMQ: LRB_TK(K_IMQ); break;
    case K_PHS3:
      if constexpr (MM > 0) LRB_TK(K_PHS3);
      break;
    default: LRB_TK(K_GAUSS); break;
  }
#undef LRB_TK
}

template <typename T>
static void tiled_mm(const Grid& g, int ktype, const int2* tasks, const std::vector<int64_t>& off, const int32_t* pn,
                     const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd, T* fac,
                     int32_t* status, const AugView& a, hipStream_t s) {
  if (!a.m) tiled_kt<T, 0>(g, ktype, tasks, off, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s);
  else if (a.mmax <= 4) tiled_kt<T, 4>(g, ktype, tasks, off, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s);
  else tiled_kt<T, POLY_MAX>(g, ktype, tasks, off, pn, mem_off, fac_off, mem_pool, xd, fac, status, a, s);
}

void tile_tasks(const std::vector<int32_t>& list, const std::vector<int32_t>& nn, std::vector<int2>& tasks,
                std::vector<int64_t>& off) {
  int tmax = 0;
  for (int d : nn) tmax = std::max(tmax, (d + TILE_NB - 1) / TILE_NB);
  tasks.clear();
  off.assign(1, 0);
  for (int k = 0; k < tmax; k++) {
    for (size_t i = 0; i < list.size(); i++) {
      const int T = (nn[i] + TILE_NB - 1) / TILE_NB;
      for (int s = k; s < T; s++) tasks.push_back(make_int2(list[i], s));
    }
    off.push_back((int64_t)tasks.size());
  }
}

void launch_factor_tiled(const Grid& g, int ktype, bool fp32, const int2* tasks, const std::vector<int64_t>& off,
                         const int32_t* pn, const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool,
                         const double* xd, void* fac, int32_t* status, const AugView& a, hipStream_t s) {
  if (off.size() < 2) return;
  if (fp32) tiled_mm<float>(g, ktype, tasks, off, pn, mem_off, fac_off, mem_pool, xd, (float*)fac, status, a, s);
  else tiled_mm<double>(g, ktype, tasks, off, pn, mem_off, fac_off, mem_pool, xd, (double*)fac, status, a, s);
}

}  // namespace lrb
