// Cell-ordered PU evaluation over a cell-major node table (interpolation.c:39-154
// restated for the GPU).
//
// Table: one row per data node, in cell-sorted order (row s = s-th entry of the
// per-cell node lists), BS = 4 + 2^d * nft elements:
//   [x, y, (z), pad | coefficients of the node in the patches of the 2^d corners
//    of its cell, corner order Q1, nft fields each]
// Each (patch, member) coefficient appears exactly once: a member of patch p lies
// in one cell of p, and p is the patch of one corner of that cell.  An inactive
// patch (not set up) holds zeros.
//
// Evaluation: one workgroup per particle cell c (XCD-aware order), one lane per
// particle, waves past the cell's particle count exit at once.  The 3^d
// neighbour cells q are visited in ascending cell index and their nodes in
// ascending node index; phi(|x - x_j|) is computed once per (particle, node) and
// accumulated into every corner patch of c that holds q, so each patch's
// members are summed in member order (patch = concatenation of its cells' node
// lists, partition.c:337-371).  Which row entries a neighbour contributes to
// which corner is fixed at compile time, so the row pointer is the only address
// the node loop computes: coordinates and coefficients are wave-uniform scalar
// loads at constant offsets, used as SGPR operands.
//
// T = double: absolute FP64 coordinates.  T = float (mixed precision): particle
// and node coordinates relative to their own cell's centre (shifted in FP64 before
// rounding), the row holding -2 n and |n|^2 for the node n so that r^2 takes four
// operations (row_r2), FP32 coefficients, FP64 Shepard combination.
#pragma once
#include "kernels.hpp"

#include <hip/hip_ext.h>

#include <algorithm>

namespace lrb {

// Timing of the evaluation launches: when set, the next evaluation dispatch
// records its own start and end in these events (kernel-accurate timestamps that
// need no host wait); unset inside graph capture.
extern thread_local hipEvent_t t_eval_ev0, t_eval_ev1;

#define LRB_LAUNCH_TIMED(KERNEL, GRID, BLOCK, S, ...)                                        \
  do {                                                                                        \
    if (t_eval_ev0)                                                                           \
      hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, S, t_eval_ev0, t_eval_ev1, 0, __VA_ARGS__); \
    else                                                                                      \
      hipLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, S, __VA_ARGS__);                             \
  } while (0)

// Q1 corner index of offset (a, b, c) in {0,1}^3 (DMDAGetElements order)
__host__ __device__ constexpr int q1_index(int a, int b, int c) { return (b ? (a ? 2 : 3) : (a ? 1 : 0)) + 4 * c; }

// For particle-cell corner k and neighbour cell nbi (di,dj,dk in {-1,0,1}, i fastest):
// the corner index of that patch's node within the neighbour cell, or -1 when the
// corner's patch does not hold the neighbour cell.
template <int DIM>
__host__ __device__ constexpr int corner_in_nbr(int nbi, int k) {
  const int di = nbi % 3 - 1, dj = (nbi / 3) % 3 - 1, dk = DIM == 3 ? nbi / 9 - 1 : 0;
  const int oi = ((k & 3) == 1 || (k & 3) == 2) ? 1 : 0, oj = ((k & 3) >= 2) ? 1 : 0, ok = k >> 2;
  const int a = oi - di, b = oj - dj, c = ok - dk;
  const bool in = a >= 0 && a <= 1 && b >= 0 && b <= 1 && c >= 0 && c <= 1;
  return in ? q1_index(a, b, c) : -1;
}

// how many corner patches of the particle cell hold neighbour nbi
template <int DIM>
__host__ __device__ constexpr int n_corners_in_nbr(int nbi) {
  int n = 0;
  for (int k = 0; k < (1 << DIM); k++) n += corner_in_nbr<DIM>(nbi, k) >= 0 ? 1 : 0;
  return n;
}

// Whether two rows of neighbour nbi's table entries fit the scalar registers next
// to the rest of the kernel (at most 72 of them: the FP32 centre neighbour of a
// 3- or 4-field table, the FP64 face neighbours of a 3-field table).
template <int DIM, int NFT, typename T>
__host__ __device__ constexpr bool two_rows_fit(int nbi) {
  return 2 * (4 + n_corners_in_nbr<DIM>(nbi) * NFT) * (int)(sizeof(T) / 4) <= 72;
}

// phi from the squared distance.  r^3: eps^3 is folded into the table's
// coefficients (k_build_blk), so phi = r^2 sqrt(r^2).
template <int KT, typename T>
__device__ __forceinline__ T phi_r2(T r2, T eps) {
  if constexpr (KT == K_PHS3) return r2 * sqrt_fast(r2);
  else if constexpr (KT == K_GAUSS && sizeof(T) == 4) return __builtin_amdgcn_exp2f(-(eps * eps * 1.44269504f) * r2);
  // FP32 inverse multiquadric 1 / sqrt(1 + eps^2 r^2): the hardware reciprocal square
  // root (~1 ulp) instead of a square root, a square and a correctly rounded divide
  else if constexpr (KT == K_IMQ && sizeof(T) == 4) return __builtin_amdgcn_rsqf(fma(eps * eps, r2, 1.0f));
  else return rbf<KT, T>(sqrt_fast(r2), eps);
}

// The Gaussian of a corner patch with its own shape parameter, e2k = eps_k^2 (FP64)
// or eps_k^2 log2(e) (FP32, corner_eps2).  In FP32 the hardware exp2 (v_exp_f32,
// ~1 ulp) takes the place of expf's range-reduced sequence (a dozen instructions per
// exponential, 2.4 exponentials per node): the argument's own rounding, a few ulp of
// eps_k^2 r^2, already bounds the result's accuracy at that level.
template <typename T>
__device__ __forceinline__ T gauss_local(T e2k, T r2) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(-e2k * r2);
  else return exp(-e2k * r2);
}

template <int KT>
__host__ __device__ constexpr bool compact_support_blk() { return KT == K_WC0 || KT == K_WC2 || KT == K_BUMP; }

template <int DIM, int NFT>
__host__ __device__ constexpr int blk_stride() { return 4 + (1 << DIM) * NFT; }

// |d|^2 with one fixed rounding: every launch shape computes it through this
// helper, so they agree bit for bit whatever contraction the compiler would pick
// for the plain expression (fusing d0*d0 or d1*d1 gives different roundings).
template <typename T>
__device__ __forceinline__ T dist2(T d0, T d1, T d2) { return fma(d2, d2, fma(d1, d1, d0 * d0)); }

// r^2 of a particle and the node of a table row (n0 .. n3: the row's first four
// entries).  FP64 rows hold the node's coordinates, and the difference is squared.
// FP32 rows hold a = -2 n and |n|^2, n the node relative to the centre of its cell,
// and the particle comes relative to the same centre, with cp = |p|^2 (cell_r2):
// r^2 = cp + |n|^2 + a . p, four operations instead of six.  Its rounding error is a
// few ulp of cp + |n|^2 (cell-centred coordinates keep both within a few h^2, the
// same order as the subtracting form's error at the distances that dominate phi),
// and a sum rounded below zero, where particle and node coincide, is taken by
// magnitude (a free source modifier).
template <int DIM, typename T>
__device__ __forceinline__ T row_r2(T xp, T yp, T zp, T cp, T n0, T n1, T n2, T n3) {
  if constexpr (sizeof(T) == 8) {
    const T d0 = xp - n0, d1 = yp - n1, d2 = DIM == 3 ? zp - n2 : T(0);
    return dist2(d0, d1, d2);
  } else {
    T s = cp + n3;
    if constexpr (DIM == 3) s = fma(n2, zp, s);
    s = fma(n1, yp, s);
    s = fma(n0, xp, s);
    return fabsf(s);
  }
}

// cp of row_r2: |p|^2 of the particle in the neighbour cell's centred frame (FP32;
// unused in FP64)
template <int DIM, typename T>
__device__ __forceinline__ T cell_r2(T xp, T yp, T zp) {
  if constexpr (sizeof(T) == 8) return T(0);
  return dist2(xp, yp, DIM == 3 ? zp : T(0));
}

// xp, yp, zp: particle coordinates in the frame of the neighbour cell's node rows
// (FP64: absolute; FP32: relative to the neighbour cell's centre).
template <int DIM, int KT, int NFT, typename T, int NBI>
__device__ __forceinline__ void accum_blk(T xp, T yp, T zp, const T* __restrict__ B, int m,
                                          T (&acc)[1 << DIM][NFT], T eps, T rcut2, const T (&e2)[1 << DIM]) {
  constexpr int NEN = 1 << DIM, BS = blk_stride<DIM, NFT>();
  const T cp = cell_r2<DIM, T>(xp, yp, zp);
  // the next row's leading entries are loaded one node ahead (the table ends with a
  // padding row, so row m is always readable)
  T n0 = B[0], n1 = B[1], n2 = DIM == 3 ? B[2] : T(0), n3 = B[3];
  for (int r = 0; r < m; r++) {
    const T* Br = B + (size_t)r * BS;
    const T c0 = n0, c1 = n1, c2 = n2, c3 = n3;
    n0 = Br[BS]; n1 = Br[BS + 1];
    if constexpr (DIM == 3) n2 = Br[BS + 2];
    n3 = Br[BS + 3];
    const T r2 = row_r2<DIM, T>(xp, yp, zp, cp, c0, c1, c2, c3);
    if constexpr (compact_support_blk<KT>()) {
      // every lane beyond the support: phi = 0 exactly, nothing to add
      if (!__any(r2 < rcut2)) continue;
    }
    T phi = T(0);
    if constexpr (KT != K_GAUSS_LOCAL) phi = phi_r2<KT, T>(r2, eps);
#pragma unroll
    for (int k = 0; k < NEN; k++) {
      const int kp = corner_in_nbr<DIM>(NBI, k);
      if (kp >= 0) {
        if constexpr (KT == K_GAUSS_LOCAL) phi = gauss_local<T>(e2[k], r2);
#pragma unroll
        for (int f = 0; f < NFT; f++) acc[k][f] += phi * Br[4 + kp * NFT + f];
      }
    }
  }
}

// Wait for every scalar load in flight (s_waitcnt lgkmcnt(0); vmcnt and expcnt
// left at their maxima).  Scalar loads return out of order, so this is the only
// wait there is for them.
__device__ __forceinline__ void wait_scalar_loads() { __builtin_amdgcn_s_waitcnt(0xC07F); }

// One table row's entries that neighbour NBI uses, as wave-uniform values.
template <int DIM, int NFT, typename T>
struct BlkRow {
  T v[blk_stride<DIM, NFT>()];
};

template <int DIM, int NFT, typename T, int NBI>
__device__ __forceinline__ void load_row(BlkRow<DIM, NFT, T>& row, const T* __restrict__ p) {
  constexpr int NEN = 1 << DIM;
  // all four leading entries, pad included, so that they arrive as one 16-byte load
  row.v[0] = p[0];
  row.v[1] = p[1];
  row.v[2] = p[2];
  row.v[3] = p[3];
#pragma unroll
  for (int k = 0; k < NEN; k++) {
    const int kp = corner_in_nbr<DIM>(NBI, k);
    if (kp >= 0) {
#pragma unroll
      for (int f = 0; f < NFT; f++) row.v[4 + kp * NFT + f] = p[4 + kp * NFT + f];
    }
  }
}

template <int DIM, int KT, int NFT, typename T, int NBI>
__device__ __forceinline__ void node_blk(const BlkRow<DIM, NFT, T>& row, T xp, T yp, T zp, T cp,
                                         T (&acc)[1 << DIM][NFT], T eps, T rcut2, const T (&e2)[1 << DIM]) {
  constexpr int NEN = 1 << DIM;
  const T r2 = row_r2<DIM, T>(xp, yp, zp, cp, row.v[0], row.v[1], row.v[2], row.v[3]);
  // every lane beyond a compact support: phi = 0 exactly, nothing to add
  if (compact_support_blk<KT>() && !__any(r2 < rcut2)) return;
  if constexpr (KT == K_GAUSS_LOCAL) {
    // each corner patch has its own shape parameter: phi = exp(-eps_k^2 r^2)
#pragma unroll
    for (int k = 0; k < NEN; k++) {
      const int kp = corner_in_nbr<DIM>(NBI, k);
      if (kp >= 0) {
        const T phi = gauss_local<T>(e2[k], r2);
#pragma unroll
        for (int f = 0; f < NFT; f++) acc[k][f] += phi * row.v[4 + kp * NFT + f];
      }
    }
    return;
  }
  const T phi = phi_r2<KT, T>(r2, eps);
#pragma unroll
  for (int k = 0; k < NEN; k++) {
    const int kp = corner_in_nbr<DIM>(NBI, k);
    if (kp >= 0) {
#pragma unroll
      for (int f = 0; f < NFT; f++) acc[k][f] += phi * row.v[4 + kp * NFT + f];
    }
  }
}

// accum_blk with the scalar loads of row r + 1 in flight while row r is computed:
// two rows alternate in registers, and the wait for the next row comes after the
// current one's arithmetic.  Same operations in the same node order as accum_blk,
// so the same values.  Row m, one past the neighbour's last node, may be loaded:
// it is the next cell's first row or the table's padding row (row N).
template <int DIM, int KT, int NFT, typename T, int NBI>
__device__ __forceinline__ void accum_blk_pipe(T xp, T yp, T zp, const T* __restrict__ B, int m,
                                               T (&acc)[1 << DIM][NFT], T eps, T rcut2, const T (&e2)[1 << DIM]) {
  constexpr int BS = blk_stride<DIM, NFT>();
  const T cp = cell_r2<DIM, T>(xp, yp, zp);
  BlkRow<DIM, NFT, T> ra, rb;
  load_row<DIM, NFT, T, NBI>(ra, B);
  wait_scalar_loads();
  const T* q = B + BS;   // the next row to load
  int r = m;
  // the empty asm statements keep the loads where they are written: without them
  // the compiler sinks a row's coefficient loads into the branch that uses them,
  // right before their first use
  while (r >= 2) {
    load_row<DIM, NFT, T, NBI>(rb, q);
    q += BS;
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    node_blk<DIM, KT, NFT, T, NBI>(ra, xp, yp, zp, cp, acc, eps, rcut2, e2);
    __builtin_amdgcn_sched_barrier(0);
    wait_scalar_loads();
    asm volatile("" ::"s"(rb.v[2]), "s"(rb.v[3]));   // keeps the 16-byte load whole
    load_row<DIM, NFT, T, NBI>(ra, q);
    q += BS;
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    node_blk<DIM, KT, NFT, T, NBI>(rb, xp, yp, zp, cp, acc, eps, rcut2, e2);
    __builtin_amdgcn_sched_barrier(0);
    wait_scalar_loads();
    asm volatile("" ::"s"(ra.v[2]), "s"(ra.v[3]));
    r -= 2;
  }
  if (r == 1) node_blk<DIM, KT, NFT, T, NBI>(ra, xp, yp, zp, cp, acc, eps, rcut2, e2);
}

// Two rows per wait, for rows of at most 16 entries: rows j, j + 1 are computed
// while j + 2, j + 3 load, so each load has two nodes of arithmetic to land in.
// Same node order and operations as accum_blk.  Up to two rows past the
// neighbour's last node may be loaded: the next cell's rows or the table's
// padding rows N, N + 1.
template <int DIM, int KT, int NFT, typename T, int NBI>
__device__ __forceinline__ void accum_blk_pipe2(T xp, T yp, T zp, const T* __restrict__ B, int m,
                                                T (&acc)[1 << DIM][NFT], T eps, T rcut2, const T (&e2)[1 << DIM]) {
  constexpr int BS = blk_stride<DIM, NFT>();
  const T cp = cell_r2<DIM, T>(xp, yp, zp);
  BlkRow<DIM, NFT, T> a0, a1, b0, b1;
  load_row<DIM, NFT, T, NBI>(a0, B);
  load_row<DIM, NFT, T, NBI>(a1, B + BS);
  wait_scalar_loads();
  const T* q = B + 2 * BS;   // the next row to load
  int r = m;
  while (r >= 4) {
    load_row<DIM, NFT, T, NBI>(b0, q);
    load_row<DIM, NFT, T, NBI>(b1, q + BS);
    q += 2 * BS;
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    node_blk<DIM, KT, NFT, T, NBI>(a0, xp, yp, zp, cp, acc, eps, rcut2, e2);
    node_blk<DIM, KT, NFT, T, NBI>(a1, xp, yp, zp, cp, acc, eps, rcut2, e2);
    __builtin_amdgcn_sched_barrier(0);
    wait_scalar_loads();
    asm volatile("" ::"s"(b0.v[2]), "s"(b0.v[3]), "s"(b1.v[2]), "s"(b1.v[3]));
    load_row<DIM, NFT, T, NBI>(a0, q);
    load_row<DIM, NFT, T, NBI>(a1, q + BS);
    q += 2 * BS;
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    node_blk<DIM, KT, NFT, T, NBI>(b0, xp, yp, zp, cp, acc, eps, rcut2, e2);
    node_blk<DIM, KT, NFT, T, NBI>(b1, xp, yp, zp, cp, acc, eps, rcut2, e2);
    __builtin_amdgcn_sched_barrier(0);
    wait_scalar_loads();
    asm volatile("" ::"s"(a0.v[2]), "s"(a0.v[3]), "s"(a1.v[2]), "s"(a1.v[3]));
    r -= 4;
  }
  // 0 <= r < 4 rows left; a0, a1 hold the first two of them
  if (r >= 1) node_blk<DIM, KT, NFT, T, NBI>(a0, xp, yp, zp, cp, acc, eps, rcut2, e2);
  if (r >= 2) node_blk<DIM, KT, NFT, T, NBI>(a1, xp, yp, zp, cp, acc, eps, rcut2, e2);
  if (r == 3) {
    load_row<DIM, NFT, T, NBI>(b0, q);
    wait_scalar_loads();
    node_blk<DIM, KT, NFT, T, NBI>(b0, xp, yp, zp, cp, acc, eps, rcut2, e2);
  }
}

#ifndef LRB_EVAL_PIPE
#define LRB_EVAL_PIPE 1
#endif

// Matrix-core accumulation (FP32 tables, up to 4 fields): the node loop computes phi
// on the vector ALUs as above (one lane per particle, node coordinates as scalar
// loads), and each corner patch's accumulation phi * coefficient becomes one
// v_mfma_f32_4x4x1_16b_f32 instead of NFT vector FMAs.  That instruction holds 16
// independent 4 x 4 blocks, D[v] of lane 4b + j += A(lane 4b + v) * B(lane 4b + j):
// lane 4b + i supplies A = the coefficient of field i (a vector load of the node row
// at column i; column 3 reads the next entry and lands in an unused register) and
// B = phi of its own particle, so its four result registers hold its own
// particle's sums for fields 0 .. 3.  K = 1: each accumulation is one product added
// with one rounding, as the vector FMA does it, so the values are the same bit for
// bit as the vector-ALU loops'.
typedef float ev_f4 __attribute__((ext_vector_type(4)));
#ifndef LRB_MFMA_WAVES
#define LRB_MFMA_WAVES 4
#endif

template <int DIM, int KT, int NFT, int NBI, int R>
__device__ __forceinline__ void mfma_rows(float xp, float yp, float zp, const float* __restrict__ B,
                                          ev_f4 (&acc)[1 << DIM], float eps, float rcut2,
                                          const float (&e2)[1 << DIM], int lq) {
  constexpr int NEN = 1 << DIM, BS = blk_stride<DIM, NFT>();
  const float cp = cell_r2<DIM, float>(xp, yp, zp);
  // every load of the R rows is issued before any of their arithmetic, so the
  // rows share one memory latency
  float c[R][4], cf[R][NEN];
#pragma unroll
  for (int r = 0; r < R; r++) {
    c[r][0] = B[r * BS]; c[r][1] = B[r * BS + 1]; c[r][2] = DIM == 3 ? B[r * BS + 2] : 0.f; c[r][3] = B[r * BS + 3];
#pragma unroll
    for (int k = 0; k < NEN; k++) {
      const int kp = corner_in_nbr<DIM>(NBI, k);
      cf[r][k] = kp >= 0 ? B[r * BS + 4 + kp * NFT + lq] : 0.f;
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const float r2 = row_r2<DIM, float>(xp, yp, zp, cp, c[r][0], c[r][1], c[r][2], c[r][3]);
    bool use = true;
    if constexpr (compact_support_blk<KT>()) use = __any(r2 < rcut2);
    if (use) {
      float phi = 0.f;
      if constexpr (KT != K_GAUSS_LOCAL) phi = phi_r2<KT, float>(r2, eps);
#pragma unroll
      for (int k = 0; k < NEN; k++) {
        const int kp = corner_in_nbr<DIM>(NBI, k);
        if (kp >= 0) {
          if constexpr (KT == K_GAUSS_LOCAL) phi = gauss_local<float>(e2[k], r2);
          acc[k] = __builtin_amdgcn_mfma_f32_4x4x1f32(cf[r][k], phi, acc[k], 0, 0, 0);
        }
      }
    }
  }
}

// The m rows of neighbour NBI in batches of MR (fewer corners: larger batches fit
// the registers), the tail one row at a time; same node order as accum_blk.
template <int DIM, int KT, int NFT, int NBI>
__device__ __forceinline__ void accum_mfma(float xp, float yp, float zp, const float* __restrict__ B, int m,
                                           ev_f4 (&acc)[1 << DIM], float eps, float rcut2,
                                           const float (&e2)[1 << DIM], int lq) {
  constexpr int BS = blk_stride<DIM, NFT>();
  constexpr int NK = n_corners_in_nbr<DIM>(NBI);
  constexpr int MR = NK >= 8 ? 2 : (NK >= 4 ? 4 : 8);
  int r = 0;
  for (; r + MR <= m; r += MR) mfma_rows<DIM, KT, NFT, NBI, MR>(xp, yp, zp, B + (size_t)r * BS, acc, eps, rcut2, e2, lq);
  for (; r < m; r++) mfma_rows<DIM, KT, NFT, NBI, 1>(xp, yp, zp, B + (size_t)r * BS, acc, eps, rcut2, e2, lq);
}

// The particles s0 .. s0 + 63 of cell (ci, cj, ck) with matrix-core accumulation:
// register f of acc[k] ends as field f of corner k's sum for the lane's particle.
template <int DIM, int KT, int NFT>
__device__ __forceinline__ void eval_wave_mfma(const Grid& g, int ci, int cj, int ck, float xp, float yp, float zp,
                                               const float* __restrict__ blk, const int32_t* __restrict__ cs,
                                               float eps, float rcut2, const float (&e2)[1 << DIM], int lane,
                                               ev_f4 (&acc)[1 << DIM]) {
  constexpr int NNB = DIM == 3 ? 27 : 9;
  constexpr int BS = blk_stride<DIM, NFT>();
  const int lq = lane & 3;
#pragma unroll 1
  for (int nbi = 0; nbi < NNB; nbi++) {
    const int di = nbi % 3 - 1, dj = (nbi / 3) % 3 - 1, dk = DIM == 3 ? nbi / 9 - 1 : 0;
    const int a = ci + di, bb = cj + dj, e = ck + dk;
    if (a < 0 || a >= g.nc[0] || bb < 0 || bb >= g.nc[1] || (DIM == 3 && (e < 0 || e >= g.nc[2]))) continue;
    const int q = a + g.nc[0] * (bb + g.nc[1] * e);
    const int nb = cs[q], ne = cs[q + 1];
    if (nb == ne) continue;
    const float qx = xp - (float)(di * g.h[0]), qy = yp - (float)(dj * g.h[1]);
    const float qz = DIM == 3 ? zp - (float)(dk * g.h[2]) : 0.f;
    const float* B = blk + (size_t)nb * BS;
    switch (nbi) {
#define LRB_MCASE(V) case V: if constexpr (V < NNB) accum_mfma<DIM, KT, NFT, V>(qx, qy, qz, B, ne - nb, acc, eps, rcut2, e2, lq); break;
      LRB_MCASE(0) LRB_MCASE(1) LRB_MCASE(2) LRB_MCASE(3) LRB_MCASE(4) LRB_MCASE(5) LRB_MCASE(6) LRB_MCASE(7)
      LRB_MCASE(8) LRB_MCASE(9) LRB_MCASE(10) LRB_MCASE(11) LRB_MCASE(12) LRB_MCASE(13) LRB_MCASE(14)
      LRB_MCASE(15) LRB_MCASE(16) LRB_MCASE(17) LRB_MCASE(18) LRB_MCASE(19) LRB_MCASE(20) LRB_MCASE(21)
      LRB_MCASE(22) LRB_MCASE(23) LRB_MCASE(24) LRB_MCASE(25) LRB_MCASE(26)
#undef LRB_MCASE
      default: break;
    }
  }
}


// The 3^d neighbour cells in ascending cell index, as one straight sequence of
// loops specialised per neighbour (no dispatch between them, so the accumulators
// keep their registers across the whole sequence).
template <int DIM, int KT, int NFT, typename T, int NBI>
__device__ __forceinline__ void accum_nbrs(const Grid& g, int ci, int cj, int ck, T xp, T yp, T zp,
                                           const T* __restrict__ blk, const int32_t* __restrict__ cs,
                                           T (&acc)[1 << DIM][NFT], T eps, T rcut2, const T (&e2)[1 << DIM]) {
  constexpr int NNB = DIM == 3 ? 27 : 9;
  if constexpr (NBI < NNB) {
    constexpr int di = NBI % 3 - 1, dj = (NBI / 3) % 3 - 1, dk = DIM == 3 ? NBI / 9 - 1 : 0;
    constexpr int BS = blk_stride<DIM, NFT>();
    const int a = ci + di, bb = cj + dj, e = ck + dk;
    const bool in = a >= 0 && a < g.nc[0] && bb >= 0 && bb < g.nc[1] && (DIM == 2 || (e >= 0 && e < g.nc[2]));
    if (in) {
      const int q = a + g.nc[0] * (bb + g.nc[1] * e);
      const int nb = cs[q], ne = cs[q + 1];
      if (nb < ne) {
        // FP32: particle relative to the neighbour cell's centre (node rows are
        // relative to their own cell's centre); FP64: absolute coordinates
        T qx = xp, qy = yp, qz = zp;
        if constexpr (sizeof(T) == 4) {
          qx = xp - (T)(di * g.h[0]); qy = yp - (T)(dj * g.h[1]);
          if constexpr (DIM == 3) qz = zp - (T)(dk * g.h[2]);
        }
        // two FP64 rows of the centre neighbour (all 2^d corners) would not fit the
        // scalar registers: it keeps the single-row loop
        if constexpr (LRB_EVAL_PIPE && two_rows_fit<DIM, NFT, T>(NBI))
          accum_blk_pipe<DIM, KT, NFT, T, NBI>(qx, qy, qz, blk + (size_t)nb * BS, ne - nb, acc, eps, rcut2, e2);
        else
          accum_blk<DIM, KT, NFT, T, NBI>(qx, qy, qz, blk + (size_t)nb * BS, ne - nb, acc, eps, rcut2, e2);
      }
    }
    accum_nbrs<DIM, KT, NFT, T, NBI + 1>(g, ci, cj, ck, xp, yp, zp, blk, cs, acc, eps, rcut2, e2);
  }
}

// Shape parameter squared of each corner patch of cell (ci, cj, ck) for the
// Gaussian with per-patch shape (K_GAUSS_LOCAL); zeros for the other kernels.
template <int DIM, int KT, typename T>
__device__ __forceinline__ void corner_eps2(const Grid& g, int ci, int cj, int ck, const EvalExt& ev,
                                            T (&e2)[1 << DIM]) {
#pragma unroll
  for (int k = 0; k < (1 << DIM); k++) {
    e2[k] = T(0);
    if constexpr (KT == K_GAUSS_LOCAL) {
      int a, b, c;
      corner_offset(k, a, b, c);
      const double e = ev.peps[(ci + a) + g.n[0] * ((cj + b) + g.n[1] * (DIM == 3 ? ck + c : 0))];
      e2[k] = (T)(sizeof(T) == 4 ? e * e * 1.4426950408889634 : e * e);   // FP32: log2(e) folded in
    }
  }
}

// Shepard combination of the corner sums, val = sum_k w_k acc_k (the polynomial
// parts of the corner patches, if any, come from poly_part).
template <int DIM, int NFT, typename T>
__device__ __forceinline__ void combine_corners(const double* __restrict__ w, int64_t ld, int i,
                                                const T (&acc)[1 << DIM][NFT], double (&val)[NFT]) {
#pragma unroll
  for (int f = 0; f < NFT; f++) val[f] = 0.0;
#pragma unroll
  for (int k = 0; k < (1 << DIM); k++) {
    const double wk = w[(size_t)k * ld + i];
#pragma unroll
    for (int f = 0; f < NFT; f++) val[f] += (double)acc[k][f] * wk;
  }
}

// Monomial q of the basis 1, x, y, (z), xx, xy, (xz), yy, (yz), (zz) is the product
// of coordinates mono_factor(q, 0) and mono_factor(q, 1) (-1: factor 1).
template <int DIM>
__host__ __device__ constexpr int mono_factor(int q, int which) {
  if (q == 0) return -1;
  if (q <= DIM) return which == 0 ? q - 1 : -1;
  int k = DIM + 1;
  for (int i = 0; i < DIM; i++)
    for (int j = i; j < DIM; j++) {
      if (k == q) return which == 0 ? i : j;
      k++;
    }
  return -1;
}

// Polynomial parts of the corner patches of cell (ci, cj, ck) at particle i, in FP64:
// add = sum_k w_k p_k(x) . beta_k, the monomials of the patch node's scaled local
// coordinates (x - node_k) ps times the patch's polynomial coefficients (ptab row
// [mm][NFT] of patch k).
template <int DIM, int NFT>
__device__ __forceinline__ void poly_part(const Grid& g, int ci, int cj, int ck, const double (&x)[3], int64_t ld,
                                          int i, const double* __restrict__ w, const EvalExt& ev, double (&add)[NFT]) {
#pragma unroll
  for (int f = 0; f < NFT; f++) add[f] = 0.0;
  // one corner at a time: unrolled, the 2^d corners' rows would be loaded together
  // into more scalar registers than there are, and the spill would take vector
  // registers from the whole kernel
#pragma unroll 1
  for (int k = 0; k < (1 << DIM); k++) {
    int a, b, e;
    corner_offset(k, a, b, e);
    const int pi = ci + a, pj = cj + b, pk = DIM == 3 ? ck + e : 0;
    const int p = pi + g.n[0] * (pj + g.n[1] * pk);
    const int idx[3] = {pi, pj, pk};
    double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; d++) xi[d] = (x[d] - fma((double)idx[d], g.h[d], g.O[d])) * ev.ps;
    const double* pt = ev.ptab + (size_t)p * ev.mm * NFT;
    const double wk = w[(size_t)k * ld + i];
    double sp[NFT];
#pragma unroll
    for (int f = 0; f < NFT; f++) sp[f] = 0.0;
    // monomial q = xi[fa] xi[fb] (factor 1 for -1), indices fixed at compile time so
    // that nothing here is indexed at run time (that would put arrays in scratch)
#pragma unroll
    for (int q = 0; q < POLY_MAX; q++) {
      if (q < ev.mm) {
        const int fa = mono_factor<DIM>(q, 0), fb = mono_factor<DIM>(q, 1);
        const double mq = (fa < 0 ? 1.0 : xi[fa]) * (fb < 0 ? 1.0 : xi[fb]);
#pragma unroll
        for (int f = 0; f < NFT; f++) sp[f] = fma(mq, pt[q * NFT + f], sp[f]);
      }
    }
#pragma unroll
    for (int f = 0; f < NFT; f++) add[f] = fma(wk, sp[f], add[f]);
  }
}

// The value of particle i into out (fields < nfields): the RBF part val plus, with a
// polynomial block, the polynomial part, computed here after the node loops.  The
// compiler barrier keeps its loads (weights, coordinates, the corner patches' rows)
// from being hoisted above the node loops, where they would stay live through them.
// xd: the particle's FP64 coordinates when the caller holds them (FP64 evaluation),
// else null and they are read from px.
template <int DIM, int NFT>
__device__ __forceinline__ void store_value(const Grid& g, int ci, int cj, int ck, const double* __restrict__ px,
                                            const double* xd, int64_t ld, int i, const double* __restrict__ w,
                                            const EvalExt& ev, const double (&val)[NFT], int nfields,
                                            double* __restrict__ out) {
  if (ev.ptab) {
    __asm__ volatile("" ::: "memory");
    double x[3];
    if (xd) {
      x[0] = xd[0]; x[1] = xd[1]; x[2] = xd[2];
    } else {
      x[0] = px[i]; x[1] = px[ld + i]; x[2] = DIM == 3 ? px[2 * ld + i] : 0.0;
    }
    double add[NFT];
    poly_part<DIM, NFT>(g, ci, cj, ck, x, ld, i, w, ev, add);
#pragma unroll
    for (int f = 0; f < NFT; f++)
      if (f < nfields) out[(size_t)f * ld + i] = val[f] + add[f];
    return;
  }
#pragma unroll
  for (int f = 0; f < NFT; f++)
    if (f < nfields) out[(size_t)f * ld + i] = val[f];
}

// occupancy floor (waves per SIMD) of the evaluation kernels
#ifndef LRB_EVAL_WAVES_F64
#define LRB_EVAL_WAVES_F64 6
#endif
#ifndef LRB_EVAL_WAVES_F32
#define LRB_EVAL_WAVES_F32 8
#endif
// PIPE (FP32 only): table rows per wait in the node loops of the non-centre
// neighbours, 1 or 2; WAVES: occupancy floor in waves per SIMD (the autotuner
// chooses among the combinations; each suits other launch shapes)
template <typename T>
__host__ __device__ constexpr int default_waves() { return sizeof(T) == 8 ? LRB_EVAL_WAVES_F64 : LRB_EVAL_WAVES_F32; }

template <int DIM, int KT, int NFT, typename T, int PIPE, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
void k_eval_blk(Grid g, const double* __restrict__ px, int64_t ld,
                                                  const int32_t* __restrict__ perm, const int32_t* __restrict__ pstart,
                                                  const double* __restrict__ w, const T* __restrict__ blk,
                                                  const int32_t* __restrict__ cs, int nfields,
                                                  double* __restrict__ out, const int32_t* __restrict__ cell_list,
                                                  const int32_t* __restrict__ list_count, int deal, EvalExt ev) {
  constexpr int NEN = 1 << DIM;
  constexpr int NNB = DIM == 3 ? 27 : 9;
  constexpr int BS = blk_stride<DIM, NFT>();
  // dense swarm: one block per cell, XCD-aware order (block b runs on XCD b % 8,
  // which gets a contiguous range of cells); sparse swarm: one block per listed
  // non-empty cell
  const int b = blockIdx.x;
  int c;
  if (cell_list) {
    // listed cells (ascending; the grid is a multiple of 8 with at least as many
    // blocks as listed cells): deal 0, XCD b % 8 takes a contiguous run of them;
    // deal K >= 1, runs of K consecutive entries go round-robin over the XCDs
    // (K = 1: block b takes entry b).  Round-robin balances heavy and light cells
    // over the XCDs; runs of K keep K neighbouring cells' table rows in one L2.
    // The last partial round of 8K blocks keeps b -> b, so every entry has one block.
    const int cnt = *list_count;
    int pos = b;
    if (deal == 0) {
      const int per = (cnt + 7) / 8;
      pos = (b % 8) * per + b / 8;
      if (b / 8 >= per) return;
    } else if (deal > 1) {
      const int r = b / 8, x = b % 8;
      const int full = ((int)gridDim.x / 8) / deal;
      if (r < full * deal) pos = (r / deal) * 8 * deal + x * deal + (r % deal);
    }
    if (pos >= cnt) return;
    c = cell_list[pos];
  } else {
    const int per = (gridDim.x + 7) / 8;
    c = (b % 8) * per + b / 8;
    if (c >= g.C) return;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const T eps = (T)g.eps_r;
  const double rc = 1.0 / g.eps_r;
  const T rcut2 = (T)(rc * rc * (sizeof(T) == 4 ? (1.0 + 1e-5) : (1.0 + 1e-12)));
  {
  const int pb = pstart[c], pe = pstart[c + 1];
  if (pb + wv * 64 >= pe) return;
  const int ci = c % g.nc[0], cj = (c / g.nc[0]) % g.nc[1], ck = DIM == 3 ? c / (g.nc[0] * g.nc[1]) : 0;
  T e2[NEN];
  corner_eps2<DIM, KT, T>(g, ci, cj, ck, ev, e2);
  // an empty 3^d neighbourhood leaves the particles' values untouched (interpolation.c:89)
  int nnb = 0;
  for (int nbi = 0; nbi < NNB; nbi++) {
    const int a = ci + nbi % 3 - 1, bb = cj + (nbi / 3) % 3 - 1, e = ck + (DIM == 3 ? nbi / 9 - 1 : 0);
    if (a < 0 || a >= g.nc[0] || bb < 0 || bb >= g.nc[1] || (DIM == 3 && (e < 0 || e >= g.nc[2]))) continue;
    const int q = a + g.nc[0] * (bb + g.nc[1] * e);
    nnb += cs[q + 1] - cs[q];
  }
  if (nnb == 0) return;
  // FP32: coordinates relative to the particle cell's centre
  const double ox = sizeof(T) == 4 ? g.O[0] + (ci + 0.5) * g.h[0] : 0.0;
  const double oy = sizeof(T) == 4 ? g.O[1] + (cj + 0.5) * g.h[1] : 0.0;
  const double oz = sizeof(T) == 4 && DIM == 3 ? g.O[2] + (ck + 0.5) * g.h[2] : 0.0;
  for (int s0 = pb + wv * 64; s0 < pe; s0 += nw * 64) {
    const int s = s0 + lane;
    const bool live = s < pe;
    const int i = live ? perm[s] : perm[s0];
    T xp, yp, zp;
    if constexpr (sizeof(T) == 4) {
      xp = (T)(px[i] - ox); yp = (T)(px[ld + i] - oy); zp = DIM == 3 ? (T)(px[2 * ld + i] - oz) : T(0);
    } else {
      xp = px[i]; yp = px[ld + i]; zp = DIM == 3 ? px[2 * ld + i] : T(0);
    }
    if constexpr (PIPE == 3 && sizeof(T) == 4 && NFT <= 4) {
      // matrix-core accumulation: register f of acc4[k] holds field f of the lane's particle
      ev_f4 acc4[NEN];
#pragma unroll
      for (int k = 0; k < NEN; k++) acc4[k] = ev_f4{0.f, 0.f, 0.f, 0.f};
      eval_wave_mfma<DIM, KT, NFT>(g, ci, cj, ck, xp, yp, zp, blk, cs, eps, rcut2, e2, lane, acc4);
      if (!live) continue;
      T acc[NEN][NFT];
#pragma unroll
      for (int k = 0; k < NEN; k++)
#pragma unroll
        for (int f = 0; f < NFT; f++) acc[k][f] = acc4[k][f];
      double val[NFT];
      combine_corners<DIM, NFT, T>(w, ld, i, acc, val);
      {
        const double xd[3] = {(double)xp, (double)yp, (double)zp};
        store_value<DIM, NFT>(g, ci, cj, ck, px, sizeof(T) == 8 ? xd : nullptr, ld, i, w, ev, val, nfields, out);
      }
      continue;
    }
    T acc[NEN][NFT];
#pragma unroll
    for (int k = 0; k < NEN; k++)
#pragma unroll
      for (int f = 0; f < NFT; f++) acc[k][f] = T(0);
    if constexpr (sizeof(T) == 8) {
      // FP64: straight sequence, accumulators stay in place (76 VGPRs, no spill)
      accum_nbrs<DIM, KT, NFT, T, 0>(g, ci, cj, ck, xp, yp, zp, blk, cs, acc, eps, rcut2, e2);
    } else {
      // FP32: one loop over the neighbours dispatching to the specialised loops
      // (measured faster than the straight sequence at 8 waves per SIMD)
#pragma unroll 1
      for (int nbi = 0; nbi < NNB; nbi++) {
        const int di = nbi % 3 - 1, dj = (nbi / 3) % 3 - 1, dk = DIM == 3 ? nbi / 9 - 1 : 0;
        const int a = ci + di, bb = cj + dj, e = ck + dk;
        if (a < 0 || a >= g.nc[0] || bb < 0 || bb >= g.nc[1] || (DIM == 3 && (e < 0 || e >= g.nc[2]))) continue;
        const int q = a + g.nc[0] * (bb + g.nc[1] * e);
        const int nb = cs[q], ne = cs[q + 1];
        if (nb == ne) continue;
        const T qx = xp - (T)(di * g.h[0]), qy = yp - (T)(dj * g.h[1]);
        const T qz = DIM == 3 ? zp - (T)(dk * g.h[2]) : T(0);
        const T* B = blk + (size_t)nb * BS;
        switch (nbi) {
#if LRB_EVAL_PIPE
#define LRB_CASE(V)                                                                                     \
  case V:                                                                                               \
    if constexpr (PIPE == 2 && 4 + n_corners_in_nbr<DIM>(V) * NFT <= 16)                                 \
      accum_blk_pipe2<DIM, KT, NFT, T, V>(qx, qy, qz, B, ne - nb, acc, eps, rcut2, e2);                     \
    else if constexpr (two_rows_fit<DIM, NFT, T>(V))                                                    \
      accum_blk_pipe<DIM, KT, NFT, T, V>(qx, qy, qz, B, ne - nb, acc, eps, rcut2, e2);                      \
    else                                                                                                \
      accum_blk<DIM, KT, NFT, T, V>(qx, qy, qz, B, ne - nb, acc, eps, rcut2, e2);                           \
    break;
#else
#define LRB_CASE(V) case V: accum_blk<DIM, KT, NFT, T, V>(qx, qy, qz, B, ne - nb, acc, eps, rcut2, e2); break;
#endif
          LRB_CASE(0) LRB_CASE(1) LRB_CASE(2) LRB_CASE(3) LRB_CASE(4) LRB_CASE(5) LRB_CASE(6) LRB_CASE(7) LRB_CASE(8)
          LRB_CASE(9) LRB_CASE(10) LRB_CASE(11) LRB_CASE(12) LRB_CASE(13) LRB_CASE(14) LRB_CASE(15) LRB_CASE(16)
          LRB_CASE(17) LRB_CASE(18) LRB_CASE(19) LRB_CASE(20) LRB_CASE(21) LRB_CASE(22) LRB_CASE(23) LRB_CASE(24)
          LRB_CASE(25) LRB_CASE(26)
#undef LRB_CASE
          default: break;
        }
      }
    }
    if (!live) continue;
    double val[NFT];
    combine_corners<DIM, NFT, T>(w, ld, i, acc, val);
    {
      const double xd[3] = {(double)xp, (double)yp, (double)zp};
      store_value<DIM, NFT>(g, ci, cj, ck, px, sizeof(T) == 8 ? xd : nullptr, ld, i, w, ev, val, nfields, out);
    }
  }
  }
}

// Sparse swarms (a few particles per cell): a wave serves 64/G cells at once, one
// group of G lanes per listed non-empty cell, each lane one particle of its cell
// (cells with more than G particles are served in chunks of G).  Groups walk the
// same neighbour offsets in the same order, so the corner mapping stays a
// compile-time property; each group reads its own cell's rows (vector loads,
// shared addresses inside a group).  Per particle the arithmetic and its order
// are those of k_eval_blk.
template <int DIM, int KT, int NFT, typename T, int NBI>
__device__ __forceinline__ void accum_sparse(T xp, T yp, T zp, const T* __restrict__ B, int m, int mmax,
                                             T (&acc)[1 << DIM][NFT], T eps, T rcut2, const T (&e2)[1 << DIM]) {
  constexpr int NEN = 1 << DIM, BS = blk_stride<DIM, NFT>();
  const T cp = cell_r2<DIM, T>(xp, yp, zp);
  for (int r = 0; r < mmax; r++) {
    const bool act = r < m;
    T r2 = rcut2 * T(4);
    const T* Br = B + (size_t)r * BS;
    if (act) r2 = row_r2<DIM, T>(xp, yp, zp, cp, Br[0], Br[1], DIM == 3 ? Br[2] : T(0), Br[3]);
    if constexpr (compact_support_blk<KT>()) {
      if (!__any(r2 < rcut2)) continue;
    }
    if (act) {
      T phi = T(0);
      if constexpr (KT != K_GAUSS_LOCAL) phi = phi_r2<KT, T>(r2, eps);
#pragma unroll
      for (int k = 0; k < NEN; k++) {
        const int kp = corner_in_nbr<DIM>(NBI, k);
        if (kp >= 0) {
          if constexpr (KT == K_GAUSS_LOCAL) phi = gauss_local<T>(e2[k], r2);
#pragma unroll
          for (int f = 0; f < NFT; f++) acc[k][f] += phi * Br[4 + kp * NFT + f];
        }
      }
    }
  }
}

__device__ __forceinline__ int wave_max_int(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

template <int DIM, int KT, int NFT, typename T, int G>
__global__ __launch_bounds__(256) void k_eval_sparse(Grid g, const double* __restrict__ px, int64_t ld,
                                                     const int32_t* __restrict__ perm, const int32_t* __restrict__ pstart,
                                                     const double* __restrict__ w, const T* __restrict__ blk,
                                                     const int32_t* __restrict__ cs, int nfields, double* __restrict__ out,
                                                     const int32_t* __restrict__ cell_list,
                                                     const int32_t* __restrict__ list_count, EvalExt ev) {
  constexpr int NG = 64 / G, NEN = 1 << DIM, NNB = DIM == 3 ? 27 : 9, BS = blk_stride<DIM, NFT>();
  const int lane = threadIdx.x & 63;
  const int gi = lane / G, mlane = lane % G;
  const int cnt = *list_count;
  // blocks of 4 waves over the ascending cell list; XCD b % 8 takes a contiguous
  // run of blocks (the grid is a multiple of 8 with at least the needed blocks)
  const int need = (cnt + 4 * NG - 1) / (4 * NG), per = (need + 7) / 8;
  const int lb = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if ((int)blockIdx.x / 8 >= per || lb >= need) return;
  const int wave = lb * 4 + (threadIdx.x >> 6);
  if (wave * NG >= cnt) return;
  const int item = wave * NG + gi;
  const bool gvalid = item < cnt;
  const int c = cell_list[gvalid ? item : wave * NG];
  const int ci = c % g.nc[0], cj = (c / g.nc[0]) % g.nc[1], ck = DIM == 3 ? c / (g.nc[0] * g.nc[1]) : 0;
  T e2[NEN];
  corner_eps2<DIM, KT, T>(g, ci, cj, ck, ev, e2);
  int nnb = 0;
  for (int nbi = 0; nbi < NNB; nbi++) {
    const int a = ci + nbi % 3 - 1, bb = cj + (nbi / 3) % 3 - 1, e = ck + (DIM == 3 ? nbi / 9 - 1 : 0);
    if (a < 0 || a >= g.nc[0] || bb < 0 || bb >= g.nc[1] || (DIM == 3 && (e < 0 || e >= g.nc[2]))) continue;
    const int q = a + g.nc[0] * (bb + g.nc[1] * e);
    nnb += cs[q + 1] - cs[q];
  }
  // an empty 3^d neighbourhood leaves the particles' values untouched (interpolation.c:89)
  const int pb = pstart[c], pe = (gvalid && nnb > 0) ? pstart[c + 1] : pb;
  const int nchunk = wave_max_int((pe - pb + G - 1) / G);
  const T eps = (T)g.eps_r;
  const double rc = 1.0 / g.eps_r;
  const T rcut2 = (T)(rc * rc * (sizeof(T) == 4 ? (1.0 + 1e-5) : (1.0 + 1e-12)));
  const double ox = sizeof(T) == 4 ? g.O[0] + (ci + 0.5) * g.h[0] : 0.0;
  const double oy = sizeof(T) == 4 ? g.O[1] + (cj + 0.5) * g.h[1] : 0.0;
  const double oz = sizeof(T) == 4 && DIM == 3 ? g.O[2] + (ck + 0.5) * g.h[2] : 0.0;
  for (int k = 0; k < nchunk; k++) {
    const int s = pb + k * G + mlane;
    const bool live = s < pe;
    const int i = perm[live ? s : 0];
    T xp, yp, zp;
    if constexpr (sizeof(T) == 4) {
      xp = (T)(px[i] - ox); yp = (T)(px[ld + i] - oy); zp = DIM == 3 ? (T)(px[2 * ld + i] - oz) : T(0);
    } else {
      xp = px[i]; yp = px[ld + i]; zp = DIM == 3 ? px[2 * ld + i] : T(0);
    }
    T acc[NEN][NFT];
#pragma unroll
    for (int kk = 0; kk < NEN; kk++)
#pragma unroll
      for (int f = 0; f < NFT; f++) acc[kk][f] = T(0);
#pragma unroll 1
    for (int nbi = 0; nbi < NNB; nbi++) {
      const int di = nbi % 3 - 1, dj = (nbi / 3) % 3 - 1, dk = DIM == 3 ? nbi / 9 - 1 : 0;
      const int a = ci + di, bb = cj + dj, e = ck + dk;
      int nb = 0, ne = 0;
      if (live && !(a < 0 || a >= g.nc[0] || bb < 0 || bb >= g.nc[1] || (DIM == 3 && (e < 0 || e >= g.nc[2])))) {
        const int q = a + g.nc[0] * (bb + g.nc[1] * e);
        nb = cs[q]; ne = cs[q + 1];
      }
      const int mq = ne - nb;
      const int mmax = wave_max_int(mq);
      if (mmax == 0) continue;
      T qx = xp, qy = yp, qz = zp;
      if constexpr (sizeof(T) == 4) {
        qx = xp - (T)(di * g.h[0]); qy = yp - (T)(dj * g.h[1]);
        if constexpr (DIM == 3) qz = zp - (T)(dk * g.h[2]);
      }
      const T* B = blk + (size_t)nb * BS;
      switch (nbi) {
#define LRB_CASE(V) case V: accum_sparse<DIM, KT, NFT, T, V>(qx, qy, qz, B, mq, mmax, acc, eps, rcut2, e2); break;
        LRB_CASE(0) LRB_CASE(1) LRB_CASE(2) LRB_CASE(3) LRB_CASE(4) LRB_CASE(5) LRB_CASE(6) LRB_CASE(7) LRB_CASE(8)
        LRB_CASE(9) LRB_CASE(10) LRB_CASE(11) LRB_CASE(12) LRB_CASE(13) LRB_CASE(14) LRB_CASE(15) LRB_CASE(16)
        LRB_CASE(17) LRB_CASE(18) LRB_CASE(19) LRB_CASE(20) LRB_CASE(21) LRB_CASE(22) LRB_CASE(23) LRB_CASE(24)
        LRB_CASE(25) LRB_CASE(26)
#undef LRB_CASE
        default: break;
      }
    }
    if (!live) continue;
    double val[NFT];
    combine_corners<DIM, NFT, T>(w, ld, i, acc, val);
    {
      const double xd[3] = {(double)xp, (double)yp, (double)zp};
      store_value<DIM, NFT>(g, ci, cj, ck, px, sizeof(T) == 8 ? xd : nullptr, ld, i, w, ev, val, nfields, out);
    }
  }
}

template <int DIM, int KT, int NFT, typename T>
static void launch_sparse_g(int G, const Grid& g, const double* px, int64_t ld, const int32_t* perm, const int32_t* pstart,
                            const double* w, const T* blk, const int32_t* cs, int nfields, double* out,
                            const int32_t* cell_list, const int32_t* list_count, int np, const EvalExt& ev,
                            hipStream_t s) {
  // waves: one per 64/G listed cells (at most min(C, np) cells are non-empty)
  const int ncell = std::min(g.C, np);
  const int waves = (ncell + 64 / G - 1) / (64 / G);
  const dim3 grid(((waves + 3) / 4 + 7) / 8 * 8), blkd(256);
#define LRB_SP(GG) LRB_LAUNCH_TIMED((k_eval_sparse<DIM, KT, NFT, T, GG>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs, nfields, out, cell_list, list_count, ev)
  if (G == 8) LRB_SP(8);
  else if (G == 16) LRB_SP(16);
  else LRB_SP(32);
#undef LRB_SP
}

template <int DIM, int KT, typename T>
static void eval_blk_k(const Grid& g, const double* px, int64_t ld, const int32_t* perm, const int32_t* pstart,
                       const double* w, const T* blk, int nft, const int32_t* cs, int nfields, double* out, int block,
                       int np, const int32_t* cell_list, const int32_t* list_count, int group, int pipe, int waves,
                       const EvalExt& ev, hipStream_t s) {
  if (cell_list && group > 0 && nft == DIM) {   // group < 0: listed cells dealt round-robin
    // sparse velocity evaluation: lane groups of G particles, 64/G cells per wave
    launch_sparse_g<DIM, KT, DIM, T>(group, g, px, ld, perm, pstart, w, blk, cs, nfields, out, cell_list, list_count, np,
                                     ev, s);
    return;
  }
  // dense: one block per cell; listed cells: one block per listed non-empty cell
  const int nblk = cell_list ? (std::min(g.C, np) + 7) / 8 * 8 : (g.C + 7) / 8 * 8;
  const dim3 grid(nblk), blkd(block);
  // group 0: contiguous XCD runs; group -K: round-robin in runs of K listed cells
  const int deal = group >= 0 ? 0 : -group;
#define LRB_EB(NFT, P) LRB_LAUNCH_TIMED((k_eval_blk<DIM, KT, NFT, T, P, default_waves<T>()>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs, nfields, out, cell_list, list_count, deal, ev)
  if constexpr (sizeof(T) == 4) {
    // matrix-core accumulation: up to 4 fields
    if (pipe == 3 && nft <= 4) {
      switch (nft) {
        case 1: LRB_LAUNCH_TIMED((k_eval_blk<DIM, KT, 1, T, 3, LRB_MFMA_WAVES>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs, nfields, out, cell_list, list_count, deal, ev); break;
        case 2: LRB_LAUNCH_TIMED((k_eval_blk<DIM, KT, 2, T, 3, LRB_MFMA_WAVES>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs, nfields, out, cell_list, list_count, deal, ev); break;
        case 3: LRB_LAUNCH_TIMED((k_eval_blk<DIM, KT, 3, T, 3, LRB_MFMA_WAVES>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs, nfields, out, cell_list, list_count, deal, ev); break;
        default: LRB_LAUNCH_TIMED((k_eval_blk<DIM, KT, 4, T, 3, LRB_MFMA_WAVES>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs, nfields, out, cell_list, list_count, deal, ev); break;
      }
      return;
    }
    // two rows per wait: the velocity evaluations (nft = dim) only
    if (pipe == 2 && nft == DIM) {
      if (waves == 6)
        LRB_LAUNCH_TIMED((k_eval_blk<DIM, KT, DIM, T, 2, 6>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs,
                         nfields, out, cell_list, list_count, deal, ev);
      else
        LRB_EB(DIM, 2);
      return;
    }
  }
  switch (nft) {
    case 1: LRB_EB(1, 1); break;
    case 2: LRB_EB(2, 1); break;
    case 3: LRB_EB(3, 1); break;
    case 4: LRB_EB(4, 1); break;
    default:
      // six fields, FP32: two 3D velocity snapshots in one pass (48 accumulators per
      // lane: 6 waves per SIMD, no spills)
      if constexpr (DIM == 3 && sizeof(T) == 4)
        LRB_LAUNCH_TIMED((k_eval_blk<DIM, KT, 6, T, 1, 6>), grid, blkd, s, g, px, ld, perm, pstart, w, blk, cs,
                         nfields, out, cell_list, list_count, deal, ev);
      break;
  }
#undef LRB_EB
}

// Every launch shape and field count of one kernel type, FP64 and (where the kernel
// has an FP32 form) FP32 tables; instantiated once per kernel type, each in a
// translation unit of its own (eval_blk_k*.hip).
template <int KT>
void eval_blk_for_kernel(const Grid& g, bool fp32, const double* px, int64_t ld, const int32_t* perm,
                         const int32_t* pstart, const double* w, const void* blk_, int nft, const int32_t* cs,
                         int nfields, double* out, int block, int np, const int32_t* cell_list,
                         const int32_t* list_count, int group, int pipe, int waves, const EvalExt& ev, hipStream_t s) {
#define LRB_ARGS px, ld, perm, pstart, w, blk, nft, cs, nfields, out, block, np, cell_list, list_count, group, pipe, waves, ev, s
  if constexpr (KT == K_WC0 || KT == K_WC2 || KT == K_IMQ || KT == K_GAUSS || KT == K_PHS3 || KT == K_GAUSS_LOCAL) {
    if (fp32) {
      const float* blk = (const float*)blk_;
      if (g.dim == 3) eval_blk_k<3, KT, float>(g, LRB_ARGS);
      else eval_blk_k<2, KT, float>(g, LRB_ARGS);
      return;
    }
  }
  const double* blk = (const double*)blk_;
  if (g.dim == 3) eval_blk_k<3, KT, double>(g, LRB_ARGS);
  else eval_blk_k<2, KT, double>(g, LRB_ARGS);
#undef LRB_ARGS
}

#define LRB_EVAL_BLK_DECL(KT)                                                                                        \
  extern template void eval_blk_for_kernel<KT>(const Grid&, bool, const double*, int64_t, const int32_t*,            \
                                               const int32_t*, const double*, const void*, int, const int32_t*, int, \
                                               double*, int, int, const int32_t*, const int32_t*, int, int, int,     \
                                               const EvalExt&, hipStream_t);
#define LRB_EVAL_BLK_INST(KT)                                                                                 \
  template void eval_blk_for_kernel<KT>(const Grid&, bool, const double*, int64_t, const int32_t*,            \
                                        const int32_t*, const double*, const void*, int, const int32_t*, int, \
                                        double*, int, int, const int32_t*, const int32_t*, int, int, int,     \
                                        const EvalExt&, hipStream_t);
LRB_EVAL_BLK_DECL(K_LINEAR)
LRB_EVAL_BLK_DECL(K_WC0)
LRB_EVAL_BLK_DECL(K_WC2)
LRB_EVAL_BLK_DECL(K_BUMP)
LRB_EVAL_BLK_DECL(K_MQ)
LRB_EVAL_BLK_DECL(K_IMQ)
LRB_EVAL_BLK_DECL(K_GAUSS)
LRB_EVAL_BLK_DECL(K_PHS3)
LRB_EVAL_BLK_DECL(K_GAUSS_LOCAL)

}  // namespace lrb
