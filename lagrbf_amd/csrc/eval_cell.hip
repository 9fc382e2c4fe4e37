// Cell-ordered PU evaluation (interpolation.c:39-154 restated for the GPU).
//
// Particles are grouped by cell (perm / pstart).  One workgroup per cell, one
// lane per particle.  The 3^d neighbour cells are visited in ascending cell index
// and the nodes of a cell in ascending node index, so phi(|x - x_j|) is computed
// once per (particle, node) and accumulated into every corner patch that holds
// the node's cell -- which visits each patch's members exactly in member order
// (patch = concatenation of its cells' node lists, partition.c:337-371).  Node
// coordinates and coefficients are wave-uniform loads (scalar cache, SGPR
// operands); each neighbour cell runs a loop specialised at compile time for the
// corners that contain it.
//
// T = double: absolute FP64 coordinates.  T = float (mixed precision): particle
// and node coordinates relative to their cell origin, shifted in FP64 before the
// rounding to FP32, FP32 coefficients, FP64 Shepard combination.
#include "kernels.hpp"

#include <cmath>

namespace lrb {

template <int DIM>
__host__ __device__ constexpr int nbr_di(int nbi) { return nbi % 3 - 1; }
template <int DIM>
__host__ __device__ constexpr int nbr_dj(int nbi) { return (nbi / 3) % 3 - 1; }
template <int DIM>
__host__ __device__ constexpr int nbr_dk(int nbi) { return DIM == 3 ? nbi / 9 - 1 : 0; }

// Does corner k's patch contain the neighbour cell nbi, and in which slot?
template <int DIM>
__host__ __device__ constexpr int corner_slot(int nbi, int k) {
  const int oi = ((k & 3) == 1 || (k & 3) == 2) ? 1 : 0, oj = ((k & 3) >= 2) ? 1 : 0, ok = k >> 2;
  const int sa = nbr_di<DIM>(nbi) - oi + 1, sb = nbr_dj<DIM>(nbi) - oj + 1, sc = nbr_dk<DIM>(nbi) - ok + 1;
  const bool has = sa >= 0 && sa <= 1 && sb >= 0 && sb <= 1 && (DIM == 2 || (sc >= 0 && sc <= 1));
  return has ? sa + 2 * sb + (DIM == 3 ? 4 * sc : 0) : -1;
}

template <int KT>
__host__ __device__ constexpr bool compact_support() { return KT == K_WC0 || KT == K_WC2 || KT == K_BUMP; }

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int DIM, int KT, int NF, typename T, int NBI>
__device__ __forceinline__ void accum_cell(T xp, T yp, T zp, const T* __restrict__ xs, int N, int nb, int ne, T sx,
                                           T sy, T sz, const T* const (&Xk)[1 << DIM], int nft, T (&acc)[1 << DIM][NF],
                                           T eps, T rcut2) {
  constexpr int NEN = 1 << DIM;
  for (int j = nb; j < ne; j++) {
    T d0, d1, d2;
    if constexpr (sizeof(T) == 4) {
      d0 = xp - (xs[j] + sx); d1 = yp - (xs[(size_t)N + j] + sy);
      d2 = DIM == 3 ? zp - (xs[2 * (size_t)N + j] + sz) : T(0);
    } else {
      d0 = xp - xs[j]; d1 = yp - xs[(size_t)N + j];
      d2 = DIM == 3 ? zp - xs[2 * (size_t)N + j] : T(0);
    }
    const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if constexpr (compact_support<KT>()) {
      // every lane beyond the support: phi = 0 exactly, nothing to add
      if (!__any(r2 < rcut2)) continue;
    }
    const T phi = rbf<KT, T>(sqrt_fast(r2), eps);
    const int r = j - nb;
#pragma unroll
    for (int k = 0; k < NEN; k++) {
      if (corner_slot<DIM>(NBI, k) >= 0) {
        const T* Xr = Xk[k] + (size_t)r * nft;
#pragma unroll
        for (int f = 0; f < NF; f++) acc[k][f] += phi * Xr[f];
      }
    }
  }
}

template <int DIM, int KT, int NF, typename T>
__global__ __launch_bounds__(256) void k_eval_cell(Grid g, const double* __restrict__ px, int64_t ld,
                                                   const int32_t* __restrict__ perm, const int32_t* __restrict__ pstart,
                                                   const double* __restrict__ w, const uint8_t* __restrict__ active,
                                                   const int64_t* __restrict__ mem_off, const T* __restrict__ coef,
                                                   int nft, const T* __restrict__ zeros, const T* __restrict__ xs, int N,
                                                   const int32_t* __restrict__ cs, int nfields, int f0,
                                                   double* __restrict__ out) {
  constexpr int NEN = 1 << DIM;
  constexpr int NNB = DIM == 3 ? 27 : 9;
  __shared__ int s_off[NEN][NEN];
  __shared__ long long s_base[NEN];  // element offset of corner k's coefficient rows, -1: inactive patch
  // XCD-aware order: block b runs on XCD b % 8, which gets a contiguous range of cells
  const int C = g.C;
  const int per = (gridDim.x + 7) / 8;
  const int b = blockIdx.x;
  const int c = (b % 8) * per + b / 8;
  if (c >= C) return;
  const int pb = pstart[c], pe = pstart[c + 1];
  if (pb == pe) return;
  const int ci = c % g.nc[0], cj = (c / g.nc[0]) % g.nc[1], ck = DIM == 3 ? c / (g.nc[0] * g.nc[1]) : 0;
  const int tid = threadIdx.x;
  if (tid < NEN * NEN) {
    // corner k's patch (node cell + o_k): member offset of its cell in slot q
    const int k = tid / NEN, q = tid % NEN;
    int oi, oj, ok;
    corner_offset(k, oi, oj, ok);
    const int ni = ci + oi, nj = cj + oj, nk = ck + ok;
    int off = 0;
    for (int qq = 0; qq < q; qq++) {
      const int a = qq & 1, bb = (qq >> 1) & 1, cc = qq >> 2;
      const int ii = ni - 1 + a, jj = nj - 1 + bb, kk = DIM == 3 ? nk - 1 + cc : 0;
      bool valid = ii >= 0 && ii < g.nc[0] && jj >= 0 && jj < g.nc[1];
      if (DIM == 3) valid = valid && kk >= 0 && kk < g.nc[2];
      if (valid) { const int c2 = ii + g.nc[0] * (jj + g.nc[1] * kk); off += cs[c2 + 1] - cs[c2]; }
    }
    s_off[k][q] = off;
    if (q == 0) {
      const int p = ni + g.n[0] * (nj + g.n[1] * nk);
      s_base[k] = active[p] ? (long long)mem_off[p] * nft + f0 : -1ll;
    }
  }
  __syncthreads();
  // an empty 3^d neighbourhood leaves the particles' values untouched (interpolation.c:89)
  int nnb = 0;
  for (int nbi = 0; nbi < NNB; nbi++) {
    const int a = ci + nbi % 3 - 1, bb = cj + (nbi / 3) % 3 - 1, e = ck + (DIM == 3 ? nbi / 9 - 1 : 0);
    if (a < 0 || a >= g.nc[0] || bb < 0 || bb >= g.nc[1] || (DIM == 3 && (e < 0 || e >= g.nc[2]))) continue;
    const int q = a + g.nc[0] * (bb + g.nc[1] * e);
    nnb += cs[q + 1] - cs[q];
  }
  if (nnb == 0) return;
  const T eps = (T)g.eps_r;
  const double rc = 1.0 / g.eps_r;
  const T rcut2 = (T)(rc * rc * (sizeof(T) == 4 ? (1.0 + 1e-5) : (1.0 + 1e-12)));
  // FP32: coordinates relative to the particle cell's origin
  const double ox = sizeof(T) == 4 ? g.O[0] + ci * g.h[0] : 0.0;
  const double oy = sizeof(T) == 4 ? g.O[1] + cj * g.h[1] : 0.0;
  const double oz = sizeof(T) == 4 && DIM == 3 ? g.O[2] + ck * g.h[2] : 0.0;
  // corner bases as wave-uniform (SGPR) values: the coefficient rows are then
  // read through the scalar cache, like the node coordinates
  long long base[NEN];
#pragma unroll
  for (int k = 0; k < NEN; k++) {
    const long long v = s_base[k];
    const int lo = uniform_int((int)(v & 0xffffffffll)), hi = uniform_int((int)(v >> 32));
    base[k] = (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
  }
  for (int s = pb + tid; s - tid < pe; s += blockDim.x) {
    const bool live = s < pe;
    const int i = live ? perm[s] : perm[pb];
    const T xp = (T)(px[i] - ox), yp = (T)(px[ld + i] - oy), zp = DIM == 3 ? (T)(px[2 * ld + i] - oz) : T(0);
    T acc[NEN][NF];
#pragma unroll
    for (int k = 0; k < NEN; k++)
#pragma unroll
      for (int f = 0; f < NF; f++) acc[k][f] = T(0);
#pragma unroll 1
    for (int nbi = 0; nbi < NNB; nbi++) {
      const int di = nbi % 3 - 1, dj = (nbi / 3) % 3 - 1, dk = DIM == 3 ? nbi / 9 - 1 : 0;
      const int a = ci + di, bb = cj + dj, e = ck + dk;
      if (a < 0 || a >= g.nc[0] || bb < 0 || bb >= g.nc[1] || (DIM == 3 && (e < 0 || e >= g.nc[2]))) continue;
      const int q = a + g.nc[0] * (bb + g.nc[1] * e);
      const int nb = cs[q], ne = cs[q + 1];
      if (nb == ne) continue;
      const T sx = sizeof(T) == 4 ? (T)(di * g.h[0]) : T(0);
      const T sy = sizeof(T) == 4 ? (T)(dj * g.h[1]) : T(0);
      const T sz = sizeof(T) == 4 && DIM == 3 ? (T)(dk * g.h[2]) : T(0);
      const T* Xk[NEN];
#pragma unroll
      for (int k = 0; k < NEN; k++) {
        int oi, oj, ok;
        corner_offset(k, oi, oj, ok);
        const int sa = di - oi + 1, sb = dj - oj + 1, sc = dk - ok + 1;
        const bool has = sa >= 0 && sa <= 1 && sb >= 0 && sb <= 1 && (DIM == 2 || (sc >= 0 && sc <= 1));
        const int slot = has ? uniform_int(s_off[k][sa + 2 * sb + (DIM == 3 ? 4 * sc : 0)]) : 0;
        Xk[k] = (has && base[k] >= 0) ? coef + base[k] + (long long)slot * nft : zeros;
      }
      switch (nbi) {
#define LRB_CASE(V) case V: accum_cell<DIM, KT, NF, T, V>(xp, yp, zp, xs, N, nb, ne, sx, sy, sz, Xk, nft, acc, eps, rcut2); break;
        LRB_CASE(0) LRB_CASE(1) LRB_CASE(2) LRB_CASE(3) LRB_CASE(4) LRB_CASE(5) LRB_CASE(6) LRB_CASE(7) LRB_CASE(8)
        LRB_CASE(9) LRB_CASE(10) LRB_CASE(11) LRB_CASE(12) LRB_CASE(13) LRB_CASE(14) LRB_CASE(15) LRB_CASE(16)
        LRB_CASE(17) LRB_CASE(18) LRB_CASE(19) LRB_CASE(20) LRB_CASE(21) LRB_CASE(22) LRB_CASE(23) LRB_CASE(24)
        LRB_CASE(25) LRB_CASE(26)
#undef LRB_CASE
        default: break;
      }
    }
    if (!live) continue;
    double val[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) val[f] = 0.0;
#pragma unroll
    for (int k = 0; k < NEN; k++) {
      const double wk = w[(size_t)k * ld + i];
#pragma unroll
      for (int f = 0; f < NF; f++) val[f] += (double)acc[k][f] * wk;
    }
    const int nfc = min(NF, nfields - f0);
#pragma unroll
    for (int f = 0; f < NF; f++) if (f < nfc) out[(size_t)(f0 + f) * ld + i] = val[f];
  }
}

template <int DIM, int KT, typename T>
static void eval_cell_k(const Grid& g, const double* px, int64_t ld, const int32_t* perm, const int32_t* pstart,
                        const double* w, const uint8_t* active, const int64_t* mem_off, const T* coef, int nft,
                        const T* zeros, const T* xs, int N, const int32_t* cs, int nfields, double* out, int block,
                        hipStream_t s) {
  const int per = (g.C + 7) / 8;
  const dim3 grid(per * 8), blk(block);
  for (int f0 = 0; f0 < nfields; f0 += 4) {
    const int r = nfields - f0;
#define LRB_EC(NF) hipLaunchKernelGGL((k_eval_cell<DIM, KT, NF, T>), grid, blk, 0, s, g, px, ld, perm, pstart, w, active, mem_off, coef, nft, zeros, xs, N, cs, nfields, f0, out)
    if (r == 1) LRB_EC(1);
    else if (r == 2) LRB_EC(2);
    else if (r == 3) LRB_EC(3);
    else LRB_EC(4);
#undef LRB_EC
  }
}

template <int DIM, typename T>
static void eval_cell_dim(const Grid& g, int ktype, const double* px, int64_t ld, const int32_t* perm,
                          const int32_t* pstart, const double* w, const uint8_t* active, const int64_t* mem_off,
                          const T* coef, int nft, const T* zeros, const T* xs, int N, const int32_t* cs, int nfields,
                          double* out, int block, hipStream_t s) {
#define LRB_K(K) eval_cell_k<DIM, K, T>(g, px, ld, perm, pstart, w, active, mem_off, coef, nft, zeros, xs, N, cs, nfields, out, block, s)
  if constexpr (sizeof(T) == 8) {
    switch (ktype) {
      case K_LINEAR: LRB_K(K_LINEAR); break;
      case K_WC0: LRB_K(K_WC0); break;
      case K_WC2: LRB_K(K_WC2); break;
      case K_BUMP: LRB_K(K_BUMP); break;
      case K_MQ: LRB_K(K_MQ); break;
      case K_IMQ: LRB_K(K_IMQ); break;
      default: LRB_K(K_GAUSS); break;
    }
  } else {
    switch (ktype) {
      case K_WC0: LRB_K(K_WC0); break;
      case K_WC2: LRB_K(K_WC2); break;
      case K_IMQ: LRB_K(K_IMQ); break;
      default: LRB_K(K_GAUSS); break;
    }
  }
#undef LRB_K
}

// FP32 evaluation: the positive definite kernels, r^3 (always with a polynomial
// block) and the Gaussian with per-patch shape (cell-table kernels only)
bool eval_f32_supported(int ktype) {
  return ktype == K_WC0 || ktype == K_WC2 || ktype == K_IMQ || ktype == K_GAUSS || ktype == K_PHS3 ||
         ktype == K_GAUSS_LOCAL;
}

void launch_eval_cell(const Grid& g, int ktype, bool fp32, const double* px, int64_t ld, const int32_t* perm,
                      const int32_t* pstart, const double* w, const uint8_t* active, const int64_t* mem_off,
                      const void* coef, int nft, const void* zeros, const void* xs, int N, const int32_t* cs,
                      int nfields, double* out, int block, hipStream_t s) {
  if (nfields <= 0 || g.C <= 0) return;
  if (fp32) {
    if (g.dim == 3) eval_cell_dim<3, float>(g, ktype, px, ld, perm, pstart, w, active, mem_off, (const float*)coef, nft, (const float*)zeros, (const float*)xs, N, cs, nfields, out, block, s);
    else eval_cell_dim<2, float>(g, ktype, px, ld, perm, pstart, w, active, mem_off, (const float*)coef, nft, (const float*)zeros, (const float*)xs, N, cs, nfields, out, block, s);
  } else {
    if (g.dim == 3) eval_cell_dim<3, double>(g, ktype, px, ld, perm, pstart, w, active, mem_off, (const double*)coef, nft, (const double*)zeros, (const double*)xs, N, cs, nfields, out, block, s);
    else eval_cell_dim<2, double>(g, ktype, px, ld, perm, pstart, w, active, mem_off, (const double*)coef, nft, (const double*)zeros, (const double*)xs, N, cs, nfields, out, block, s);
  }
}

// FP32 node coordinates relative to the origin of the node's own cell (cell-sorted SoA)
__global__ void k_sorted_coords32(Grid g, const double* __restrict__ x, int n, const int32_t* __restrict__ cn,
                                  const int32_t* __restrict__ cell, float* __restrict__ xs) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int j = cn[s];
  const int c = cell[j];
  const int ijk[3] = {c % g.nc[0], (c / g.nc[0]) % g.nc[1], g.dim == 3 ? c / (g.nc[0] * g.nc[1]) : 0};
  for (int d = 0; d < g.dim; d++) xs[(size_t)d * n + s] = (float)(x[(size_t)j * g.dim + d] - (g.O[d] + ijk[d] * g.h[d]));
}

void launch_sorted_coords32(const Grid& g, const double* x, int n, const int32_t* cn, const int32_t* cell, float* xs,
                            hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sorted_coords32, dim3((n + 255) / 256), dim3(256), 0, s, g, x, n, cn, cell, xs);
}

// FP32 copy of the (FP64-refined) coefficients of the listed patches
__global__ void k_coef_to_f32(const int32_t* __restrict__ list, const int32_t* __restrict__ pn,
                              const int64_t* __restrict__ mem_off, int nf, const double* __restrict__ coef,
                              float* __restrict__ coef32) {
  const int p = list[blockIdx.x];
  const size_t base = (size_t)mem_off[p] * nf, m = (size_t)pn[p] * nf;
  for (size_t t = threadIdx.x; t < m; t += blockDim.x) coef32[base + t] = (float)coef[base + t];
}

void launch_coef_to_f32(const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off, int nf,
                        const double* coef, float* coef32, hipStream_t s) {
  if (nlist > 0 && nf > 0) hipLaunchKernelGGL(k_coef_to_f32, dim3(nlist), dim3(256), 0, s, list, pn, mem_off, nf, coef, coef32);
}

}  // namespace lrb
