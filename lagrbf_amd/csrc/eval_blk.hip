// Cell-ordered PU evaluation over a cell-major node table (interpolation.c:39-154
// restated for the GPU).
//
// Table: one row per data node, in cell-sorted order (row s = s-th entry of the
// per-cell node lists), BS = 4 + 2^d * nft elements:
//   [x, y, (z), pad | coefficients of the node in the patches of the 2^d corners
//    of its cell, corner order Q1, nft fields each]
// (FP32 rows: [-2 x, -2 y, (-2 z), x^2 + y^2 (+ z^2) | ...], x, y, z relative to
// the centre of the node's cell.)
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
#include "eval_blk_impl.hpp"

#include <hip/hip_ext.h>

#include <algorithm>

namespace lrb {

thread_local hipEvent_t t_eval_ev0 = nullptr, t_eval_ev1 = nullptr;

void set_eval_timing_events(hipEvent_t start, hipEvent_t stop) {
  t_eval_ev0 = start;
  t_eval_ev1 = stop;
}

// Table row s: node cn[s] of cell q = dcell[cn[s]], r = s - cs[q] its rank in q;
// rows N .. N + BLK_PAD_ROWS - 1 are zero padding.
// Coefficient of corner kp: patch p = node kp of q; member index = (nodes of the
// cells of p before q, ascending cell index) + r.
template <typename T>
__global__ void k_build_blk(Grid g, const double* __restrict__ x, int N, const int32_t* __restrict__ cn,
                            const int32_t* __restrict__ dcell, const int32_t* __restrict__ cs,
                            const uint8_t* __restrict__ active, const int64_t* __restrict__ mem_off,
                            const double* __restrict__ coef, int nft, int cstride, int f0, double cscale,
                            T* __restrict__ blk) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N + BLK_PAD_ROWS) return;
  const int dim = g.dim, BS = 4 + g.nen * nft;
  if (s >= N) {
    for (int e = 0; e < BS; e++) blk[(size_t)s * BS + e] = T(0);
    return;
  }
  const int j = cn[s], q = dcell[j], r = s - cs[q];
  const int qi = q % g.nc[0], qj = (q / g.nc[0]) % g.nc[1], qk = dim == 3 ? q / (g.nc[0] * g.nc[1]) : 0;
  const int qq[3] = {qi, qj, qk};
  T* row = blk + (size_t)s * BS;
  if (sizeof(T) == 4) {
    // -2 n and |n|^2 of the node n relative to its cell's centre, |n|^2 from the
    // rounded coordinates
    double n2 = 0.0;
    for (int d = 0; d < 3; d++) {
      float v = 0.f;
      if (d < dim) v = (float)(x[(size_t)j * dim + d] - (g.O[d] + (qq[d] + 0.5) * g.h[d]));
      row[d] = (T)(-2.f * v);
      n2 += (double)v * (double)v;
    }
    row[3] = (T)n2;
  } else {
    for (int d = 0; d < 4; d++) row[d] = (T)(d < dim ? x[(size_t)j * dim + d] : 0.0);
  }
  for (int kp = 0; kp < g.nen; kp++) {
    int a, b, c;
    corner_offset(kp, a, b, c);
    const int pi = qi + a, pj = qj + b, pk = qk + c;
    const int p = pi + g.n[0] * (pj + g.n[1] * pk);
    T* dst = row + 4 + kp * nft;
    if (!active[p]) {
      for (int f = 0; f < nft; f++) dst[f] = T(0);
      continue;
    }
    int off = 0;
    for (int e2 = 0; e2 <= (dim == 3 ? 1 : 0); e2++)
      for (int e1 = 0; e1 <= 1; e1++)
        for (int e0 = 0; e0 <= 1; e0++) {
          const int ii = pi - 1 + e0, jj = pj - 1 + e1, kk = dim == 3 ? pk - 1 + e2 : 0;
          if (ii < 0 || ii >= g.nc[0] || jj < 0 || jj >= g.nc[1] || (dim == 3 && (kk < 0 || kk >= g.nc[2]))) continue;
          const int c2 = ii + g.nc[0] * (jj + g.nc[1] * kk);
          if (c2 < q) off += cs[c2 + 1] - cs[c2];
        }
    const double* src = coef + (size_t)(mem_off[p] + off + r) * cstride + f0;
    for (int f = 0; f < nft; f++) dst[f] = (T)(src[f] * cscale);
  }
}

// ---------------------------------------------------------------------------
// Non-empty particle cells of a grouped swarm, in ascending cell order: per-block
// counts, an exclusive scan of the counts, then an ordered write.  `list` holds C
// entries followed by ceil(C / 256) + 1 entries of scratch (the block offsets).
__global__ __launch_bounds__(256) void k_ne_count(const int32_t* __restrict__ pstart, int C,
                                                  int32_t* __restrict__ bcount) {
  __shared__ int32_t wc[4];
  const int c = blockIdx.x * 256 + threadIdx.x;
  const bool ne = c < C && pstart[c + 1] > pstart[c];
  const unsigned long long m = __ballot(ne);
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (int32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) bcount[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// One block: exclusive scan of nb block counts in place, total into *count.
__global__ __launch_bounds__(256) void k_ne_scan(int32_t* __restrict__ bcount, int nb, int32_t* __restrict__ count) {
  __shared__ int32_t sh[256];
  const int t = threadIdx.x;
  int32_t run = 0;
  for (int base = 0; base < nb; base += 256) {
    const int32_t v = base + t < nb ? bcount[base + t] : 0;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const int32_t x = t >= o ? sh[t - o] : 0;
      __syncthreads();
      sh[t] += x;
      __syncthreads();
    }
    if (base + t < nb) bcount[base + t] = run + sh[t] - v;
    run += sh[255];
    __syncthreads();
  }
  if (t == 0) *count = run;
}

__global__ __launch_bounds__(256) void k_ne_write(const int32_t* __restrict__ pstart, int C,
                                                  const int32_t* __restrict__ boff, int32_t* __restrict__ list) {
  __shared__ int32_t wc[4];
  const int c = blockIdx.x * 256 + threadIdx.x;
  const bool ne = c < C && pstart[c + 1] > pstart[c];
  const unsigned long long m = __ballot(ne);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) wc[wv] = (int32_t)__popcll(m);
  __syncthreads();
  int off = boff[blockIdx.x];
  for (int k = 0; k < wv; k++) off += wc[k];
  if (ne) list[off + (int)__popcll(m & ((1ull << lane) - 1ull))] = c;
}

// Kernels only: the sequence may sit in a captured graph.
void launch_nonempty_cells(const int32_t* pstart, int C, int32_t* list, int32_t* count, hipStream_t s) {
  if (C <= 0) {
    launch_fill_i32(count, 1, 0, s);
    return;
  }
  const int nb = (C + 255) / 256;
  int32_t* boff = list + C;
  hipLaunchKernelGGL(k_ne_count, dim3(nb), dim3(256), 0, s, pstart, C, boff);
  hipLaunchKernelGGL(k_ne_scan, dim3(1), dim3(256), 0, s, boff, nb, count);
  hipLaunchKernelGGL(k_ne_write, dim3(nb), dim3(256), 0, s, pstart, C, boff, list);
}

size_t nonempty_list_capacity(int C) { return (size_t)std::max(C, 1) + (size_t)(std::max(C, 1) + 255) / 256 + 1; }

// six fields (two 3D velocity snapshots in one pass) in FP32 only: the FP64 kernel
// would need 96 accumulator registers per lane
bool eval_blk_supported(int nft, int dim, bool f32) { return (nft >= 1 && nft <= 4) || (nft == 6 && dim == 3 && f32); }
size_t blk_row(int dim, int nft) { return 4 + ((size_t)1 << dim) * (size_t)nft; }

void launch_build_blk(const Grid& g, bool fp32, const double* x, int N, const int32_t* cn, const int32_t* dcell,
                      const int32_t* cs, const uint8_t* active, const int64_t* mem_off, const double* coef, int nft,
                      int cstride, int f0, double cscale, void* blk, hipStream_t s) {
  if (N <= 0) return;
  // N + BLK_PAD_ROWS threads: rows 0..N-1 and the zero padding rows
  const int nt = N + BLK_PAD_ROWS;
  if (fp32)
    hipLaunchKernelGGL(k_build_blk<float>, dim3((nt + 255) / 256), dim3(256), 0, s, g, x, N, cn, dcell, cs, active,
                       mem_off, coef, nft, cstride, f0, cscale, (float*)blk);
  else
    hipLaunchKernelGGL(k_build_blk<double>, dim3((nt + 255) / 256), dim3(256), 0, s, g, x, N, cn, dcell, cs, active,
                       mem_off, coef, nft, cstride, f0, cscale, (double*)blk);
}

void launch_eval_blk(const Grid& g, int ktype, bool fp32, const double* px, int64_t ld, const int32_t* perm,
                     const int32_t* pstart, const double* w, const void* blk, int nft, const int32_t* cs, int nfields,
                     double* out, int block, int np, const int32_t* cell_list, const int32_t* list_count,
                     int group, int pipe, int waves, const EvalExt& ev, hipStream_t s) {
  if (nfields <= 0 || g.C <= 0 || np <= 0) return;
#define LRB_K(K) eval_blk_for_kernel<K>(g, fp32, px, ld, perm, pstart, w, blk, nft, cs, nfields, out, block, np, cell_list, list_count, group, pipe, waves, ev, s)
  switch (ktype) {
    case K_LINEAR: LRB_K(K_LINEAR); break;
    case K_WC0: LRB_K(K_WC0); break;
    case K_WC2: LRB_K(K_WC2); break;
    case K_BUMP: LRB_K(K_BUMP); break;
    case K_MQ: LRB_K(K_MQ); break;
    case K_IMQ: LRB_K(K_IMQ); break;
    case K_PHS3: LRB_K(K_PHS3); break;
    case K_GAUSS_LOCAL: LRB_K(K_GAUSS_LOCAL); break;
    default: LRB_K(K_GAUSS); break;
  }
#undef LRB_K
}

}  // namespace lrb
