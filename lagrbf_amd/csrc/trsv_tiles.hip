// Triangular solves L L^T u = h over many workgroups for the B22 factors of the
// large saddle-point patches (factor dimension above 512, the tiled Cholesky of
// chol_tiled.hip).  One workgroup per patch (k_solve_aug) streams a factor of
// thousands of columns through one compute unit, one load round trip after
// another; here the factor is cut into the tiles of 64 rows and columns the
// factor was built in, and one launch per tile column covers every listed patch:
//
//   forward, column tile k: task (patch, i), i > k: y_i -= L_ik z_k, z_k the
//     solved tile k; the task of tile k + 1 then solves its diagonal tile,
//     z_{k+1} = L_{k+1,k+1}^-1 y_{k+1} (every column before it is applied).  A
//     first launch solves the diagonal tile 0.
//   backward, step s: for a patch of T tiles, kp = T - 1 - s holds the final x_kp;
//     task (patch, j), j < kp: z_j -= L_{kp,j}^T x_kp; the task of j = kp - 1 then
//     solves x_j = L_jj^-T z_j.  A first launch solves the last diagonal tile.
//
// Diagonal tiles are solved in LDS by blocks of 16 with the inverses the factor
// keeps in the upper triangle of its 16 x 16 diagonal blocks (element (J + c, J +
// r), r > c, holds (L_JJ^-1)[r][c]).  The vectors are those of k_solve_aug's
// workspace: y of patch p at ws + mem_off (3 + 3 nf) + (3 + nf) n, field stride n,
// B22 rows from m on.  Each task's sums run in a fixed order, so the result does
// not depend on how the tasks are scheduled.
#include "kernels.hpp"

#include <algorithm>
#include <vector>

namespace lrb {

static constexpr int XNT = 256;   // threads per task
static constexpr int XT = 64;     // tile
static constexpr int XF = 8;      // fields

template <typename TL>
struct TrsvView {
  const TL* L;   // B22 factor, column-major, leading dimension nn
  double* y;     // B22 part of the workspace vector, field stride ld
  int nn, ld;
};

template <typename TL>
__device__ __forceinline__ TrsvView<TL> trsv_view(int p, const int32_t* pn, const AugView& a, const int64_t* mem_off,
                                                  const int64_t* fac_off, const TL* fac, double* ws, int nf) {
  TrsvView<TL> v;
  const int n = pn[p], m = a.m[p];
  v.nn = n - m;
  v.ld = n;
  v.L = fac + fac_off[p];
  v.y = ws + (size_t)mem_off[p] * (3 + 3 * (size_t)nf) + (3 + (size_t)nf) * n + m;
  return v;
}

// (L_JJ^-1)[r][c] into lb[r * 17 + c] for the w <= 16 columns at J (zero elsewhere)
template <typename TL>
__device__ __forceinline__ void trsv_inverse(const TrsvView<TL>& v, int J, int w, double* lb) {
  const int t = threadIdx.x;
  if (t < 256) {
    const int r = t >> 4, c = t & 15;
    double e = 0.0;
    if (r < w && c <= r) {
      const double d = (double)v.L[(size_t)(J + r) * v.nn + J + c];
      e = r == c ? 1.0 / d : d;
    }
    lb[r * 17 + c] = e;
  }
}

// z_o = L_oo^-1 y_o for the diagonal tile at row/column o (w rows), in LDS ys[r * XF + f]
template <typename TL>
__device__ __forceinline__ void trsv_diag_fwd(const TrsvView<TL>& v, int o, int w, int nf, double* ys, double* lb, double* zb) {
  const int tid = threadIdx.x;
  for (int e = tid; e < XT * XF; e += XNT) {
    const int r = e / XF, f = e % XF;
    ys[e] = (r < w && f < nf) ? v.y[(size_t)f * v.ld + o + r] : 0.0;
  }
  for (int b = 0; b * 16 < w; b++) {
    const int J = 16 * b, wb = min(16, w - J);
    __syncthreads();
    trsv_inverse<TL>(v, o + J, wb, lb);
    __syncthreads();
    if (tid < 16 * XF) {
      const int r = tid / XF, f = tid % XF;
      double z = 0.0;
#pragma unroll
      for (int c = 0; c < 16; c++) z += lb[r * 17 + c] * ys[(J + c) * XF + f];   // lb is zero past the block
      zb[r * XF + f] = z;
    }
    __syncthreads();
    if (tid < 16 * XF) {
      const int r = tid / XF, f = tid % XF;
      if (r < wb) ys[(J + r) * XF + f] = zb[r * XF + f];
    }
    // rows of the tile below the block
    for (int e = tid; e < (w - J - wb) * XF; e += XNT) {
      const int r = J + wb + e / XF, f = e % XF;
      double a = 0.0;
#pragma unroll
      for (int c = 0; c < 16; c++)
        if (c < wb) a += (double)v.L[(size_t)(o + J + c) * v.nn + o + r] * zb[c * XF + f];
      ys[r * XF + f] -= a;
    }
  }
  __syncthreads();
  for (int e = tid; e < XT * XF; e += XNT) {
    const int r = e / XF, f = e % XF;
    if (r < w && f < nf) v.y[(size_t)f * v.ld + o + r] = ys[e];
  }
}

// x_o = L_oo^-T z_o for the diagonal tile at o (w rows)
template <typename TL>
__device__ __forceinline__ void trsv_diag_bwd(const TrsvView<TL>& v, int o, int w, int nf, double* ys, double* lb, double* zb) {
  const int tid = threadIdx.x;
  for (int e = tid; e < XT * XF; e += XNT) {
    const int r = e / XF, f = e % XF;
    ys[e] = (r < w && f < nf) ? v.y[(size_t)f * v.ld + o + r] : 0.0;
  }
  for (int J = ((w - 1) / 16) * 16; J >= 0; J -= 16) {
    const int wb = min(16, w - J);
    __syncthreads();
    trsv_inverse<TL>(v, o + J, wb, lb);
    // z_c - sum over the tile's rows below the block of L[r][J + c] x_r
    if (tid < 16 * XF) {
      const int c = tid / XF, f = tid % XF;
      double a = 0.0;
      if (c < wb)
        for (int r = J + wb; r < w; r++) a += (double)v.L[(size_t)(o + J + c) * v.nn + o + r] * ys[r * XF + f];
      zb[c * XF + f] = c < wb ? ys[(J + c) * XF + f] - a : 0.0;
    }
    __syncthreads();
    if (tid < 16 * XF) {
      const int r = tid / XF, f = tid % XF;
      double x = 0.0;
#pragma unroll
      for (int c2 = 0; c2 < 16; c2++) x += lb[c2 * 17 + r] * zb[c2 * XF + f];   // (L_JJ^-T)[r][c2]
      if (r < wb) ys[(J + r) * XF + f] = x;
    }
  }
  __syncthreads();
  for (int e = tid; e < XT * XF; e += XNT) {
    const int r = e / XF, f = e % XF;
    if (r < w && f < nf) v.y[(size_t)f * v.ld + o + r] = ys[e];
  }
}

// Forward step over column tile k (diag_only: the first launch, tasks (p, 0)).
template <typename TL>
__global__ __launch_bounds__(XNT) void k_trsv_fwd(const int2* __restrict__ tasks, int k, int diag_only,
                                                  const int32_t* __restrict__ pn, AugView a,
                                                  const int64_t* __restrict__ mem_off,
                                                  const int64_t* __restrict__ fac_off, const TL* __restrict__ fac,
                                                  double* __restrict__ ws, int nf) {
  __shared__ double zs[XT * XF], red[4 * XT * XF];
  __shared__ double ys[XT * XF], lb[16 * 17], zb[16 * XF];
  const int2 tk = tasks[blockIdx.x];
  const int p = tk.x, i = tk.y;
  const TrsvView<TL> v = trsv_view<TL>(p, pn, a, mem_off, fac_off, fac, ws, nf);
  const int tid = threadIdx.x;
  const int oi = XT * i, wi = min(XT, v.nn - oi);
  if (wi <= 0) return;
  if (diag_only) {
    trsv_diag_fwd<TL>(v, oi, wi, nf, ys, lb, zb);
    return;
  }
  const int ok = XT * k;   // a full tile: k < i
  for (int e = tid; e < XT * XF; e += XNT) {
    const int c = e / XF, f = e % XF;
    zs[e] = f < nf ? v.y[(size_t)f * v.ld + ok + c] : 0.0;
  }
  __syncthreads();
  {
    const int r = tid & (XT - 1), part = tid / XT;
    double acc[XF];
#pragma unroll
    for (int f = 0; f < XF; f++) acc[f] = 0.0;
    if (r < wi) {
      TL l[16];
#pragma unroll
      for (int c = 0; c < 16; c++) l[c] = v.L[(size_t)(ok + part * 16 + c) * v.nn + oi + r];
#pragma unroll
      for (int c = 0; c < 16; c++)
#pragma unroll
        for (int f = 0; f < XF; f++) acc[f] += (double)l[c] * zs[(part * 16 + c) * XF + f];
    }
#pragma unroll
    for (int f = 0; f < XF; f++) red[(part * XT + r) * XF + f] = acc[f];
  }
  __syncthreads();
  for (int e = tid; e < XT * XF; e += XNT) {
    const int r = e / XF, f = e % XF;
    if (r < wi && f < nf) {
      const double s = red[(0 * XT + r) * XF + f] + red[(1 * XT + r) * XF + f] + red[(2 * XT + r) * XF + f] +
                       red[(3 * XT + r) * XF + f];
      v.y[(size_t)f * v.ld + oi + r] -= s;
    }
  }
  if (i == k + 1) {
    __syncthreads();   // the tile's updated rows, written above, are read back below
    trsv_diag_fwd<TL>(v, oi, wi, nf, ys, lb, zb);
  }
}

// Backward step s (diag_only: the first launch, tasks (p, T - 1)).
template <typename TL>
__global__ __launch_bounds__(XNT) void k_trsv_bwd(const int2* __restrict__ tasks, int s, int diag_only,
                                                  const int32_t* __restrict__ pn, AugView a,
                                                  const int64_t* __restrict__ mem_off,
                                                  const int64_t* __restrict__ fac_off, const TL* __restrict__ fac,
                                                  double* __restrict__ ws, int nf) {
  __shared__ double xsh[XT * XF], red[4 * XT * XF];
  __shared__ double ys[XT * XF], lb[16 * 17], zb[16 * XF];
  const int2 tk = tasks[blockIdx.x];
  const int p = tk.x, j = tk.y;
  const TrsvView<TL> v = trsv_view<TL>(p, pn, a, mem_off, fac_off, fac, ws, nf);
  const int tid = threadIdx.x;
  const int oj = XT * j, wj = min(XT, v.nn - oj);
  if (wj <= 0) return;
  if (diag_only) {
    trsv_diag_bwd<TL>(v, oj, wj, nf, ys, lb, zb);
    return;
  }
  const int T = (v.nn + XT - 1) / XT;
  const int kp = T - 1 - s;   // j < kp: tile j is full
  const int ok = XT * kp, wk = min(XT, v.nn - ok);
  for (int e = tid; e < XT * XF; e += XNT) {
    const int r = e / XF, f = e % XF;
    xsh[e] = (r < wk && f < nf) ? v.y[(size_t)f * v.ld + ok + r] : 0.0;
  }
  __syncthreads();
  {
    const int c = tid & (XT - 1), part = tid / XT;
    double acc[XF];
#pragma unroll
    for (int f = 0; f < XF; f++) acc[f] = 0.0;
    const TL* Lc = v.L + (size_t)(oj + c) * v.nn + ok + part * 16;
#pragma unroll 4
    for (int r = 0; r < 16; r++) {
      if (part * 16 + r < wk) {
        const double l = (double)Lc[r];
#pragma unroll
        for (int f = 0; f < XF; f++) acc[f] += l * xsh[(part * 16 + r) * XF + f];
      }
    }
#pragma unroll
    for (int f = 0; f < XF; f++) red[(part * XT + c) * XF + f] = acc[f];
  }
  __syncthreads();
  for (int e = tid; e < XT * XF; e += XNT) {
    const int c = e / XF, f = e % XF;
    if (f < nf) {
      const double sum = red[(0 * XT + c) * XF + f] + red[(1 * XT + c) * XF + f] + red[(2 * XT + c) * XF + f] +
                         red[(3 * XT + c) * XF + f];
      v.y[(size_t)f * v.ld + oj + c] -= sum;
    }
  }
  if (j == kp - 1) {
    __syncthreads();
    trsv_diag_bwd<TL>(v, oj, wj, nf, ys, lb, zb);
  }
}

int trsv_tile() { return XT; }

// Task lists of the two sweeps for the patches l (factor dimensions nn[i]): offsets
// off[s] .. off[s + 1] of launch s, forward first (1 + (Tmax - 1) launches), then
// backward (the same count).
void trsv_tasks(const std::vector<int32_t>& l, const std::vector<int>& nn, std::vector<int2>& tasks,
                std::vector<int>& off) {
  tasks.clear();
  off.clear();
  int tmax = 0;
  std::vector<int> T(l.size());
  for (size_t q = 0; q < l.size(); q++) {
    T[q] = (nn[q] + XT - 1) / XT;
    tmax = std::max(tmax, T[q]);
  }
  off.push_back(0);
  for (size_t q = 0; q < l.size(); q++)
    if (T[q] > 0) tasks.push_back(make_int2(l[q], 0));
  off.push_back((int)tasks.size());
  for (int k = 0; k + 1 < tmax; k++) {
    for (size_t q = 0; q < l.size(); q++)
      for (int i = k + 1; i < T[q]; i++) tasks.push_back(make_int2(l[q], i));
    off.push_back((int)tasks.size());
  }
  for (size_t q = 0; q < l.size(); q++)
    if (T[q] > 0) tasks.push_back(make_int2(l[q], T[q] - 1));
  off.push_back((int)tasks.size());
  for (int s = 0; s + 1 < tmax; s++) {
    for (size_t q = 0; q < l.size(); q++)
      for (int j = 0; j < T[q] - 1 - s; j++) tasks.push_back(make_int2(l[q], j));
    off.push_back((int)tasks.size());
  }
}

template <typename TL>
static void trsv_run(const int2* dtasks, const std::vector<int>& off, const int32_t* pn, const AugView& a,
                     const int64_t* mem_off, const int64_t* fac_off, const TL* fac, double* ws, int nf, hipStream_t s) {
  const int nl = (int)off.size() - 1;   // launches: forward nl / 2, backward nl / 2
  const int nfw = nl / 2;
  for (int t = 0; t < nl; t++) {
    const int cnt = off[t + 1] - off[t];
    if (cnt <= 0) continue;
    const bool fw = t < nfw;
    const int step = fw ? t - 1 : t - nfw - 1;   // -1: the diagonal-only launch
    if (fw)
      hipLaunchKernelGGL((k_trsv_fwd<TL>), dim3(cnt), dim3(XNT), 0, s, dtasks + off[t], step < 0 ? 0 : step,
                         step < 0 ? 1 : 0, pn, a, mem_off, fac_off, fac, ws, nf);
    else
      hipLaunchKernelGGL((k_trsv_bwd<TL>), dim3(cnt), dim3(XNT), 0, s, dtasks + off[t], step < 0 ? 0 : step,
                         step < 0 ? 1 : 0, pn, a, mem_off, fac_off, fac, ws, nf);
  }
}

void launch_trsv_tiles(bool fp32, const int2* dtasks, const std::vector<int>& off, const int32_t* pn,
                       const AugView& a, const int64_t* mem_off, const int64_t* fac_off, const void* fac, double* ws,
                       int nf, hipStream_t s) {
  if (off.size() < 2 || nf <= 0) return;
  if (fp32) trsv_run<float>(dtasks, off, pn, a, mem_off, fac_off, (const float*)fac, ws, nf, s);
  else trsv_run<double>(dtasks, off, pn, a, mem_off, fac_off, (const double*)fac, ws, nf, s);
}

}  // namespace lrb
