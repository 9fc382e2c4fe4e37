// gfx950 kernels of the PU-RBF interpolation + Lagrangian advection path.
//
// Layout: particles are SoA ([d][Np], leading dimension ld), data nodes keep the
// caller's AoS [N][dim] order, patch state lives in pools indexed by per-patch
// offsets.  Wave size is 64 (CDNA); blocks are multiples of 64.
#include "kernels.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace lrb {

static constexpr int WAVE = 64;

// --------------------------------------------------------------------------
// block reductions
// --------------------------------------------------------------------------
struct ArgMax { double v; int i; };

__device__ __forceinline__ ArgMax better(ArgMax a, ArgMax b) {
  // larger |value| wins; ties keep the smaller index (LAPACK idamax: first max)
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

template <int NT>
__device__ ArgMax block_argmax(ArgMax x, ArgMax* red) {
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    ArgMax y; y.v = __shfl_xor(x.v, o); y.i = __shfl_xor(x.i, o);
    x = better(x, y);
  }
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  if (lane == 0) red[wid] = x;
  __syncthreads();
  ArgMax r = red[0];
  for (int w = 1; w < NT / WAVE; w++) r = better(r, red[w]);
  __syncthreads();
  return r;
}

template <int NT>
__device__ double block_sum(double x, double* red) {
  for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  if (lane == 0) red[wid] = x;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < NT / WAVE; w++) r += red[w];
  __syncthreads();
  return r;
}

// --------------------------------------------------------------------------
// binning: particle/node -> cell, per-cell node lists, patch membership
// --------------------------------------------------------------------------
__global__ void k_locate_aos(Grid g, const double* __restrict__ x, int n, int32_t* __restrict__ cell) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = x + (size_t)i * g.dim;
  cell[i] = locate(g, p[0], p[1], g.dim == 3 ? p[2] : 0.0);
}

__global__ void k_locate_particles(Grid g, const double* __restrict__ px, int64_t ld, int Np,
                                   int32_t* __restrict__ cell, int32_t* __restrict__ alive,
                                   unsigned long long* removed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || !alive[i]) return;
  int c = locate(g, px[i], px[ld + i], g.dim == 3 ? px[2 * ld + i] : 0.0);
  cell[i] = c;
  if (c < 0) { alive[i] = 0; atomicAdd(removed, 1ull); }
}

__global__ void k_hist(const int32_t* __restrict__ cell, int n, int32_t* __restrict__ counts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && cell[i] >= 0) atomicAdd(&counts[cell[i]], 1);
}

// Exclusive scan of n ints in place: per-block scan + block totals.
static constexpr int SCAN_NT = 256, SCAN_PER = 4, SCAN_CHUNK = SCAN_NT * SCAN_PER;

__global__ __launch_bounds__(SCAN_NT) void k_scan_blocks(int32_t* a, int n, int32_t* sums) {
  __shared__ int32_t s[SCAN_NT];
  const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
  int32_t v[SCAN_PER], run = 0;
  for (int k = 0; k < SCAN_PER; k++) { v[k] = (base + k < n) ? a[base + k] : 0; run += v[k]; }
  s[threadIdx.x] = run;
  __syncthreads();
  for (int o = 1; o < SCAN_NT; o <<= 1) {
    int32_t t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int32_t ex = s[threadIdx.x] - run;
  for (int k = 0; k < SCAN_PER; k++) { if (base + k < n) a[base + k] = ex; ex += v[k]; }
  if (threadIdx.x == SCAN_NT - 1) sums[blockIdx.x] = s[SCAN_NT - 1];
}

__global__ void k_scan_add(int32_t* a, int n, const int32_t* sums) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += sums[i / SCAN_CHUNK];
}

static void exclusive_scan(int32_t* a, int n, int32_t* tmp, hipStream_t s) {
  int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_NT), 0, s, a, n, tmp);
  if (nb > 1) {
    exclusive_scan(tmp, nb, tmp + nb, s);
    hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, s, a, n, tmp);
  }
}

__global__ void k_scatter(const int32_t* __restrict__ cell, int n, int32_t* __restrict__ cursor,
                          int32_t* __restrict__ nodes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell[i];
  if (c >= 0) nodes[atomicAdd(&cursor[c], 1)] = i;
}

// Restore ascending node index inside each cell (the atomic scatter is unordered).
// Cells up to SEG_SMALL nodes: one wave per cell, every element's rank = number of
// smaller elements.  Larger cells (clusters put thousands of nodes into one cell,
// where the rank count is quadratic): listed, then one block per listed cell sorts
// the cell in LDS (bitonic, padded to a power of two) up to SEG_LDS nodes, and
// ranks in global memory with the whole block beyond.  Node ids are distinct, so
// every variant gives the same order.
static constexpr int SEG_SMALL = 256;
static constexpr int SEG_LDS = 16384;
static constexpr int SEG_NT = 1024;

__global__ __launch_bounds__(256) void k_segsort(const int32_t* __restrict__ cell_start, int C,
                                                 const int32_t* __restrict__ unsorted, int32_t* __restrict__ sorted) {
  const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
  if (wid >= C) return;
  const int b = cell_start[wid], e = cell_start[wid + 1], m = e - b;
  if (m > SEG_SMALL) return;
  for (int t = lane; t < m; t += WAVE) {
    int v = unsorted[b + t], r = 0;
    for (int u = 0; u < m; u++) r += unsorted[b + u] < v;
    sorted[b + r] = v;
  }
}

// big[0]: number of cells above SEG_SMALL nodes (zeroed by the caller), big[1..]: the cells
__global__ void k_big_cells(const int32_t* __restrict__ cell_start, int C, int32_t* __restrict__ big) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  if (cell_start[c + 1] - cell_start[c] > SEG_SMALL) big[1 + atomicAdd(&big[0], 1)] = c;
}

__global__ __launch_bounds__(SEG_NT) void k_segsort_big(const int32_t* __restrict__ cell_start,
                                                        const int32_t* __restrict__ big,
                                                        const int32_t* __restrict__ unsorted,
                                                        int32_t* __restrict__ sorted) {
  __shared__ int32_t key[SEG_LDS];
  const int tid = threadIdx.x;
  const int nbig = big[0];
  for (int bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
    const int cell = big[1 + bi];
    const int beg = cell_start[cell], m = cell_start[cell + 1] - beg;
    if (m <= SEG_LDS) {
      int np2 = 1;
      while (np2 < m) np2 <<= 1;
      for (int i = tid; i < np2; i += SEG_NT) key[i] = i < m ? unsorted[beg + i] : INT32_MAX;
      __syncthreads();
      for (int k = 2; k <= np2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int i = tid; i < np2; i += SEG_NT) {
            const int ij = i ^ j;
            if (ij > i) {
              const int a = key[i], b2 = key[ij];
              if ((a > b2) == ((i & k) == 0)) { key[i] = b2; key[ij] = a; }
            }
          }
          __syncthreads();
        }
      for (int i = tid; i < m; i += SEG_NT) sorted[beg + i] = key[i];
    } else {
      for (int t = tid; t < m; t += SEG_NT) {
        const int v = unsorted[beg + t];
        int r = 0;
        for (int u = 0; u < m; u++) r += unsorted[beg + u] < v;
        sorted[beg + r] = v;
      }
    }
    __syncthreads();
  }
}

void launch_cell_lists(const int32_t* cell, int n, int C, int32_t* cell_start, int32_t* cell_nodes,
                       int32_t* scratch, hipStream_t s) {
  // scratch layout: [cursor C+1][unsorted n][scan tmp]; after the scatter the cursor
  // region holds the list of large cells
  int32_t* cursor = scratch;
  int32_t* unsorted = scratch + (C + 1);
  int32_t* tmp = unsorted + (n > 0 ? n : 1);
  hipMemsetAsync(cell_start, 0, sizeof(int32_t) * (C + 1), s);
  if (n > 0) hipLaunchKernelGGL(k_hist, dim3((n + 255) / 256), dim3(256), 0, s, cell, n, cell_start);
  exclusive_scan(cell_start, C + 1, tmp, s);
  hipMemcpyAsync(cursor, cell_start, sizeof(int32_t) * (C + 1), hipMemcpyDeviceToDevice, s);
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, s, cell, n, cursor, unsorted);
    hipLaunchKernelGGL(k_segsort, dim3((C + 3) / 4), dim3(256), 0, s, cell_start, C, unsorted, cell_nodes);
    int32_t* big = cursor;
    hipMemsetAsync(big, 0, sizeof(int32_t), s);
    hipLaunchKernelGGL(k_big_cells, dim3((C + 255) / 256), dim3(256), 0, s, cell_start, C, big);
    hipLaunchKernelGGL(k_segsort_big, dim3(1024), dim3(SEG_NT), 0, s, cell_start, big, unsorted, cell_nodes);
  }
}

// Visit the (up to 2^d) cells touching PU node p in ascending cell index
// (partition.c:54-87 builds the inverse map by ascending element index).
template <typename F>
__device__ __forceinline__ void for_patch_cells(const Grid& g, int p, F&& f) {
  int a = p % g.n[0], b = (p / g.n[0]) % g.n[1], c = g.dim == 3 ? p / (g.n[0] * g.n[1]) : 0;
  int k0 = g.dim == 3 ? c - 1 : 0, k1 = g.dim == 3 ? c : 0;
  for (int k = k0; k <= k1; k++) {
    if (g.dim == 3 && (k < 0 || k >= g.nc[2])) continue;
    for (int j = b - 1; j <= b; j++) {
      if (j < 0 || j >= g.nc[1]) continue;
      for (int i = a - 1; i <= a; i++) {
        if (i < 0 || i >= g.nc[0]) continue;
        f(i + g.nc[0] * (j + g.nc[1] * k));
      }
    }
  }
}

__global__ void k_patch_count(Grid g, const int32_t* __restrict__ cs, int32_t* __restrict__ pcount) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.P) return;
  int cnt = 0;
  for_patch_cells(g, p, [&](int c) { cnt += cs[c + 1] - cs[c]; });
  pcount[p] = cnt;
}

__global__ void k_patch_gather(Grid g, const int32_t* __restrict__ list, int nlist, const int32_t* __restrict__ cs,
                               const int32_t* __restrict__ cn, const int64_t* __restrict__ mem_off,
                               int32_t* __restrict__ mem) {
  const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
  if (wid >= nlist) return;
  const int p = list[wid];
  int32_t* out = mem + mem_off[p];
  int o = 0;
  for_patch_cells(g, p, [&](int c) {
    const int b = cs[c], m = cs[c + 1] - b;
    for (int t = lane; t < m; t += WAVE) out[o + t] = cn[b + t];
    o += m;
  });
}

__global__ void k_touch(Grid g, const int32_t* __restrict__ cell, const int32_t* __restrict__ alive, int Np,
                        uint8_t* __restrict__ touched) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || !alive[i]) return;
  const int c = cell[i];
  for (int k = 0; k < g.nen; k++) touched[corner_node(g, c, k)] = 1;
}

void launch_locate_aos(const Grid& g, const double* x, int n, int32_t* cell, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_locate_aos, dim3((n + 255) / 256), dim3(256), 0, s, g, x, n, cell);
}
void launch_locate_particles(const Grid& g, double* px, int64_t ld, int Np, int32_t* cell, int32_t* alive,
                             unsigned long long* removed, hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_locate_particles, dim3((Np + 255) / 256), dim3(256), 0, s, g, px, ld, Np, cell, alive, removed);
}
void launch_patch_count(const Grid& g, const int32_t* cs, int32_t* pcount, hipStream_t s) {
  hipLaunchKernelGGL(k_patch_count, dim3((g.P + 255) / 256), dim3(256), 0, s, g, cs, pcount);
}
void launch_patch_gather(const Grid& g, const int32_t* list, int nlist, const int32_t* cs, const int32_t* cn,
                         const int64_t* mem_off, int32_t* mem, hipStream_t s) {
  if (nlist > 0) hipLaunchKernelGGL(k_patch_gather, dim3((nlist + 3) / 4), dim3(256), 0, s, g, list, nlist, cs, cn, mem_off, mem);
}
void launch_touch(const Grid& g, const int32_t* cell, const int32_t* alive, int Np, uint8_t* touched, hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_touch, dim3((Np + 255) / 256), dim3(256), 0, s, g, cell, alive, Np, touched);
}

// --------------------------------------------------------------------------
// Patch assembly + Bunch-Kaufman LDL^T (dsytf2 'L'), one workgroup per patch,
// matrix column-major in the patch's factor slot.  Restates the factorisation
// PETSc's SeqDense Cholesky performs through dsytrf (partition_factor.c:37-55).
// --------------------------------------------------------------------------
static constexpr int FNT = 256;

template <int KT>
__global__ __launch_bounds__(FNT) void k_factor_bk(Grid g, const int32_t* __restrict__ list, const int32_t* __restrict__ pn,
                                                   const int64_t* __restrict__ mem_off, const int64_t* __restrict__ fac_off,
                                                   const int32_t* __restrict__ mem_pool, const double* __restrict__ xd,
                                                   double* __restrict__ fac, int32_t* __restrict__ piv_pool,
                                                   int32_t* __restrict__ info_out, AugView aug) {
  __shared__ ArgMax red[FNT / WAVE];
  __shared__ double redd[FNT / WAVE];
  __shared__ int s_kp, s_kstep;
  const int p = list[blockIdx.x];
  const int n0 = pn[p];
  if (n0 == 0) return;
  // with a polynomial block of m terms the factored matrix is the (n0 + m) square
  // saddle-point matrix [A P; P^T 0] (extension)
  const int m = aug.m ? aug.m[p] : 0;
  const int n = n0 + m;
  double* A = fac + fac_off[p];
  const int32_t* mem = mem_pool + mem_off[p];
  int32_t* piv = piv_pool + bk_piv_base(aug, p, mem_off[p]);
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
  const int dim = g.dim;
  const double eps = aug.eps ? aug.eps[p] : g.eps_r;
  // assemble the lower triangle: A_ij = phi(|x_i - x_j|, eps_r)  (partition_solve.c:47-82)
  for (int j = wid; j < n0; j += FNT / WAVE) {
    const double* xj = xd + (size_t)mem[j] * dim;
    const double xj0 = xj[0], xj1 = xj[1], xj2 = dim == 3 ? xj[2] : 0.0;
    for (int i = j + lane; i < n0; i += WAVE) {
      const double* xi = xd + (size_t)mem[i] * dim;
      double d0 = xi[0] - xj0, d1 = xi[1] - xj1, d2 = dim == 3 ? xi[2] - xj2 : 0.0;
      double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      A[(size_t)j * n + i] = rbf<KT, double>(r, eps);
    }
  }
  if (m > 0) {
    // rows n0.. of the lower triangle: P^T (monomials of the member shifted to the
    // patch node, scaled by ps), then the zero block
    const int deg = aug.deg[p];
    const int pi = p % g.n[0], pj = (p / g.n[0]) % g.n[1], pk = dim == 3 ? p / (g.n[0] * g.n[1]) : 0;
    const double c[3] = {g.O[0] + pi * g.h[0], g.O[1] + pj * g.h[1], dim == 3 ? g.O[2] + pk * g.h[2] : 0.0};
    for (int i = tid; i < n0; i += FNT) {
      const double* x = xd + (size_t)mem[i] * dim;
      double xi[3], pv[POLY_MAX];
      for (int d = 0; d < dim; d++) xi[d] = (x[d] - c[d]) * aug.ps;
      poly_eval(dim, deg, xi, pv);
      for (int q = 0; q < m; q++) A[(size_t)i * n + n0 + q] = pv[q];
    }
    for (int e = tid; e < m * m; e += FNT) {
      const int r = e / m, q = e % m;
      if (q >= r) A[(size_t)(n0 + r) * n + n0 + q] = 0.0;
    }
  }
  __syncthreads();
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  int info = 0;
  int k = 0;
  while (k < n) {
    // colmax over A(k+1:n, k)
    ArgMax am; am.v = -1.0; am.i = n;
    for (int i = k + 1 + tid; i < n; i += FNT) { ArgMax c; c.v = fabs(A[(size_t)k * n + i]); c.i = i; am = better(am, c); }
    am = block_argmax<FNT>(am, red);
    const double absakk = fabs(A[(size_t)k * n + k]);
    const double colmax = (k < n - 1) ? am.v : 0.0;
    const int imax = (k < n - 1) ? am.i : k;
    int kp = k, kstep = 1;
    bool zero = (fmax(absakk, colmax) == 0.0) || isnan(absakk);
    if (!zero && !(absakk >= alpha * colmax)) {
      // rowmax: off-diagonal entries of row/column imax in the trailing matrix
      double rm = 0.0;
      for (int j = k + tid; j < imax; j += FNT) rm = fmax(rm, fabs(A[(size_t)j * n + imax]));
      for (int j = imax + 1 + tid; j < n; j += FNT) rm = fmax(rm, fabs(A[(size_t)imax * n + j]));
      for (int o = WAVE / 2; o > 0; o >>= 1) rm = fmax(rm, __shfl_xor(rm, o));
      if (lane == 0) redd[wid] = rm;
      __syncthreads();
      double rowmax = redd[0];
      for (int w = 1; w < FNT / WAVE; w++) rowmax = fmax(rowmax, redd[w]);
      __syncthreads();
      if (absakk >= alpha * colmax * (colmax / rowmax)) kp = k;
      else if (fabs(A[(size_t)imax * n + imax]) >= alpha * rowmax) kp = imax;
      else { kp = imax; kstep = 2; }
    }
    if (zero) { if (info == 0) info = k + 1; }
    if (!zero) {
      const int kk = k + kstep - 1;
      if (kp != kk) {
        for (int i = kp + 1 + tid; i < n; i += FNT) {
          double t = A[(size_t)kk * n + i]; A[(size_t)kk * n + i] = A[(size_t)kp * n + i]; A[(size_t)kp * n + i] = t;
        }
        for (int j = kk + 1 + tid; j < kp; j += FNT) {
          double t = A[(size_t)kk * n + j]; A[(size_t)kk * n + j] = A[(size_t)j * n + kp]; A[(size_t)j * n + kp] = t;
        }
        if (tid == 0) {
          double t = A[(size_t)kk * n + kk]; A[(size_t)kk * n + kk] = A[(size_t)kp * n + kp]; A[(size_t)kp * n + kp] = t;
          if (kstep == 2) { t = A[(size_t)k * n + k + 1]; A[(size_t)k * n + k + 1] = A[(size_t)k * n + kp]; A[(size_t)k * n + kp] = t; }
        }
        __syncthreads();
      }
      if (kstep == 1) {
        if (k < n - 1) {
          const double d11 = 1.0 / A[(size_t)k * n + k];
          const double* ck = A + (size_t)k * n;
          for (int j = k + 1 + wid; j < n; j += FNT / WAVE) {
            const double t = -d11 * ck[j];
            double* cj = A + (size_t)j * n;
            for (int i = j + lane; i < n; i += WAVE) cj[i] += ck[i] * t;
          }
          __syncthreads();
          for (int i = k + 1 + tid; i < n; i += FNT) A[(size_t)k * n + i] *= d11;
        }
      } else if (k < n - 2) {
        const double* c0 = A + (size_t)k * n;
        const double* c1 = A + (size_t)(k + 1) * n;
        double d21 = c0[k + 1];
        const double d11 = c1[k + 1] / d21;
        const double d22 = c0[k] / d21;
        const double t = 1.0 / (d11 * d22 - 1.0);
        d21 = t / d21;
        for (int j = k + 2 + wid; j < n; j += FNT / WAVE) {
          const double wk = d21 * (d11 * c0[j] - c1[j]);
          const double wkp1 = d21 * (d22 * c1[j] - c0[j]);
          double* cj = A + (size_t)j * n;
          for (int i = j + lane; i < n; i += WAVE) cj[i] = cj[i] - c0[i] * wk - c1[i] * wkp1;
        }
        __syncthreads();
        for (int j = k + 2 + tid; j < n; j += FNT) {
          const double a0 = c0[j], a1 = c1[j];
          A[(size_t)k * n + j] = d21 * (d11 * a0 - a1);
          A[(size_t)(k + 1) * n + j] = d21 * (d22 * a1 - a0);
        }
      }
    }
    if (tid == 0) {
      if (kstep == 1) piv[k] = kp + 1;
      else { piv[k] = -(kp + 1); piv[k + 1] = -(kp + 1); }
    }
    __syncthreads();
    k += kstep;
  }
  if (tid == 0 && info_out) info_out[p] = info;
}

void launch_factor_bk(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                      const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                      double* fac, int32_t* piv, int32_t* info, const AugView& a, hipStream_t s) {
  if (nlist <= 0) return;
#define LRB_FBK(K) hipLaunchKernelGGL(k_factor_bk<K>, dim3(nlist), dim3(FNT), 0, s, g, list, pn, mem_off, fac_off, mem_pool, xd, fac, piv, info, a)
  switch (ktype) {
    case K_LINEAR: LRB_FBK(K_LINEAR); break;
    case K_WC0: LRB_FBK(K_WC0); break;
    case K_WC2: LRB_FBK(K_WC2); break;
    case K_BUMP: LRB_FBK(K_BUMP); break;
    case K_MQ: LRB_FBK(K_MQ); break;
    case K_IMQ: LRB_FBK(K_IMQ); break;
    case K_PHS3: LRB_FBK(K_PHS3); break;
    default: LRB_FBK(K_GAUSS); break;
  }
#undef LRB_FBK
}

// dsytrs 'L' (petsc_dense.c:66) for nf right-hand sides gathered from the data
// values (PartitionMatSetRHS, partition_factor.c:57-90).  X row-major [n][nf].
__global__ __launch_bounds__(FNT) void k_solve_bk(const int32_t* __restrict__ list, const int32_t* __restrict__ pn,
                                                  const int64_t* __restrict__ mem_off, const int64_t* __restrict__ fac_off,
                                                  const int32_t* __restrict__ mem_pool, const double* __restrict__ fac,
                                                  const int32_t* __restrict__ piv_pool, const double* __restrict__ values,
                                                  int nf, double* __restrict__ coef, AugView aug) {
  __shared__ double redd[FNT / WAVE];
  const int p = list[blockIdx.x];
  const int n0 = pn[p];
  if (n0 == 0 || nf == 0) return;
  // saddle-point system (extension): right-hand side [f; 0], rows n0.. (beta) in aug.beta
  const int m = aug.m ? aug.m[p] : 0;
  const int n = n0 + m;
  const double* A = fac + fac_off[p];
  const int32_t* mem = mem_pool + mem_off[p];
  const int32_t* piv = piv_pool + bk_piv_base(aug, p, mem_off[p]);
  double* B = coef + mem_off[p] * nf;
  double* Bb = m > 0 ? aug.beta + (size_t)p * aug.bstride : nullptr;
#define BR(i) ((i) < n0 ? B + (size_t)(i) * nf : Bb + (size_t)((i) - n0) * nf)
  const int tid = threadIdx.x;
  for (int t = tid; t < n * nf; t += FNT) { int i = t / nf, f = t % nf; BR(i)[f] = i < n0 ? values[(size_t)mem[i] * nf + f] : 0.0; }
  __syncthreads();
  // forward: L D y = b
  int k = 0;
  while (k < n) {
    if (piv[k] > 0) {
      const int kp = piv[k] - 1;
      if (kp != k && tid < nf) { double t = BR(k)[tid]; BR(k)[tid] = BR(kp)[tid]; BR(kp)[tid] = t; }
      __syncthreads();
      const double* ck = A + (size_t)k * n;
      for (int f = 0; f < nf; f++) {
        const double bk = BR(k)[f];
        for (int i = k + 1 + tid; i < n; i += FNT) BR(i)[f] -= ck[i] * bk;
      }
      __syncthreads();
      if (tid < nf) BR(k)[tid] *= 1.0 / ck[k];
      __syncthreads();
      k += 1;
    } else {
      const int kp = -piv[k] - 1;
      if (kp != k + 1 && tid < nf) { double t = BR(k + 1)[tid]; BR(k + 1)[tid] = BR(kp)[tid]; BR(kp)[tid] = t; }
      __syncthreads();
      const double* c0 = A + (size_t)k * n;
      const double* c1 = A + (size_t)(k + 1) * n;
      for (int f = 0; f < nf; f++) {
        const double b0 = BR(k)[f], b1 = BR(k + 1)[f];
        for (int i = k + 2 + tid; i < n; i += FNT) { double v = BR(i)[f] - c0[i] * b0; BR(i)[f] = v - c1[i] * b1; }
      }
      __syncthreads();
      if (tid < nf) {
        const double akm1k = c0[k + 1], akm1 = c0[k] / akm1k, ak = c1[k + 1] / akm1k, denom = akm1 * ak - 1.0;
        const double bkm1 = BR(k)[tid] / akm1k, bk = BR(k + 1)[tid] / akm1k;
        BR(k)[tid] = (ak * bkm1 - bk) / denom;
        BR(k + 1)[tid] = (akm1 * bk - bkm1) / denom;
      }
      __syncthreads();
      k += 2;
    }
  }
  // backward: L^T x = y
  k = n - 1;
  while (k >= 0) {
    if (piv[k] > 0) {
      const double* ck = A + (size_t)k * n;
      for (int f = 0; f < nf; f++) {
        double s = 0.0;
        for (int i = k + 1 + tid; i < n; i += FNT) s += BR(i)[f] * ck[i];
        s = block_sum<FNT>(s, redd);
        if (tid == 0) BR(k)[f] -= s;
      }
      __syncthreads();
      const int kp = piv[k] - 1;
      if (kp != k && tid < nf) { double t = BR(k)[tid]; BR(k)[tid] = BR(kp)[tid]; BR(kp)[tid] = t; }
      __syncthreads();
      k -= 1;
    } else {
      const double* ck = A + (size_t)k * n;
      const double* ckm1 = A + (size_t)(k - 1) * n;
      for (int f = 0; f < nf; f++) {
        double s = 0.0, s1 = 0.0;
        for (int i = k + 1 + tid; i < n; i += FNT) { s += BR(i)[f] * ck[i]; s1 += BR(i)[f] * ckm1[i]; }
        s = block_sum<FNT>(s, redd);
        s1 = block_sum<FNT>(s1, redd);
        if (tid == 0) { BR(k)[f] -= s; BR(k - 1)[f] -= s1; }
      }
      __syncthreads();
      const int kp = -piv[k] - 1;
      if (kp != k && tid < nf) { double t = BR(k)[tid]; BR(k)[tid] = BR(kp)[tid]; BR(kp)[tid] = t; }
      __syncthreads();
      k -= 2;
    }
  }
#undef BR
}

void launch_solve_bk(const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off, const int64_t* fac_off,
                     const int32_t* mem_pool, const double* fac, const int32_t* piv, const double* values, int nf,
                     double* coef, const AugView& a, hipStream_t s) {
  if (nlist > 0) hipLaunchKernelGGL(k_solve_bk, dim3(nlist), dim3(FNT), 0, s, list, pn, mem_off, fac_off, mem_pool, fac, piv, values, nf, coef, a);
}

// --------------------------------------------------------------------------
// Shepard weights (swarm_interp.c:283-380): Wendland-C2 weight of the cell's
// 2^d corner nodes normalised by the sum over every PU node; nodes outside the
// support contribute exact zeros, so the sum runs over the support window in
// ascending node order.
// --------------------------------------------------------------------------
// PU node coordinate O + i*h, as the host builds the node arrays (DMDASetUniformCoordinates,
// partition_of_unity.c:143): product and sum rounded separately.
__device__ __forceinline__ double pu_coord(const Grid& g, int d, int i) {
#pragma clang fp contract(off)
  return g.O[d] + (double)i * g.h[d];
}

__device__ __forceinline__ void weights_point(const Grid& g, const double* __restrict__ pux, const double* __restrict__ puy,
                                              const double* __restrict__ puz, double x, double y, double z, int cell,
                                              double* w) {
  // Every PU node with a nonzero weight lies within R of x, so inside the window of
  // cells ci - m .. ci + 1 + m per axis, m = ceil(R / h) (R widened by 1e-12); the
  // nodes outside it add exact zeros to the reference's sum over all PU nodes
  // (swarm_interp.c:327-372), which is accumulated here in the same ascending
  // node order.  The window is then narrowed per particle to the nodes within R
  // per axis: with R below the cell size that is 1-2 nodes per axis instead of 4.
  const int ci = cell % g.nc[0], cj = (cell / g.nc[0]) % g.nc[1], ck = g.dim == 3 ? cell / (g.nc[0] * g.nc[1]) : 0;
  const double R = (1.0 / g.eps_w) * (1.0 + 1e-12);
  const int cc[3] = {ci, cj, ck};
  const double xc[3] = {x, y, z};
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < g.dim; d++) {
    const int m = (int)ceil(R / g.h[d]);
    lo[d] = max(cc[d] - m, 0);
    hi[d] = min(cc[d] + 1 + m, g.n[d] - 1);
    // narrowed per particle: node i can lie within R only if |x_d - (O_d + i h_d)| <= R,
    // i.e. t - R/h <= i <= t + R/h with t = (x_d - O_d) / h_d (bounds widened by 1e-6
    // of a cell against rounding); the nodes dropped add exact zeros
    const double t = (xc[d] - g.O[d]) / g.h[d], rh = R / g.h[d];
    if (fabs(t) < 1e9) {   // a NaN coordinate keeps the cell window (and NaN weights)
      lo[d] = max(lo[d], (int)ceil(t - rh - 1e-6));
      hi[d] = min(hi[d], (int)floor(t + rh + 1e-6));
    }
  }
  // w is indexed with compile-time indices only (unrolled selects), so it stays in
  // registers: a runtime index lets the compiler move the array to LDS
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < 8; k++) w[k] = 0.0;
  for (int kk = lo[2]; kk <= hi[2]; kk++) {
    const double dz = g.dim == 3 ? z - pu_coord(g, 2, kk) : 0.0;
    const int ek = kk - ck;
    for (int j = lo[1]; j <= hi[1]; j++) {
      const double dy = y - pu_coord(g, 1, j);
      const int ej = j - cj;
      for (int i = lo[0]; i <= hi[0]; i++) {
        const double dx = x - pu_coord(g, 0, i);
        const double wt = rbf<K_WC2, double>(sqrt_fast(dx * dx + dy * dy + dz * dz), g.eps_w);
        sum += wt;
        const int ei = i - ci;
        if ((unsigned)ei <= 1u && (unsigned)ej <= 1u && (unsigned)ek <= 1u) {
          const int q = (ej ? (ei ? 2 : 3) : (ei ? 1 : 0)) + 4 * ek;
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (k == q) w[k] = wt;
        }
      }
    }
  }
  if (g.wnorm) {
    sum = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k < g.nen) sum += w[k];
  }
  const double inv = sum > 0.0 ? 1.0 / sum : 0.0;
#pragma unroll
  for (int k = 0; k < 8; k++) w[k] *= inv;
}

__global__ void k_weights(Grid g, const double* __restrict__ pux, const double* __restrict__ puy, const double* __restrict__ puz,
                          const double* __restrict__ px, int64_t ld, int Np, const int32_t* __restrict__ cell,
                          const int32_t* __restrict__ alive, double* __restrict__ w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || !alive[i]) return;
  double wl[8];
  weights_point(g, pux, puy, puz, px[i], px[ld + i], g.dim == 3 ? px[2 * ld + i] : 0.0, cell[i], wl);
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (k < g.nen) w[(size_t)k * ld + i] = wl[k];
}

void launch_weights(const Grid& g, const double* pux, const double* puy, const double* puz, const double* px, int64_t ld,
                    int Np, const int32_t* cell, const int32_t* alive, double* w, hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_weights, dim3((Np + 255) / 256), dim3(256), 0, s, g, pux, puy, puz, px, ld, Np, cell, alive, w);
}

// --------------------------------------------------------------------------
// PU evaluation in the reference's summation order (interpolation.c:39-154):
// per corner patch, interpolant_f = sum_j phi(|x - x_j|) X_j,f over the patch's
// members, then values_f += interpolant_f * w_k.  Particles whose 3^d cell
// neighbourhood holds no node keep their previous values.
// --------------------------------------------------------------------------
template <int KT, int NF>
__global__ __launch_bounds__(256) void k_eval_ref(Grid g, const double* __restrict__ px, int64_t ld, int Np,
                                                  const int32_t* __restrict__ cell, const int32_t* __restrict__ alive,
                                                  const double* __restrict__ w, PatchView pv,
                                                  const double* __restrict__ xd, const int32_t* __restrict__ cs,
                                                  int nfields, int f0, double* __restrict__ out, AugView aug) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || !alive[i]) return;
  const int c = cell[i];
  const int ci = c % g.nc[0], cj = (c / g.nc[0]) % g.nc[1], ck = g.dim == 3 ? c / (g.nc[0] * g.nc[1]) : 0;
  int nnb = 0;
  for (int dk = (g.dim == 3 ? -1 : 0); dk <= (g.dim == 3 ? 1 : 0); dk++)
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        const int a = ci + di, b = cj + dj, e = ck + dk;
        if (a < 0 || a >= g.nc[0] || b < 0 || b >= g.nc[1]) continue;
        if (g.dim == 3 && (e < 0 || e >= g.nc[2])) continue;
        const int q = a + g.nc[0] * (b + g.nc[1] * e);
        nnb += cs[q + 1] - cs[q];
      }
  if (nnb == 0) return;
  const double x = px[i], y = px[ld + i], z = g.dim == 3 ? px[2 * ld + i] : 0.0;
  const int dim = g.dim;
  const int nfc = min(NF, nfields - f0);
  double val[NF];
  for (int f = 0; f < NF; f++) val[f] = 0.0;
  for (int k = 0; k < g.nen; k++) {
    const int p = corner_node(g, c, k);
    if (!pv.active[p]) continue;
    const int n = pv.n[p];
    const int64_t off = pv.mem_off[p];
    const int32_t* mem = pv.members + off;
    const double* X = pv.coef + off * pv.nf_total + f0;
    double it[NF];
    for (int f = 0; f < NF; f++) it[f] = 0.0;
    for (int nd = 0; nd < n; nd++) {
      const double* xj = xd + (size_t)mem[nd] * dim;
      const double d0 = x - xj[0], d1 = y - xj[1], d2 = dim == 3 ? z - xj[2] : 0.0;
      const double phi = rbf<KT, double>(sqrt(d0 * d0 + d1 * d1 + d2 * d2), aug.eps ? aug.eps[p] : g.eps_r);
      const double* Xr = X + (size_t)nd * pv.nf_total;
#pragma unroll
      for (int f = 0; f < NF; f++) if (f < nfc) it[f] += phi * Xr[f];
    }
    const int m = aug.m ? aug.m[p] : 0;
    if (m > 0) {
      // polynomial part of the patch (extension)
      const int pi = p % g.n[0], pj = (p / g.n[0]) % g.n[1], pk = dim == 3 ? p / (g.n[0] * g.n[1]) : 0;
      double xi[3], pvv[POLY_MAX];
      xi[0] = (x - (g.O[0] + pi * g.h[0])) * aug.ps;
      xi[1] = (y - (g.O[1] + pj * g.h[1])) * aug.ps;
      xi[2] = dim == 3 ? (z - (g.O[2] + pk * g.h[2])) * aug.ps : 0.0;
      poly_eval(dim, aug.deg[p], xi, pvv);
      const double* bt = aug.beta + (size_t)p * aug.bstride + f0;
      for (int q = 0; q < m; q++)
#pragma unroll
        for (int f = 0; f < NF; f++) if (f < nfc) it[f] += pvv[q] * bt[(size_t)q * pv.nf_total + f];
    }
    const double wk = w[(size_t)k * ld + i];
#pragma unroll
    for (int f = 0; f < NF; f++) val[f] += it[f] * wk;
  }
#pragma unroll
  for (int f = 0; f < NF; f++) if (f < nfc) out[(size_t)(f0 + f) * ld + i] = val[f];
}

template <int KT>
static void eval_ref_k(const Grid& g, const double* px, int64_t ld, int Np, const int32_t* cell, const int32_t* alive,
                       const double* w, const PatchView& pv, const double* xd, const int32_t* cs, int nfields,
                       double* out, const AugView& a, hipStream_t s) {
  const dim3 grid((Np + 255) / 256), blk(256);
  for (int f0 = 0; f0 < nfields; f0 += 4) {
    const int r = nfields - f0;
    if (r == 1) hipLaunchKernelGGL((k_eval_ref<KT, 1>), grid, blk, 0, s, g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, f0, out, a);
    else if (r == 2) hipLaunchKernelGGL((k_eval_ref<KT, 2>), grid, blk, 0, s, g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, f0, out, a);
    else if (r == 3) hipLaunchKernelGGL((k_eval_ref<KT, 3>), grid, blk, 0, s, g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, f0, out, a);
    else hipLaunchKernelGGL((k_eval_ref<KT, 4>), grid, blk, 0, s, g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, f0, out, a);
  }
}

void launch_eval_ref(const Grid& g, int ktype, const double* px, int64_t ld, int Np, const int32_t* cell,
                     const int32_t* alive, const double* w, const PatchView& pv, const double* xd,
                     const int32_t* cs, int nfields, double* out, const AugView& a, hipStream_t s) {
  if (Np <= 0 || nfields <= 0) return;
  switch (ktype) {
    case K_LINEAR: eval_ref_k<K_LINEAR>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
    case K_WC0: eval_ref_k<K_WC0>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
    case K_WC2: eval_ref_k<K_WC2>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
    case K_BUMP: eval_ref_k<K_BUMP>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
    case K_MQ: eval_ref_k<K_MQ>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
    case K_IMQ: eval_ref_k<K_IMQ>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
    case K_PHS3: eval_ref_k<K_PHS3>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
    default: eval_ref_k<K_GAUSS>(g, px, ld, Np, cell, alive, w, pv, xd, cs, nfields, out, a, s); break;
  }
}

// --------------------------------------------------------------------------
// Cell-ordered PU evaluation.  Particles are grouped by cell (perm/pstart); one
// workgroup per cell, one lane per particle.  The 3^d neighbour cells are visited
// in ascending cell index and, inside a cell, nodes in ascending index: phi is
// computed once per (particle, node) and accumulated into every corner patch that
// contains the node's cell.  Each patch's members are thereby visited in exactly
// the member order of interpolation.c:123-129.  Node coordinates and coefficients
// are wave-uniform loads (scalar cache / SGPR operands).
// --------------------------------------------------------------------------
// Particles grouped by cell for the cell-ordered evaluation; dead particles go to
// the tail bucket C.  Lanes of a wave holding the same key in a row (the common
// case once the swarm is stored in cell order) share one atomic.
struct WaveRun {
  int head;   // lane that starts this lane's run
  int len;    // run length (valid at the head lane)
};

__device__ __forceinline__ WaveRun wave_run(int key) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == WAVE - 1 ? ~0ull : ((1ull << (lane + 1)) - 1);
  WaveRun r;
  r.head = WAVE - 1 - __clzll((long long)(heads & upto));
  const unsigned long long above = lane == WAVE - 1 ? 0ull : (heads & ~upto);
  r.len = (above ? __ffsll((long long)above) - 1 : WAVE) - lane;
  return r;
}

__device__ __forceinline__ int particle_key(const int32_t* cell, const int32_t* alive, int n, int C, int i) {
  if (i >= n) return -1;
  return alive[i] ? cell[i] : C;
}

// Live particles: per-cell counts, one atomic per same-cell run of a wave.  Dead
// particles (finished or removed) take no atomics: each block writes its dead
// count, and a scan of those counts gives every dead particle its tail slot (one
// counter shared by all dead particles would serialise once most are dead).
// Blocks of 256 threads.
__global__ __launch_bounds__(256) void k_hist_alive(const int32_t* __restrict__ cell, const int32_t* __restrict__ alive,
                                                    int n, int C, int32_t* __restrict__ counts,
                                                    int32_t* __restrict__ bdead) {
  __shared__ int32_t wd[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int key = particle_key(cell, alive, n, C, i);
  const bool dead = key == C;
  const int lkey = dead ? -1 : key;
  const WaveRun r = wave_run(lkey);
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  if (lkey >= 0 && r.head == lane) atomicAdd(&counts[lkey], r.len);
  const unsigned long long m = __ballot(dead);
  if (lane == 0) wd[wv] = (int32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) bdead[blockIdx.x] = wd[0] + wd[1] + wd[2] + wd[3];
}

// dead_base: the first dead slot (the live total, pstart[C]); doff: exclusive scan
// of the blocks' dead counts
__global__ __launch_bounds__(256) void k_pscatter(const int32_t* __restrict__ cell, const int32_t* __restrict__ alive,
                                                  int n, int C, int32_t* __restrict__ cursor,
                                                  const int32_t* __restrict__ dead_base,
                                                  const int32_t* __restrict__ doff, int32_t* __restrict__ perm) {
  __shared__ int32_t wd[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int key = particle_key(cell, alive, n, C, i);
  const bool dead = key == C;
  const int lkey = dead ? -1 : key;
  const WaveRun r = wave_run(lkey);
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  int base = 0;
  if (lkey >= 0 && r.head == lane) base = atomicAdd(&cursor[lkey], r.len);
  base = __shfl(base, r.head);
  if (lkey >= 0) perm[base + lane - r.head] = i;
  const unsigned long long m = __ballot(dead);
  if (lane == 0) wd[wv] = (int32_t)__popcll(m);
  __syncthreads();
  if (dead) {
    int off = *dead_base + doff[blockIdx.x];
    for (int k = 0; k < wv; k++) off += wd[k];
    const unsigned long long below = lane == 0 ? 0ull : (m & ((1ull << lane) - 1ull));
    perm[off + (int)__popcll(below)] = i;
  }
}

__global__ void k_set_i32(int32_t* __restrict__ a, int32_t v) { *a = v; }

__global__ void k_zero_i32(int32_t* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_copy_i32(const int32_t* __restrict__ src, int32_t* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

size_t sort_particles_scratch(int C, int Np) {
  const size_t nb = ((size_t)std::max(Np, 1) + 255) / 256;
  return ((size_t)C + 2) + ((size_t)C + 2) / 1024 + 64 + nb + nb / 1024 + 64;
}

// Kernels only (no memset or copy nodes), so the sequence can sit in a captured graph.
void launch_sort_particles(const int32_t* cell, const int32_t* alive, int Np, int C, int32_t* pstart, int32_t* perm,
                           int32_t* scratch, hipStream_t s) {
  // pstart: C + 2 entries (cells, dead bucket, end); scratch (sort_particles_scratch):
  // cursor C + 2, scan blocks of pstart, per-block dead counts, their scan blocks
  const int n = C + 2;
  const int nb = (std::max(Np, 1) + 255) / 256;
  int32_t* cursor = scratch;
  int32_t* tmp = scratch + n;
  int32_t* bdead = tmp + (size_t)n / 1024 + 64;
  int32_t* tmp2 = bdead + nb;
  hipLaunchKernelGGL(k_zero_i32, dim3((n + 255) / 256), dim3(256), 0, s, pstart, n);
  if (Np > 0) hipLaunchKernelGGL(k_hist_alive, dim3(nb), dim3(256), 0, s, cell, alive, Np, C, pstart, bdead);
  // live counts only: after the scan pstart[C] = pstart[C + 1] = live total
  exclusive_scan(pstart, n, tmp, s);
  if (Np > 0) exclusive_scan(bdead, nb, tmp2, s);
  hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, s, pstart + C + 1, Np);
  hipLaunchKernelGGL(k_copy_i32, dim3((n + 255) / 256), dim3(256), 0, s, pstart, cursor, n);
  if (Np > 0)
    hipLaunchKernelGGL(k_pscatter, dim3(nb), dim3(256), 0, s, cell, alive, Np, C, cursor, pstart + C, bdead, perm);
}

// Store the swarm in the order of perm (slot s receives particle perm[s]).
__global__ void k_reorder(ReorderSet r, const int32_t* __restrict__ perm, int n, int64_t ld) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int src = perm[s];
  for (int a = 0; a < r.nd; a++)
    for (int row = 0; row < r.rows[a]; row++) r.dst[a][(size_t)row * ld + s] = r.src[a][(size_t)row * ld + src];
  for (int a = 0; a < r.ni; a++) r.idst[a][s] = r.isrc[a][src];
}

__global__ void k_iota(int32_t* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

hipError_t launch_reorder(const ReorderSet& r, const int32_t* perm, int n, int64_t ld, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!perm) return hipErrorInvalidValue;
  if (r.nd < 0 || r.nd > 6 || r.ni < 0 || r.ni > 4) return hipErrorInvalidValue;
  for (int a = 0; a < r.nd; a++)
    if (r.rows[a] > 0 && (!r.src[a] || !r.dst[a])) return hipErrorInvalidValue;
  for (int a = 0; a < r.ni; a++)
    if (!r.isrc[a] || !r.idst[a]) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_reorder, dim3((n + 255) / 256), dim3(256), 0, s, r, perm, n, ld);
  return hipSuccess;
}

void launch_iota(int32_t* a, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, a, n);
}

__global__ void k_sorted_coords(const double* __restrict__ x, int n, int dim, const int32_t* __restrict__ cn,
                                double* __restrict__ xs) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int j = cn[s];
  for (int d = 0; d < dim; d++) xs[(size_t)d * n + s] = x[(size_t)j * dim + d];
}

void launch_sorted_coords(const double* x, int n, int dim, const int32_t* cn, double* xs, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sorted_coords, dim3((n + 255) / 256), dim3(256), 0, s, x, n, dim, cn, xs);
}


// --------------------------------------------------------------------------
// RK stage loop (runge_kutta_{1,2,4}.c, dormand_prince.c): x = x0 + s*sum_j a_j v_j,
// optional sphere projection (advection.c:472-491), re-location (removal of
// points that left the domain, DMSwarmMigrate) and Shepard weights, fused.
// The update is evaluated without FMA contraction, in the reference's order.
// --------------------------------------------------------------------------
__device__ __forceinline__ void project_point(double* x, int dim, double rmin, double rmax) {
#pragma clang fp contract(off)
  double pos = 0.0, nrm[3];
  for (int d = 0; d < dim; d++) pos += x[d] * x[d];
  if (pos < rmax * rmax && pos > rmin * rmin) return;
  pos = sqrt(pos);
  for (int d = 0; d < dim; d++) nrm[d] = x[d] / pos;
  if (pos > rmax) for (int d = 0; d < dim; d++) x[d] = rmax * nrm[d];
  else for (int d = 0; d < dim; d++) x[d] = rmin * nrm[d];
}

__global__ void k_stage(Grid g, const double* __restrict__ pux, const double* __restrict__ puy, const double* __restrict__ puz,
                        double* px, const double* x0, double* __restrict__ vel, int64_t ld, int Np,
                        int32_t* __restrict__ cell, int32_t* __restrict__ alive, unsigned long long* removed,
                        double* __restrict__ w, StageArgs a) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || !alive[i]) return;
  const int dim = g.dim;
  // every operand is loaded before anything is stored, so the loads are in flight
  // together (the targets may alias the sources: x0 is px for one-stage updates, and
  // vel one of the stage velocities; each element is read before it is written)
  double x[3] = {0.0, 0.0, 0.0}, xs[3] = {0.0, 0.0, 0.0}, vs[7][3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d < dim) {
      x[d] = px[d * ld + i];
      xs[d] = x0[d * ld + i];
    }
#pragma unroll
    for (int j = 0; j < 7; j++) vs[j][d] = (d < dim && j < a.nv) ? a.v[j][d * ld + i] : 0.0;
  }
  const double scale = a.dt_req ? (a.dt_req[a.req[i]] * a.mul) / a.div : a.scale;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d >= dim) break;
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < 7; j++)
      if (j < a.nv) v += a.a[j] * vs[j][d];
    if (a.write_vel) vel[d * ld + i] = v;
    x[d] = xs[d] + scale * v;
    if (a.proj_in_loop) project_point(x, dim, a.rmin, a.rmax);
  }
  if (a.project && !a.proj_in_loop) project_point(x, dim, a.rmin, a.rmax);
  for (int d = 0; d < dim; d++) px[d * ld + i] = x[d];
  const int c = locate(g, x[0], x[1], x[2]);
  cell[i] = c;
  if (c < 0) {
    alive[i] = 0;
    if (a.suspend) a.suspend[i] |= 1;
    else atomicAdd(removed, 1ull);
    return;
  }
  double wl[8];
  weights_point(g, pux, puy, puz, x[0], x[1], x[2], c, wl);
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (k < g.nen) w[(size_t)k * ld + i] = wl[k];
}

void launch_stage(const Grid& g, const double* pux, const double* puy, const double* puz, double* px, const double* x0,
                  double* vel, int64_t ld, int Np, int32_t* cell, int32_t* alive, unsigned long long* removed, double* w,
                  const StageArgs& a, hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_stage, dim3((Np + 255) / 256), dim3(256), 0, s, g, pux, puy, puz, px, x0, vel, ld, Np, cell, alive, removed, w, a);
}

__global__ void k_save(double* __restrict__ px, double* __restrict__ x0, const double* __restrict__ vel, double* __restrict__ v0,
                       int64_t ld, int Np, int dim, const int32_t* __restrict__ alive) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || !alive[i]) return;
  for (int d = 0; d < dim; d++) {
    x0[d * ld + i] = px[d * ld + i];
    if (v0) v0[d * ld + i] = vel[d * ld + i];
  }
}

void launch_save(double* px, double* x0, const double* vel, double* v0, int64_t ld, int Np, int dim, const int32_t* alive,
                 hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_save, dim3((Np + 255) / 256), dim3(256), 0, s, px, x0, vel, v0, ld, Np, dim, alive);
}

__global__ void k_copy_rows(const double* __restrict__ src, int64_t lds, double* __restrict__ dst, int64_t ldd, int nrows,
                            int Np, const int32_t* __restrict__ alive) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || (alive && !alive[i])) return;
  for (int r = 0; r < nrows; r++) dst[r * ldd + i] = src[r * lds + i];
}

void launch_copy_rows(const double* src, int64_t lds, double* dst, int64_t ldd, int nrows, int Np, const int32_t* alive,
                      hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_copy_rows, dim3((Np + 255) / 256), dim3(256), 0, s, src, lds, dst, ldd, nrows, Np, alive);
}

// CFL dt = min_p cfl * (1/eps_w) / (|v_p| + 1e-16), advection.c:430-470.
// Positive doubles order like their bit patterns, so atomicMin on the bits works.
__global__ __launch_bounds__(256) void k_cfl(const double* __restrict__ vel, int64_t ld, int Np, int dim,
                                             const int32_t* __restrict__ alive, double c, unsigned long long* out) {
#pragma clang fp contract(off)
  __shared__ double red[256 / WAVE];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double dt = DBL_MAX;
  if (i < Np && alive[i]) {
    double nv = 0.0;
    for (int d = 0; d < dim; d++) nv += vel[d * ld + i] * vel[d * ld + i];
    nv = sqrt(nv);
    dt = c / (nv + 1e-16);
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) { double y = __shfl_xor(dt, o); dt = dt < y ? dt : y; }
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = dt;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < 256 / WAVE; k++) m = m < red[k] ? m : red[k];
    atomicMin(out, (unsigned long long)__double_as_longlong(m));
  }
}

void launch_cfl(const double* vel, int64_t ld, int Np, int dim, const int32_t* alive, double c, unsigned long long* out,
                hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_cfl, dim3((Np + 255) / 256), dim3(256), 0, s, vel, ld, Np, dim, alive, c, out);
}

// DOPRI error estimate, dormand_prince.c:37-102 (max over particles).
struct VPtrs { const double* v[7]; };
__global__ __launch_bounds__(256) void k_dopri_err(const double* __restrict__ px, const double* __restrict__ x0, VPtrs vp,
                                                   int64_t ld, int Np, int dim, const int32_t* __restrict__ alive, double dt,
                                                   double atol, double rtol, unsigned long long* maxb, unsigned int* count) {
#pragma clang fp contract(off)
  const double Bs[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
  __shared__ unsigned long long red_m[256 / WAVE];
  __shared__ unsigned int red_c[256 / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  // bit pattern of E (>= 0, so bit order = value order); 0 for no contribution
  unsigned long long eb = 0ull;
  unsigned int cnt = 0u;
  if (i < Np && alive[i]) {
    double E = 0.0;
    for (int d = 0; d < dim; d++) {
      double vl = 0.0;
      for (int j = 0; j < 7; j++) vl += Bs[j] * vp.v[j][d * ld + i];
      const double a0 = fabs(x0[d * ld + i]), a1 = fabs(px[d * ld + i]);
      const double mv = a0 < a1 ? a1 : a0;
      const double tol = atol + rtol * mv;
      const double xl = x0[d * ld + i] + dt * vl;
      const double e = px[d * ld + i] - xl;
      E += (e * e) / (tol * tol);
    }
    E = sqrt(1.0 / dim * E);
    if (E >= 0.0) eb = (unsigned long long)__double_as_longlong(E);
    cnt = 1u;
  }
  // block reduction, then one atomic per block (max and count are exact in any order)
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(eb, o);
    eb = eb > y ? eb : y;
    cnt += __shfl_xor(cnt, o);
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) { red_m[threadIdx.x / WAVE] = eb; red_c[threadIdx.x / WAVE] = cnt; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red_m[0];
    unsigned int c = red_c[0];
    for (int k = 1; k < (int)(blockDim.x / WAVE); k++) { m = m > red_m[k] ? m : red_m[k]; c += red_c[k]; }
    if (c) {
      atomicMax(maxb, m);
      atomicAdd(count, c);
    }
  }
}

void launch_dopri_err(const double* px, const double* x0, const double* const* vadv, int64_t ld, int Np, int dim,
                      const int32_t* alive, double dt, double atol, double rtol, unsigned long long* maxb,
                      unsigned int* count, hipStream_t s) {
  VPtrs vp;
  for (int j = 0; j < 7; j++) vp.v[j] = vadv[j];
  if (Np > 0) hipLaunchKernelGGL(k_dopri_err, dim3((Np + 255) / 256), dim3(256), 0, s, px, x0, vp, ld, Np, dim, alive, dt, atol, rtol, maxb, count);
}

// Per-request form of k_dopri_err (serving): each particle uses its request's dt and
// reduces into its request's maximum.
// Per-request maxima are reduced in LDS first (one slot per request of the
// batch), then one global atomic per request present in the block.
__global__ __launch_bounds__(256) void k_dopri_err_req(const double* __restrict__ px, const double* __restrict__ x0,
                                                       VPtrs vp, int64_t ld, int Np, int dim,
                                                       const int32_t* __restrict__ alive, const int32_t* __restrict__ req,
                                                       const double* __restrict__ dt_req, double atol, double rtol,
                                                       unsigned long long* maxb_req, int nslots) {
#pragma clang fp contract(off)
  extern __shared__ unsigned long long smax[];
  for (int k = threadIdx.x; k < nslots; k += blockDim.x) smax[k] = 0ull;
  __syncthreads();
  const double Bs[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < Np && alive[i]) {
    const int r = req[i];
    const double dt = dt_req[r];
    double E = 0.0;
    for (int d = 0; d < dim; d++) {
      double vl = 0.0;
      for (int j = 0; j < 7; j++) vl += Bs[j] * vp.v[j][d * ld + i];
      const double a0 = fabs(x0[d * ld + i]), a1 = fabs(px[d * ld + i]);
      const double mv = a0 < a1 ? a1 : a0;
      const double tol = atol + rtol * mv;
      const double xl = x0[d * ld + i] + dt * vl;
      const double e = px[d * ld + i] - xl;
      E += (e * e) / (tol * tol);
    }
    E = sqrt(1.0 / dim * E);
    // E >= 0: the order of the bit patterns is the order of the values
    if (E >= 0.0) atomicMax(&smax[r], (unsigned long long)__double_as_longlong(E));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nslots; k += blockDim.x)
    if (smax[k]) atomicMax(&maxb_req[k], smax[k]);
}

void launch_dopri_err_req(const double* px, const double* x0, const double* const* vadv, int64_t ld, int Np, int dim,
                          const int32_t* alive, const int32_t* req, const double* dt_req, double atol, double rtol,
                          unsigned long long* maxb_req, int nslots, hipStream_t s) {
  VPtrs vp;
  for (int j = 0; j < 7; j++) vp.v[j] = vadv[j];
  if (Np > 0)
    hipLaunchKernelGGL(k_dopri_err_req, dim3((Np + 255) / 256), dim3(256), sizeof(unsigned long long) * nslots, s, px, x0,
                       vp, ld, Np, dim, alive, req, dt_req, atol, rtol, maxb_req, nslots);
}

// Particles of finished requests: final position into the request's output rows
// (row = the particle's index within its request), then out of the swarm.
__global__ void k_retire(const double* __restrict__ px, int64_t ld, int Np, int dim, int32_t* __restrict__ alive,
                         const int32_t* __restrict__ req, const int32_t* __restrict__ porig,
                         const uint8_t* __restrict__ done, double* const* __restrict__ out_req) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np) return;
  const int r = req[i];
  if (!done[r]) return;
  if (alive[i]) {
    double* o = out_req[r] + (size_t)porig[i] * dim;
    for (int d = 0; d < dim; d++) o[d] = px[d * ld + i];
  }
  alive[i] = 0;
}

void launch_retire(const double* px, int64_t ld, int Np, int dim, int32_t* alive, const int32_t* req,
                   const int32_t* porig, const uint8_t* done, double* const* out_req, hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_retire, dim3((Np + 255) / 256), dim3(256), 0, s, px, ld, Np, dim, alive, req, porig,
                                 done, out_req);
}

__global__ void k_fill_i32_v(int32_t* a, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

void launch_fill_i32(int32_t* a, int n, int32_t v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_i32_v, dim3((n + 255) / 256), dim3(256), 0, s, a, n, v);
}

// --------------------------------------------------------------------------
__global__ void k_aos_to_soa(const double* __restrict__ aos, int n, int bs, double* __restrict__ soa, int64_t ld) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int b = 0; b < bs; b++) soa[b * ld + i] = aos[(size_t)i * bs + b];
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, int64_t ld, int n, int bs, double* __restrict__ aos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int b = 0; b < bs; b++) aos[(size_t)i * bs + b] = soa[b * ld + i];
}
void launch_aos_to_soa(const double* aos, int n, int bs, double* soa, int64_t ld, hipStream_t s) {
  if (n > 0 && bs > 0) hipLaunchKernelGGL(k_aos_to_soa, dim3((n + 255) / 256), dim3(256), 0, s, aos, n, bs, soa, ld);
}
void launch_soa_to_aos(const double* soa, int64_t ld, int n, int bs, double* aos, hipStream_t s) {
  if (n > 0 && bs > 0) hipLaunchKernelGGL(k_soa_to_aos, dim3((n + 255) / 256), dim3(256), 0, s, soa, ld, n, bs, aos);
}

}  // namespace lrb
