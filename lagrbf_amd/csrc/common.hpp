// Device-side building blocks shared by the gfx950 kernels of liblagrbf.so.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace lrb {

// KernelType, include/kernels/kernels_rbf.h:34-42, plus K_PHS3 (extension, no
// reference counterpart): the polyharmonic spline phi = (eps r)^3, used with a
// polynomial block.  K_GAUSS_LOCAL is internal to the evaluation: a Gaussian whose
// shape parameter belongs to the patch (eps_p = c / h_local), not to the grid.
enum KernelType { K_LINEAR = 0, K_WC0 = 1, K_WC2 = 2, K_BUMP = 3, K_MQ = 4, K_IMQ = 5, K_GAUSS = 6, K_PHS3 = 7,
                  K_GAUSS_LOCAL = 8 };
constexpr int K_MAX_OPTION = 7;   // largest kernel_type accepted as an option

// Polynomial block of a patch system (extension): monomials of degree <= deg in
// the member coordinates shifted to the patch's PU node and scaled by 1 / (largest
// cell width), ordered 1, x, y, (z), xx, xy, (xz), yy, (yz), (zz).
constexpr int POLY_MAX = 10;
__host__ __device__ inline int poly_terms(int dim, int deg) {
  return deg < 0 ? 0 : (deg == 0 ? 1 : (deg == 1 ? dim + 1 : (dim + 1) * (dim + 2) / 2));
}
// Degree a patch of n members takes under a requested degree D: the largest k <= D
// whose term count is at most max(1, n / 2); -1 for D < 0 or an empty patch.
__host__ __device__ inline int patch_degree(int dim, int D, int n) {
  if (D < 0 || n <= 0) return -1;
  const int lim = n / 2 > 1 ? n / 2 : 1;
  for (int k = D; k >= 0; k--)
    if (poly_terms(dim, k) <= lim) return k;
  return -1;
}
template <typename T>
__host__ __device__ inline void poly_eval(int dim, int deg, const T* xi, T* out) {
  if (deg < 0) return;
  int k = 0;
  out[k++] = T(1);
  if (deg >= 1)
    for (int a = 0; a < dim; a++) out[k++] = xi[a];
  if (deg >= 2)
    for (int a = 0; a < dim; a++)
      for (int b = a; b < dim; b++) out[k++] = xi[a] * xi[b];
}

// Uniform PU grid: n nodes per axis on [O, L] (L is the max corner), nc = n-1 cells.
struct Grid {
  int dim, nen;
  int n[3], nc[3];
  double O[3], L[3], h[3];
  double eps_w, eps_r;
  int P, C;
  int wnorm;   // Shepard normalisation: 0 over every PU node (swarm_interp.c:327-372), 1 over the
               // cell's 2^d corners only (extension: the used weights then sum to 1)
};

// phi(r, eps) of kernels_rbf.c:31-91 for a kernel fixed at compile time.
template <int K, typename T>
__device__ __forceinline__ T rbf(T r, T eps) {
  if constexpr (K == K_LINEAR) {
    return eps * r;
  } else if constexpr (K == K_WC0) {
    // max(u, 0): u <= 0 gives exactly 0, as the reference's cut-off
    const T u = fmax(T(1) - eps * r, T(0));
    return u * u;
  } else if constexpr (K == K_WC2) {
    const T u = fmax(T(1) - eps * r, T(0));
    const T u2 = u * u;
    return (u2 * u2) * (T(5) - T(4) * u);
  } else if constexpr (K == K_BUMP) {
    T u = T(1) - eps * r;
    return u < T(0) ? T(0) : exp(T(-1) / (u * (T(2) - u)));
  } else if constexpr (K == K_MQ) {
    T u = eps * r;
    return sqrt(T(1) + u * u);
  } else if constexpr (K == K_IMQ) {
    T u = eps * r;
    return T(1) / sqrt(T(1) + u * u);
  } else if constexpr (K == K_PHS3) {
    T u = eps * r;
    return u * u * u;
  } else {
    T u = eps * r;
    return exp(-u * u);
  }
}

// Runtime-dispatched variant for the cold paths (assembly fallbacks, tests).
template <typename T>
__device__ __forceinline__ T rbf_rt(int k, T r, T eps) {
  switch (k) {
    case K_LINEAR: return rbf<K_LINEAR, T>(r, eps);
    case K_WC0: return rbf<K_WC0, T>(r, eps);
    case K_WC2: return rbf<K_WC2, T>(r, eps);
    case K_BUMP: return rbf<K_BUMP, T>(r, eps);
    case K_MQ: return rbf<K_MQ, T>(r, eps);
    case K_IMQ: return rbf<K_IMQ, T>(r, eps);
    case K_PHS3: return rbf<K_PHS3, T>(r, eps);
    default: return rbf<K_GAUSS, T>(r, eps);
  }
}

// sqrt for the hot loops: v_rsq_f64 (about 23 correct bits) refined by two
// Newton steps on sqrt (r <- r + (x - r^2) / (2 sqrt x)), within 1 ulp for
// x >= 1e-300; the IEEE-rounded libm sequence costs twice the instructions.
// x = 0 gives exactly 0 (rsq of x + 1e-300 is finite, r = x * y = 0); for x >= 1e-284
// the added 1e-300 is below half an ulp and changes nothing.
__device__ __forceinline__ double sqrt_fast(double x) {
  const double y = __builtin_amdgcn_rsq(x + 1e-300);
  const double hy = 0.5 * y;
  double r = x * y;
  r = fma(hy, fma(-r, r, x), r);
  r = fma(hy, fma(-r, r, x), r);
  return r;
}
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

// Particle -> cell, lexicographic cell index (i fastest) of the uniform Q1 grid,
// as DMLocatePoints on the PU DMDA (swarm_utils.c:995-1030).  A coordinate with
// O <= x <= L is inside, x == L falls in the last cell, anything else (incl. NaN)
// returns -1.  IEEE subtract, correctly rounded divide and floor: host and device
// agree bit for bit.
__device__ __forceinline__ int locate_axis(double x, double O, double L, double h, int nc) {
  if (!(x >= O && x <= L)) return -1;
  double t = (x - O) / h;   // no multiply => nothing to contract
  int i = (int)floor(t);
  if (i >= nc) i = nc - 1;
  if (i < 0) i = 0;
  return i;
}

__device__ __forceinline__ int locate(const Grid& g, double x, double y, double z) {
  int i = locate_axis(x, g.O[0], g.L[0], g.h[0], g.nc[0]);
  int j = locate_axis(y, g.O[1], g.L[1], g.h[1], g.nc[1]);
  if (i < 0 || j < 0) return -1;
  if (g.dim == 3) {
    int k = locate_axis(z, g.O[2], g.L[2], g.h[2], g.nc[2]);
    if (k < 0) return -1;
    return i + g.nc[0] * (j + g.nc[1] * k);
  }
  return i + g.nc[0] * j;
}

// Q1 node k of a cell (DMDAGetElements order: counter-clockwise bottom face, then top).
__device__ __forceinline__ int corner_node(const Grid& g, int cell, int k) {
  int i = cell % g.nc[0];
  int j = (cell / g.nc[0]) % g.nc[1];
  int kk = g.dim == 3 ? cell / (g.nc[0] * g.nc[1]) : 0;
  int di = ((k & 3) == 1 || (k & 3) == 2) ? 1 : 0;
  int dj = ((k & 3) >= 2) ? 1 : 0;
  int dk = k >> 2;
  return (i + di) + g.n[0] * ((j + dj) + g.n[1] * (kk + dk));
}

// Offset of Q1 corner k: (di, dj, dk)
__host__ __device__ __forceinline__ void corner_offset(int k, int& di, int& dj, int& dk) {
  di = ((k & 3) == 1 || (k & 3) == 2) ? 1 : 0;
  dj = ((k & 3) >= 2) ? 1 : 0;
  dk = k >> 2;
}

}  // namespace lrb
