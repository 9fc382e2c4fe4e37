// Matrix-core building blocks of the Cholesky kernels (chol_mfma.hip, the
// single-workgroup factor; chol_tiled.hip, the multi-workgroup one): the 16x16x4
// MFMA in FP32 and FP64 with its lane layout, lane broadcasts, and the in-register
// factor + inverse of a 16 x 16 diagonal block.
#pragma once
#include <hip/hip_runtime.h>

namespace lrb {

typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef double mf_d4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; its 4
// result registers hold D[row(l, v)][l % 16].  The row mapping differs between the
// FP32 (consecutive rows) and FP64 (rows 4 apart) forms.
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  using V = mf_f4;
  static __device__ __forceinline__ V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <>
struct Mfma<double> {
  using V = mf_d4;
  static __device__ __forceinline__ V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// LDS ordering between the lanes of one wave (the diagonal block is worked on by
// wave 0 alone, without workgroup barriers)
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Lane value broadcast: x of lane `src` (a compile-time constant in the callers).
__device__ __forceinline__ float lane_bcast(float x, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}
__device__ __forceinline__ double lane_bcast(double x, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
  return __hiloint2double(hi, lo);
}

// One wave: Cholesky factor of the 16 x 16 diagonal block held in the MFMA tile
// `tile` (D[c][r] = P[r][c]) and the inverse of that factor.  Lane r < 16 holds
// row r of the block in registers; the column step broadcasts the pivot and the
// column with readlane, so no LDS round trip lies on the serial path.  On return
// Li (LDS, row-major) holds L_JJ^-1 for the strip solves, and the factor slot holds
// L_JJ (column-major, lower) with the strictly lower part of L_JJ^-1 in its upper
// triangle, element (J + c, J + r) for r > c.  A non-positive pivot in the first w
// columns sets *fail (first one only).
template <typename T>
__device__ __forceinline__ void diag_block_rp(const typename Mfma<T>::V& tile, T* __restrict__ L, int n, int J,
                                              int w, int lane, T* Pd, T* Li, int* fail) {
  using M = Mfma<T>;
  constexpr int PLD = 17;
  const int cl = lane & 15;
#pragma unroll
  for (int v = 0; v < 4; v++) Pd[cl * PLD + M::row(lane, v)] = tile[v];
  wave_sync_lds();
  T row[16];
#pragma unroll
  for (int c = 0; c < 16; c++) row[c] = Pd[cl * PLD + c];
  // right-looking over the columns: row[c] becomes L[r][c] for c <= r
#pragma unroll
  for (int c = 0; c < 16; c++) {
    const T pcc = lane_bcast(row[c], c);
    const bool ok = pcc > T(0);
    if (!ok && c < w && lane == 0 && *fail == 0) *fail = J + c + 1;
    const T dd = sqrt(ok ? pcc : T(1));
    const T lc = row[c] / dd;   // lane c: dd itself
    row[c] = lc;
#pragma unroll
    for (int c2 = c + 1; c2 < 16; c2++) row[c2] -= lc * lane_bcast(lc, c2);
  }
  if (lane < 16) {
#pragma unroll
    for (int c = 0; c < 16; c++) Pd[lane * PLD + c] = c <= lane ? row[c] : T(0);
  }
  // X = L^-1 in place, right-looking over the rows: at step i row i is divided by
  // L[i][i] and broadcast, and every row r > i takes L[r][i] (its entry i, not read
  // after this step) times row i off its entries j <= i
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const T rd = T(1) / lane_bcast(row[i], i);
    const T m = row[i];
#pragma unroll
    for (int j = 0; j < i; j++) {
      const T xij = lane_bcast(row[j], i) * rd;
      row[j] = cl == i ? xij : (cl > i ? row[j] - m * xij : row[j]);
    }
    row[i] = cl == i ? rd : (cl > i ? -m * rd : row[i]);
  }
  if (lane < 16) {
#pragma unroll
    for (int c = 0; c < 16; c++) Li[lane * PLD + c] = c <= lane ? row[c] : T(0);
  }
  wave_sync_lds();
  for (int e = lane; e < w * w; e += 64) {
    const int c = e / w, r = e % w;
    L[(size_t)(J + c) * n + J + r] = r >= c ? Pd[r * PLD + c] : Li[c * PLD + r];
  }
}

}  // namespace lrb
