// Per-particle adaptive Dormand-Prince 5(4) with a velocity blended linearly in
// time between two data snapshots.  No reference counterpart: the reference's
// DOPRI (dormand_prince.c:179-310) takes one global step size; here every
// particle carries its own time, step size and accept/reject decision, with the
// same tableau, error norm and step-size controller applied per particle.
#include "kernels.hpp"

namespace lrb {

struct VPtrs7 { const double* v[7]; };
static constexpr int DP45_OUT_RETRIES = 2;   // forced rejections of one particle before it is removed
static constexpr int WAVE = 64;   // wavefront width of gfx950

// out = (1 - tau) vA + tau vB at the particle's stage time t + c dt,
// tau = (t - tA) / (tB - tA) (out may be vA)
__global__ void k_blend(double* out, const double* va, const double* __restrict__ vb, int64_t ld, int Np, int dim,
                        const int32_t* __restrict__ alive, const double* __restrict__ tp, const double* __restrict__ dtp,
                        double cst, double tA, double tB) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np || !alive[i]) return;
  const double t = tp[i] + cst * dtp[i];
  const double tau = (t - tA) / (tB - tA);
  for (int d = 0; d < dim; d++) {
    const size_t k = (size_t)d * ld + i;
    out[k] = (1.0 - tau) * va[k] + tau * vb[k];
  }
}

void launch_blend(double* out, const double* va, const double* vb, int64_t ld, int Np, int dim, const int32_t* alive,
                  const double* tp, const double* dtp, double cst, double tA, double tB, hipStream_t s) {
  if (Np > 0)
    hipLaunchKernelGGL(k_blend, dim3((Np + 255) / 256), dim3(256), 0, s, out, va, vb, ld, Np, dim, alive, tp, dtp, cst,
                       tA, tB);
}

// Accept/reject of one attempt per particle.  px holds y1 = x0 + dt sum b_j k_j
// (the seventh stage point) and k[6] the velocity there.  Error norm and
// controller as dormand_prince.c:37-102 and 179-214, per particle:
//   E = sqrt(1/d sum_d (e_d / (atol + rtol max(|x0_d|, |y1_d|)))^2),
//   e = y1 - (x0 + dt sum b*_j k_j),
//   dt <- dt min(fac_max', max(fac_min, fac E^(-1/5))), fac_max' = fac_max if E <= 1 else 1.
// Accepted (E <= 1): x0 <- y1, t += dt, k1 <- k7 (first same as last); a particle
// whose time reaches t_end finishes (alive 0, fin 1).  Rejected: px <- x0.  A
// particle suspended because a trial point of stages 2-6 left the domain is
// rejected the same way with dt <- dt fac_min, twice per integration at most: the
// third (or a step size below 1e-10 max(1, |t_end|)) removes it, its path leaving
// the domain; a step's end point (stage 7) outside the domain removes a particle
// (k_stage).  susp: bit 0 suspended in this attempt, the rest the count.
// The next step size is clipped to t_end - t.  counts: accepted, rejected, still active.
__global__ __launch_bounds__(256) void k_dp45_control(Grid g, int32_t* __restrict__ cell, double* __restrict__ px,
                                                      double* __restrict__ x0, VPtrs7 kv,
                                                      double* __restrict__ k1, int64_t ld, int Np, int dim,
                                                      int32_t* __restrict__ alive, int32_t* __restrict__ fin,
                                                      int32_t* __restrict__ susp,
                                                      unsigned long long* __restrict__ removed,
                                                      double* __restrict__ tp, double* __restrict__ dtp, double t_end,
                                                      double atol, double rtol, double fac, double fac_min,
                                                      double fac_max, unsigned long long* __restrict__ counts) {
#pragma clang fp contract(off)
  __shared__ unsigned int red[3][256 / WAVE];
  const double Bs[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int acc = 0, rej = 0, act = 0;
  if (i < Np && !alive[i] && susp && (susp[i] & 1)) {
    // a trial point left the domain: rejected, back to x0, the step size shrunk by
    // fac_min (the controller's factor for a non-finite error).  The third such
    // rejection of the particle's integration, or a step size below
    // 1e-10 max(1, |t_end|), means its path leaves the domain: removed (a path that
    // leaves does so at its stage points c = 1 and at the step's end point alike,
    // so without a bound it would creep toward the wall by ever shorter steps)
    for (int d = 0; d < dim; d++) {
      const size_t k = (size_t)d * ld + i;
      px[k] = x0[k];
    }
    const int nrej = (susp[i] >> 1) + 1;
    susp[i] = nrej << 1;
    rej = 1;
    double dn = dtp[i] * ((1.0 < fac_min) ? 1.0 : fac_min);
    if (nrej > DP45_OUT_RETRIES || dn <= 1e-10 * fmax(1.0, fabs(t_end))) {
      susp[i] = 0;
      atomicAdd(removed, 1ull);
    } else {
      // its cell id is the out-of-domain point's (-1): x0's again, for the grouping
      cell[i] = locate(g, x0[i], x0[ld + i], dim == 3 ? x0[2 * ld + i] : 0.0);
      alive[i] = 1;
      const double rest = t_end - tp[i];
      if (dn > rest) dn = rest;
      dtp[i] = dn;
      act = 1;
    }
  } else if (i < Np && alive[i]) {
    const double dt = dtp[i];
    double E = 0.0;
    for (int d = 0; d < dim; d++) {
      const size_t k = (size_t)d * ld + i;
      double vl = 0.0;
      for (int j = 0; j < 7; j++) vl += Bs[j] * kv.v[j][k];
      const double a0 = fabs(x0[k]), a1 = fabs(px[k]);
      const double mv = a0 < a1 ? a1 : a0;
      const double tol = atol + rtol * mv;
      const double xl = x0[k] + dt * vl;
      const double e = px[k] - xl;
      E += (e * e) / (tol * tol);
    }
    E = sqrt(1.0 / dim * E);
    const bool ok = E <= 1.0;   // NaN: rejected
    double t = tp[i];
    if (ok) {
      for (int d = 0; d < dim; d++) {
        const size_t k = (size_t)d * ld + i;
        x0[k] = px[k];
        k1[k] = kv.v[6][k];
      }
      t = t + dt;
      tp[i] = t;
      acc = 1;
    } else {
      for (int d = 0; d < dim; d++) {
        const size_t k = (size_t)d * ld + i;
        px[k] = x0[k];
      }
      rej = 1;
    }
    const double fmx = ok ? fac_max : 1.0;
    const double g = fac * pow(1.0 / E, 1.0 / 5.0);
    const double inner = (fac_min < g) ? g : fac_min;   // NaN g -> fac_min
    const double f = (fmx < inner) ? fmx : inner;
    double dn = dt * f;
    const double rest = t_end - t;
    if (rest <= 1e-12 * fmax(1.0, fabs(t_end))) {
      alive[i] = 0;
      fin[i] = 1;
    } else {
      if (dn > rest) dn = rest;
      dtp[i] = dn;
      act = 1;
    }
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    acc += __shfl_xor(acc, o);
    rej += __shfl_xor(rej, o);
    act += __shfl_xor(act, o);
  }
  const int wv = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) { red[0][wv] = acc; red[1][wv] = rej; red[2][wv] = act; }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long s = 0;
    for (int k = 0; k < (int)(blockDim.x / WAVE); k++) s += red[threadIdx.x][k];
    if (s) atomicAdd(&counts[threadIdx.x], s);
  }
}

void launch_dp45_control(const Grid& g, int32_t* cell, double* px, double* x0, const double* const* k, double* k1,
                         int64_t ld, int Np, int dim,
                         int32_t* alive, int32_t* fin, int32_t* susp, unsigned long long* removed, double* tp,
                         double* dtp, double t_end, double atol, double rtol, double fac, double fac_min,
                         double fac_max, unsigned long long* counts, hipStream_t s) {
  VPtrs7 kv;
  for (int j = 0; j < 7; j++) kv.v[j] = k[j];
  if (Np > 0)
    hipLaunchKernelGGL(k_dp45_control, dim3((Np + 255) / 256), dim3(256), 0, s, g, cell, px, x0, kv, k1, ld, Np, dim,
                       alive, fin,
                       susp, removed, tp, dtp, t_end, atol, rtol, fac, fac_min, fac_max, counts);
}

// Per-particle state at the start: t = t0, dt = min(dt0, t_end - t0), not finished.
__global__ void k_dp45_init(int Np, const int32_t* __restrict__ alive, double* __restrict__ tp, double* __restrict__ dtp,
                            int32_t* __restrict__ fin, int32_t* __restrict__ self, double t0, double dt0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Np) return;
  tp[i] = t0;
  dtp[i] = dt0;
  fin[i] = 0;
  self[i] = i;
}

void launch_dp45_init(int Np, const int32_t* alive, double* tp, double* dtp, int32_t* fin, int32_t* self, double t0,
                      double dt0, hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_dp45_init, dim3((Np + 255) / 256), dim3(256), 0, s, Np, alive, tp, dtp, fin, self, t0, dt0);
}

// End of the integration: finished particles are live again (removed ones stay out).
__global__ void k_dp45_finish(int Np, int32_t* __restrict__ alive, const int32_t* __restrict__ fin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < Np && fin[i]) alive[i] = 1;
}

void launch_dp45_finish(int Np, int32_t* alive, const int32_t* fin, hipStream_t s) {
  if (Np > 0) hipLaunchKernelGGL(k_dp45_finish, dim3((Np + 255) / 256), dim3(256), 0, s, Np, alive, fin);
}

}  // namespace lrb
