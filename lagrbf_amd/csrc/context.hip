// Host orchestration behind the lagrbf_* C ABI: options, device state, lazy patch
// set-up, PU evaluation and the RK stage loop.  All state lives in HBM; the host
// only keeps per-patch bookkeeping (P entries) and the swarm point ids.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lagrbf.h"
#include "../../include/lagrbf_io.h"
#include "ctx.hpp"
#include "io.hpp"

namespace lrb {

static thread_local std::string g_last_error;

int set_err(int code, const char* fn, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  char full[1200];
  snprintf(full, sizeof(full), "ERROR: %s() failed with error code %d: %s", fn, code, buf);
  g_last_error = full;
  fprintf(stderr, "%s\n", full);
  return code;
}

static const char* kVelocity = "velocity";


// ---------------------------------------------------------------------------
static int ctx_from_options(Ctx& c, const OptView& o) {
  const char* fn = "lagrbf_create";
  bool bad = false;
  // PUSetFromOptions, partition_of_unity.c:68-111
  int dim = 0;
  if (!o.get_int("pu_mesh_dim", dim, bad)) return set_err(E_WRONG, fn, "Option -pu_mesh_dim not provided");
  if (dim < 2 || dim > 3) return set_err(E_RANGE, fn, "Spatial dimension must be 2 or 3. Found: %d", dim);
  std::vector<int> n;
  if (!o.get_int_array("pu_mesh_n", n, bad)) return set_err(E_WRONG, fn, "Option -pu_mesh_n not provided");
  if ((int)n.size() != dim) return set_err(E_SIZ, fn, "Option -pu_mesh_n. Expected %d entries, found %d", dim, (int)n.size());
  std::vector<double> O(dim, 0.0), L(dim, 1.0), t;
  if (o.get_real_array("pu_mesh_O", t, bad)) { if ((int)t.size() != dim) return set_err(E_SIZ, fn, "Option -pu_mesh_O. Expected %d entries", dim); O = t; }
  if (o.get_real_array("pu_mesh_L", t, bad)) { if ((int)t.size() != dim) return set_err(E_SIZ, fn, "Option -pu_mesh_L. Expected %d entries", dim); L = t; }
  double overlap = 1.0, shape = 1.0;
  o.get_real("pu_mesh_overlap", overlap, bad);
  o.get_real("pu_mesh_shape_factor", shape, bad);
  if (!(shape > 0.0)) return set_err(E_RANGE, fn, "Shape factor must be positive. Found: %g", shape);
  if (bad) return set_err(E_WRONG, fn, "malformed PU option value");
  Grid& g = c.g;
  g.dim = dim; g.nen = 1 << dim; g.P = 1; g.C = 1;
  double box = 0.0;
  for (int d = 0; d < 3; d++) {
    g.n[d] = d < dim ? n[d] : 1;
    g.nc[d] = d < dim ? n[d] - 1 : 1;
    g.O[d] = d < dim ? O[d] : 0.0;
    g.L[d] = d < dim ? L[d] : 0.0;
    if (d < dim) {
      if (n[d] < 2) return set_err(E_RANGE, fn, "-pu_mesh_n must be >= 2 per axis");
      if (!(L[d] > O[d])) return set_err(E_RANGE, fn, "-pu_mesh_L must exceed -pu_mesh_O");
      g.h[d] = (L[d] - O[d]) / (double)(n[d] - 1);
      g.P *= n[d];
      g.C *= n[d] - 1;
      box = std::max(box, L[d] - O[d]);
    } else {
      g.h[d] = 1.0;
    }
  }
  // PartitionOfUnityComputeSupport / ComputeShapeFactor, partition_of_unity.c:322-355
  double ov = overlap * box;
  double radius = ov / std::pow((double)g.P, 1.0 / dim);
  g.eps_w = 1.0 / radius;
  g.eps_r = shape * g.eps_w;
  // LagRBFSetFromOptions, lagrangian_rbf.c:135-195
  o.get_int("n_interpolation_fields", c.nf, bad);
  o.get_int("kernel_type", c.ktype, bad);
  o.get_int("solve_method", c.solve_method, bad);
  o.get_real("rmin", c.rmin, bad);
  o.get_real("rmax", c.rmax, bad);
  o.get_real("cfl_fac", c.cfl_fac, bad);
  o.get_bool("advection", c.advection, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed option value");
  if (c.ktype < 0 || c.ktype > K_MAX_OPTION) return set_err(E_RANGE, fn, "Unknown kernel type %d", c.ktype);
  // extensions: polynomial block ([A P; P^T 0]) and per-patch shape parameter
  o.get_int("lagrbf_poly_degree", c.poly_deg, bad);
  o.get_real("lagrbf_local_eps", c.local_eps, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed -lagrbf_poly_degree / -lagrbf_local_eps option");
  if (c.poly_deg < -1 || c.poly_deg > 2)
    return set_err(E_RANGE, fn, "-lagrbf_poly_degree must be -1 (none), 0, 1 or 2 (found %d)", c.poly_deg);
  if (!(c.local_eps >= 0.0)) return set_err(E_RANGE, fn, "-lagrbf_local_eps must be >= 0");
  if (c.local_eps > 0.0 && c.ktype != K_GAUSS)
    return set_err(E_SUP, fn, "-lagrbf_local_eps applies to the Gaussian kernel (kernel_type 6)");
  std::string wn;
  if (o.get_string("lagrbf_weights", wn)) {
    if (wn == "all") c.g.wnorm = 0;
    else if (wn == "corners") c.g.wnorm = 1;
    else return set_err(E_WRONG, fn, "-lagrbf_weights must be all or corners (found %s)", wn.c_str());
  }
  if (c.poly_deg >= 0 && c.nf > 8)
    return set_err(E_SUP, fn, "a polynomial block handles up to 8 interpolated fields (found %d)", c.nf);
  c.poly_mm = poly_terms(dim, c.poly_deg);
  if (c.solve_method < 0 || c.solve_method > 2) return set_err(E_RANGE, fn, "Unknown method type: %d", c.solve_method);
  if (c.nf < 0) return set_err(E_RANGE, fn, "-n_interpolation_fields must be >= 0");
  if (c.advection) {
    c.scheme = 3;
    o.get_int("advection_scheme", c.scheme, bad);
    if (c.scheme < 0 || c.scheme > 3) return set_err(E_WRONG, fn, "Advection scheme %d Create function not found", c.scheme);
    static const int ord[4] = {1, 2, 4, 7};
    c.order = ord[c.scheme];
    o.get_string("velocity_name", c.velocity_name);
    o.get_real("dt", c.adv_dt, bad);
    o.get_real("dopri_dt_fac", c.dopri_fac, bad);
    o.get_real("dopri_dt_fac_min", c.dopri_fac_min, bad);
    o.get_real("dopri_dt_fac_max", c.dopri_fac_max, bad);
    o.get_real("dopri_atol", c.dopri_atol, bad);
    o.get_real("dopri_rtol", c.dopri_rtol, bad);
    if (bad) return set_err(E_WRONG, fn, "malformed advection option value");
  }
  o.get_int("lagrbf_device", c.device, bad);
  std::string pr;
  if (o.get_string("lagrbf_precision", pr)) {
    if (pr == "fp64") c.precision = 0;
    else if (pr == "mixed") c.precision = 1;
    else return set_err(E_WRONG, fn, "-lagrbf_precision must be fp64 or mixed (found %s)", pr.c_str());
  }
  std::string mf;
  if (o.get_string("lagrbf_mixed_factor", mf)) {
    if (mf == "fp32") c.mixed_factor64 = false;
    else if (mf == "fp64") c.mixed_factor64 = true;
    else return set_err(E_WRONG, fn, "-lagrbf_mixed_factor must be fp32 or fp64 (found %s)", mf.c_str());
  }
  std::string fm;
  if (o.get_string("lagrbf_factor", fm)) {
    if (fm == "auto") c.factor_mode = 0;
    else if (fm == "bk") c.factor_mode = 1;
    else return set_err(E_WRONG, fn, "-lagrbf_factor must be auto or bk (found %s)", fm.c_str());
  }
  o.get_int("lagrbf_refine_max_iter", c.refine_max_iter, bad);
  o.get_real("lagrbf_refine_tol", c.refine_tol, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed lagrbf_refine_* option");
  std::string ev;
  if (o.get_string("lagrbf_eval", ev)) {
    if (ev == "auto") c.eval_mode = 0;
    else if (ev == "cell") c.eval_mode = 1;
    else if (ev == "ref") c.eval_mode = 2;
    else return set_err(E_WRONG, fn, "-lagrbf_eval must be auto, cell or ref (found %s)", ev.c_str());
  }
  o.get_int("lagrbf_factor_nb", c.factor_nb, bad);
  if (bad || !(c.factor_nb == -1 || c.factor_nb == 0 || c.factor_nb == 16))
    return set_err(E_WRONG, fn, "-lagrbf_factor_nb must be -1 (tune), 0 (vector ALU) or 16 (matrix cores)");
  c.chol_big_min = chol_max_n();
  o.get_int("lagrbf_chol_big_min", c.chol_big_min, bad);
  c.chol_big_min = std::max(0, std::min(c.chol_big_min, chol_max_n()));
  o.get_bool("lagrbf_stepped_refine", c.stepped_refine, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed -lagrbf_stepped_refine option");
  o.get_bool("lagrbf_lean_solve", c.lean_solve, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed -lagrbf_lean_solve option");
  o.get_bool("lagrbf_tiled_solve", c.tiled_solve, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed -lagrbf_tiled_solve option");
  std::string bf;
  if (o.get_string("lagrbf_big_factor", bf)) {
    if (bf == "tiled") c.big_factor = 1;
    else if (bf == "vector") c.big_factor = 0;
    else return set_err(E_WRONG, fn, "-lagrbf_big_factor must be tiled or vector (found %s)", bf.c_str());
  }
  o.get_bool("lagrbf_reorder", c.reorder, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed -lagrbf_reorder option");
  o.get_bool("lagrbf_graph", c.use_graph, bad);
  if (bad) return set_err(E_WRONG, fn, "malformed -lagrbf_graph option");
  o.get_bool("lagrbf_setup_all", c.setup_all, bad);
  o.get_int("lagrbf_eval_variant", c.eval_variant, bad);
  if (c.eval_variant < -1 || c.eval_variant >= N_EVAL_VARIANTS)
    return set_err(E_RANGE, fn, "-lagrbf_eval_variant must be -1 (tune) or 0..%d", N_EVAL_VARIANTS - 1);
  c.tuner.forced = c.eval_variant;
  o.get_bool("lagrbf_time_blend", c.blend, bad);
  o.get_int("lagrbf_blend_fused", c.blend_fused, bad);
  std::vector<double> bt;
  if (o.get_real_array("lagrbf_blend_times", bt, bad)) {
    if (bt.size() != 2 || !(bt[1] != bt[0]))
      return set_err(E_WRONG, fn, "-lagrbf_blend_times needs two distinct times tA,tB");
    c.blend_t[0] = bt[0];
    c.blend_t[1] = bt[1];
  }
  if (bad) return set_err(E_WRONG, fn, "malformed -lagrbf_setup_all / -lagrbf_time_blend / -lagrbf_blend_times option");
  if (c.blend && c.nf != 2 * c.g.dim)
    return set_err(E_WRONG, fn, "-lagrbf_time_blend needs n_interpolation_fields = 2 x dim (found %d)", c.nf);
  return 0;
}

static int ctx_init_device(Ctx& c) {
  const char* fn = "lagrbf_create";
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return set_err(E_LIB, fn, "no HIP device available: liblagrbf runs on the GPU only");
  if (c.device < 0 || c.device >= ndev) return set_err(E_RANGE, fn, "lagrbf_device %d out of range", c.device);
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));
  // PU node coordinates per axis, x_i = O + i*h (DMDASetUniformCoordinates)
  std::vector<double> ax[3];
  for (int d = 0; d < 3; d++) {
    ax[d].resize(c.g.n[d]);
    for (int i = 0; i < c.g.n[d]; i++) ax[d][i] = c.g.O[d] + i * c.g.h[d];
  }
  HIPCHK(c.pux.reserve(c.g.n[0])); HIPCHK(c.puy.reserve(c.g.n[1])); HIPCHK(c.puz.reserve(c.g.n[2]));
  HIPCHK(hipMemcpy(c.pux.p, ax[0].data(), sizeof(double) * c.g.n[0], hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.puy.p, ax[1].data(), sizeof(double) * c.g.n[1], hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.puz.p, ax[2].data(), sizeof(double) * c.g.n[2], hipMemcpyHostToDevice));
  const int P = c.g.P;
  c.h_setup.assign(P, 0); c.h_solved.assign(P, 0); c.h_active.assign(P, 0);
  c.h_n.assign(P, 0); c.h_mem_off.assign(P, 0); c.h_fac_off.assign(P, 0); c.h_pcount.assign(P, 0);
  HIPCHK(c.p_n.reserve(P)); HIPCHK(c.p_active.reserve(P)); HIPCHK(c.p_mem_off.reserve(P));
  HIPCHK(c.p_fac_off.reserve(P)); HIPCHK(c.touched.reserve(P)); HIPCHK(c.list.reserve(P)); HIPCHK(c.info.reserve(P));
  HIPCHK(c.pcount.reserve(P));
  c.h_method.assign(P, 0);
  c.h_m.assign(P, 0); c.h_deg.assign(P, -1); c.h_eps.assign(P, c.g.eps_r); c.h_aux_off.assign(P, 0);
  c.h_fnb.assign(P, 0);
  HIPCHK(c.p_fnb.reserve(P));
  HIPCHK(hipMemset(c.p_fnb.p, 0, P));
  HIPCHK(c.p_m.reserve(P)); HIPCHK(c.p_deg.reserve(P)); HIPCHK(c.p_eps.reserve(P)); HIPCHK(c.p_aux_off.reserve(P));
  HIPCHK(hipMemset(c.p_m.p, 0, sizeof(int32_t) * P));
  HIPCHK(hipMemset(c.p_aux_off.p, 0, sizeof(int64_t) * P));
  HIPCHK(hipMemcpy(c.p_eps.p, c.h_eps.data(), sizeof(double) * P, hipMemcpyHostToDevice));
  if (c.poly_deg >= 0) {
    HIPCHK(c.beta_pool.reserve((size_t)P * c.poly_mm * std::max(c.nf, 1)));
    HIPCHK(hipMemset(c.beta_pool.p, 0, sizeof(double) * (size_t)P * c.poly_mm * std::max(c.nf, 1)));
  }
  HIPCHK(c.pstatus.reserve(P)); HIPCHK(c.piters.reserve(P));
  HIPCHK(c.list.reserve((size_t)N_METHODS * P));
  HIPCHK(hipMemset(c.p_n.p, 0, sizeof(int32_t) * P));
  HIPCHK(hipMemset(c.p_active.p, 0, P));
  HIPCHK(hipMemset(c.p_mem_off.p, 0, sizeof(int64_t) * P));
  HIPCHK(hipMemset(c.p_fac_off.p, 0, sizeof(int64_t) * P));
  HIPCHK(c.cell_start.reserve(c.g.C + 1));
  HIPCHK(hipMemset(c.cell_start.p, 0, sizeof(int32_t) * (c.g.C + 1)));
  HIPCHK(c.counters.reserve(4));
  HIPCHK(hipMemset(c.counters.p, 0, sizeof(unsigned long long) * 4));
  return 0;
}

// ---------------------------------------------------------------------------
static int check_bbox(const Ctx& c, int n, const double* x, const char* fn) {
  const int dim = c.g.dim;
  for (int p = 0; p < n; p++)
    for (int d = 0; d < dim; d++) {
      double v = x[(size_t)p * dim + d];
      // DMSwarmAssignRankFromDMDA, swarm_utils.c:862-870 (NaN rejected too)
      if (!(v >= c.g.O[d] && v <= c.g.L[d]))
        return set_err(E_WRONG, fn, "Point %d, with coordinates (%g,%g) does not belong to any DMDA local bounding box.",
                       p, x[(size_t)p * dim], x[(size_t)p * dim + 1]);
    }
  return 0;
}

int fill_data(Ctx& c, int n, const double* coords, const double* values, const int64_t* pidx) {
  const char* fn = "FillDataSwarm";
  const int dim = c.g.dim, nf = c.nf;
  CHK(check_bbox(c, n, coords, fn));
  c.blk_valid = false; c.table_gen++;
  c.tuner = EvalTuner();
  c.tuner.forced = c.eval_variant;
  if (c.setup_N >= 0 && c.solve_method != 2 && n != c.setup_N && c.mem_used > 0)
    return set_err(E_SIZ, fn, "solve_method %d keeps patch factors across data fills; node count changed from %d to %d "
                   "(call with solve_method 2 or destroy the context)", c.solve_method, c.setup_N, n);
  c.N = n;
  HIPCHK(c.dx.reserve((size_t)n * dim));
  HIPCHK(c.dval.reserve((size_t)n * std::max(nf, 1)));
  HIPCHK(c.dcell.reserve(n));
  HIPCHK(c.cell_nodes.reserve(n));
  const int C = c.g.C;
  HIPCHK(c.scratch.reserve((size_t)C + 1 + (size_t)std::max(n, 1) + (size_t)(C + 1) / 1024 + 64));
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(c.dx.p, coords, sizeof(double) * (size_t)n * dim, hipMemcpyHostToDevice, c.s));
    if (values && nf > 0) HIPCHK(hipMemcpyAsync(c.dval.p, values, sizeof(double) * (size_t)n * nf, hipMemcpyHostToDevice, c.s));
    else if (nf > 0) HIPCHK(hipMemsetAsync(c.dval.p, 0, sizeof(double) * (size_t)n * nf, c.s));
  }
  c.h_dpid.resize(n);
  for (int i = 0; i < n; i++) c.h_dpid[i] = pidx ? pidx[i] : (int64_t)i + 1;
  // DMSwarmPICSetCellId + CellInfoSetUp: node -> cell, per-cell node lists
  launch_locate_aos(c.g, c.dx.p, n, c.dcell.p, c.s);
  launch_cell_lists(c.dcell.p, n, C, c.cell_start.p, c.cell_nodes.p, c.scratch.p, c.s);
  launch_patch_count(c.g, c.cell_start.p, c.pcount.p, c.s);
  HIPCHK(hipMemcpyAsync(c.h_pcount.data(), c.pcount.p, sizeof(int32_t) * c.g.P, hipMemcpyDeviceToHost, c.s));
  // node coordinates in cell order (SoA) for the cell-ordered evaluation
  HIPCHK(c.xs_sorted.reserve((size_t)std::max(n, 1) * dim));
  launch_sorted_coords(c.dx.p, n, dim, c.cell_nodes.p, c.xs_sorted.p, c.s);
  HIPCHK(hipStreamSynchronize(c.s));
  int maxp = 1;
  for (int p = 0; p < c.g.P; p++) maxp = std::max(maxp, c.h_pcount[p]);
  HIPCHK(c.zeros.reserve((size_t)maxp * std::max(nf, 1)));
  HIPCHK(hipMemset(c.zeros.p, 0, sizeof(double) * (size_t)maxp * std::max(nf, 1)));
  if (c.precision) {
    HIPCHK(c.xs32.reserve((size_t)std::max(n, 1) * dim));
    launch_sorted_coords32(c.g, c.dx.p, n, c.cell_nodes.p, c.dcell.p, c.xs32.p, c.s);
    HIPCHK(c.zeros32.reserve((size_t)maxp * std::max(nf, 1)));
    HIPCHK(hipMemset(c.zeros32.p, 0, sizeof(float) * (size_t)maxp * std::max(nf, 1)));
    HIPCHK(hipStreamSynchronize(c.s));
  }
  // patches kept from an earlier fill (solve_method 0/1) were built from the cell
  // lists of that fill: the cell-ordered evaluation needs them unchanged
  if (c.mem_used > 0 && c.memb_current) {
    std::vector<int32_t> cs(C + 1), cn(n);
    HIPCHK(hipMemcpy(cs.data(), c.cell_start.p, sizeof(int32_t) * (C + 1), hipMemcpyDeviceToHost));
    if (n > 0) HIPCHK(hipMemcpy(cn.data(), c.cell_nodes.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (cs != c.h_cell_start_prev || cn != c.h_cell_nodes_prev) c.memb_current = false;
  }
  c.data_filled = true;
  return 0;
}

static int alloc_particles(Ctx& c) {
  const int dim = c.g.dim;
  const size_t ld = (size_t)c.ld;
  HIPCHK(c.px.reserve(ld * dim));
  HIPCHK(c.w.reserve(ld * c.g.nen));
  HIPCHK(c.fields.reserve(ld * std::max(c.nf, 1)));
  HIPCHK(c.pcell.reserve(ld));
  HIPCHK(c.alive.reserve(ld));
  HIPCHK(c.perm.reserve(ld));
  HIPCHK(c.pstart.reserve((size_t)c.g.C + 2));
  HIPCHK(c.pscratch.reserve(sort_particles_scratch(c.g.C, (int)ld)));
  HIPCHK(c.porig.reserve(ld));
  HIPCHK(c.cell_list.reserve(nonempty_list_capacity(c.g.C)));
  HIPCHK(c.list_count.reserve(1));
  c.sorted_valid = false;
  HIPCHK(c.vel.reserve(ld * dim));
  if (c.advection) {
    HIPCHK(c.x0.reserve(ld * dim));
    for (int s = 0; s < c.order; s++) HIPCHK(c.vadv[s].reserve(ld * dim));
  }
  return 0;
}

__global__ void k_fill_i32(int32_t* a, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

static int fill_points(Ctx& c, int n, const double* coords, const double* values, const int64_t* pidx) {
  const char* fn = "FillInterpSwarm";
  const int dim = c.g.dim, nf = c.nf;
  CHK(check_bbox(c, n, coords, fn));
  c.Np = n;
  c.ld = ((int64_t)n + 63) / 64 * 64;
  if (c.ld == 0) c.ld = 64;
  CHK(alloc_particles(c));
  DBuf<double> tmp;
  HIPCHK(tmp.reserve((size_t)std::max(n, 1) * std::max(dim, std::max(nf, 1))));
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(tmp.p, coords, sizeof(double) * (size_t)n * dim, hipMemcpyHostToDevice, c.s));
    launch_aos_to_soa(tmp.p, n, dim, c.px.p, c.ld, c.s);
    HIPCHK(hipMemsetAsync(c.fields.p, 0, sizeof(double) * c.ld * std::max(nf, 1), c.s));
    if (values && nf > 0) {
      HIPCHK(hipMemcpyAsync(tmp.p, values, sizeof(double) * (size_t)n * nf, hipMemcpyHostToDevice, c.s));
      launch_aos_to_soa(tmp.p, n, nf, c.fields.p, c.ld, c.s);
    }
    HIPCHK(hipMemsetAsync(c.w.p, 0, sizeof(double) * c.ld * c.g.nen, c.s));
    HIPCHK(hipMemsetAsync(c.vel.p, 0, sizeof(double) * c.ld * dim, c.s));
    hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, c.s, c.alive.p, n, 1);
    launch_iota(c.porig.p, n, c.s);
  }
  c.h_ppid.resize(n);
  for (int i = 0; i < n; i++) c.h_ppid[i] = pidx ? pidx[i] : (int64_t)i + 1;
  // DMSwarmPICSetCellId + SwarmInterpSetWeights (interface_utils.c:118-138)
  launch_locate_particles(c.g, c.px.p, c.ld, n, c.pcell.p, c.alive.p, c.counters.p, c.s);
  c.sorted_valid = false;
  c.cell_ordered = false;
  launch_weights(c.g, c.pux.p, c.puy.p, c.puz.p, c.px.p, c.ld, n, c.pcell.p, c.alive.p, c.w.p, c.s);
  HIPCHK(hipStreamSynchronize(c.s));
  c.points_filled = true;
  return 0;
}

// ---------------------------------------------------------------------------
AugView aug_view(Ctx& c, bool with_poly) {
  AugView a{};
  const bool poly = with_poly && c.poly_deg >= 0;
  a.m = poly ? c.p_m.p : nullptr;
  a.deg = poly ? c.p_deg.p : nullptr;
  a.eps = c.local_eps > 0.0 ? c.p_eps.p : nullptr;
  a.aux_off = c.p_aux_off.p;
  a.aux = c.aux_pool.p;
  a.beta = c.beta_pool.p;
  a.bstride = c.poly_mm * std::max(c.nf, 1);
  a.mmax = c.poly_mm;
  a.fnb = c.p_fnb.p;
  double hm = c.g.h[0];
  for (int d = 1; d < c.g.dim; d++) hm = std::max(hm, c.g.h[d]);
  a.ps = 1.0 / hm;
  return a;
}

int eval_ktype(const Ctx& c) { return (c.local_eps > 0.0 && c.ktype == K_GAUSS) ? K_GAUSS_LOCAL : c.ktype; }
bool eval_f32(const Ctx& c) { return c.precision && eval_f32_supported(eval_ktype(c)); }

EvalExt eval_ext(const Ctx& c, int group) {
  EvalExt e;
  if (c.poly_deg >= 0) {
    e.ptab = group < 0 ? c.ptab.p : c.ptabg[group].p;
    e.mm = c.poly_mm;
    e.deg = c.poly_deg;
    double hm = c.g.h[0];
    for (int d = 1; d < c.g.dim; d++) hm = std::max(hm, c.g.h[d]);
    e.ps = 1.0 / hm;
    e.pcell = c.pcell.p;
    e.palive = c.alive.p;
  }
  if (c.local_eps > 0.0) e.peps = c.p_eps.p;
  return e;
}

// Shape parameter of patch p with -lagrbf_local_eps c: c / h_local, h_local =
// (volume of the cells touching the PU node / members)^(1/d).
static double patch_eps(const Ctx& c, int p, int n) {
  if (!(c.local_eps > 0.0) || n <= 0) return c.g.eps_r;
  const Grid& g = c.g;
  const int a[3] = {p % g.n[0], (p / g.n[0]) % g.n[1], g.dim == 3 ? p / (g.n[0] * g.n[1]) : 0};
  double vol = 1.0;
  for (int d = 0; d < g.dim; d++) vol *= g.h[d] * (double)((a[d] > 0 ? 1 : 0) + (a[d] < g.nc[d] ? 1 : 0));
  return c.local_eps / std::pow(vol / (double)n, 1.0 / g.dim);
}

// Elements of a patch's factor slot: Bunch-Kaufman factors the (n + m) square
// saddle-point matrix, the null-space Cholesky the (n - m) square B22.
static size_t slot_size(const Ctx& c, int p, int method) {
  const size_t n = (size_t)c.h_n[p], m = (size_t)c.h_m[p];
  if (method == M_BK64) return (n + m) * (n + m);
  if (method_aug(method)) return (n - m) * (n - m);
  return n * n;
}

// Factor slots in the FP64 pool (Bunch-Kaufman, FP64 Cholesky) or the FP32 pool
// (FP32 Cholesky of the mixed path).
static int alloc_factor_slots(Ctx& c, const std::vector<int32_t>& l, int m) {
  if (l.empty()) return 0;
  size_t& used = method_fp32(m) ? c.fac32_used : c.fac_used;
  size_t need = used;
  for (int p : l) { c.h_fac_off[p] = (int64_t)need; need += slot_size(c, p, m); }
  if (method_fp32(m)) HIPCHK(c.fac32_pool.grow(need + 1, used, c.s));
  else HIPCHK(c.fac_pool.grow(need + 1, used, c.s));
  used = need;
  return 0;
}

static int upload_list(Ctx& c, const std::vector<int32_t>& l, int seg) {
  HIPCHK(hipMemcpyAsync(c.list.p + (size_t)seg * c.g.P, l.data(), sizeof(int32_t) * l.size(), hipMemcpyHostToDevice, c.s));
  return 0;
}

// ---------------------------------------------------------------------------
// Factor kernel choice for the small-patch Cholesky (the matrix-core kernel with
// panels of 16 or 32 columns; the vector-ALU kernel by option only), per precision, kernel,
// polynomial block and size bucket of the factor dimension (<= 128, 256, 320, 384,
// 512).  The first factorisation in a bucket times every candidate on a sample of
// its patches (up to 256; factoring is idempotent, the sample is factored again
// with the winner); the choice is kept for the process.  Where the lean solves
// apply (solve_llt.hip, aug_lean.hip) only the matrix-core factor is a candidate:
// those solves need its inverse diagonal blocks, and the column-oriented solve of
// the other factor costs more than any factor time it saves (at the clustered
// workload's 90,780 FP32 saddle-point patches of up to 96 members: 118 ms of
// solves against 5 ms).
struct FactorTuneKey {
  int ktype, mmax, fp32, bucket;
  bool operator<(const FactorTuneKey& o) const {
    return std::tie(ktype, mmax, fp32, bucket) < std::tie(o.ktype, o.mmax, o.fp32, o.bucket);
  }
};
static std::mutex g_ftune_mu;
static std::map<FactorTuneKey, int> g_ftune;
static constexpr int FACTOR_NB[] = {16, 0};
static constexpr int N_FACTOR_NB = (int)(sizeof(FACTOR_NB) / sizeof(FACTOR_NB[0]));

// size buckets of the factor dimension and the largest dimension of each: the
// matrix-core kernel is sized by it (16-row strips per wave), and the strips the
// largest bucket provides for 512 rows cost two waves per SIMD at ~300 rows
static constexpr int FACTOR_BUCKET_N[] = {128, 256, 320, 384, 512};
static constexpr int N_FACTOR_BUCKETS = 5;
static int factor_bucket(int nn) {
  int b = 0;
  while (b + 1 < N_FACTOR_BUCKETS && nn > FACTOR_BUCKET_N[b]) b++;
  return b;
}

static int factor_nb_for(Ctx& c, bool fp32, int bucket, const int32_t* dlist, int nsample, void* fac,
                         const AugView& a) {
  if (c.factor_nb >= 0) return c.factor_nb;
  if (c.lean_solve && (!a.m || solve_aug_lean_supported(c.ktype, fp32))) return 16;
  const FactorTuneKey key{c.ktype, a.m ? a.mmax : 0, fp32 ? 1 : 0, bucket};
  {
    std::lock_guard<std::mutex> lk(g_ftune_mu);
    auto it = g_ftune.find(key);
    if (it != g_ftune.end()) return it->second;
  }
  hipEvent_t e[N_FACTOR_NB + 1];
  for (auto& x : e) if (hipEventCreate(&x) != hipSuccess) return 16;
  float best = 1e30f, tot = 0.f;
  int nb_best = 16;
  for (int v = 0; v < N_FACTOR_NB; v++) {
    (void)hipEventRecord(e[v], c.s);
    launch_factor_chol(c.g, c.ktype, fp32, dlist, nsample, c.p_n.p, c.p_mem_off.p, c.p_fac_off.p, c.mem_pool.p, c.dx.p,
                       fac, c.pstatus.p, false, a, FACTOR_NB[v], FACTOR_BUCKET_N[bucket], c.s);
  }
  (void)hipEventRecord(e[N_FACTOR_NB], c.s);
  (void)hipEventSynchronize(e[N_FACTOR_NB]);
  for (int v = 0; v < N_FACTOR_NB; v++) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e[v], e[v + 1]);
    tot += ms;
    if (ms < best) { best = ms; nb_best = FACTOR_NB[v]; }
  }
  for (auto& x : e) (void)hipEventDestroy(x);
  c.factor_trial_ms += tot;
  std::lock_guard<std::mutex> lk(g_ftune_mu);
  g_ftune[key] = nb_best;
  return nb_best;
}

// Large patches of Cholesky method m (host list l, the same list on the device at
// dl): the tiled factor over many workgroups, or one workgroup per patch with
// -lagrbf_big_factor vector.  Records each patch's panel width for its solve.
static int factor_big(Ctx& c, int m, const std::vector<int32_t>& l, const int32_t* dl) {
  if (l.empty()) return 0;
  const AugView a = aug_view(c, method_aug(m));
  void* fac = method_fp32(m) ? (void*)c.fac32_pool.p : (void*)c.fac_pool.p;
  if (!c.big_factor) {
    for (int p : l) c.h_fnb[p] = 0;
    launch_factor_chol(c.g, c.ktype, method_fp32(m), dl, (int)l.size(), c.p_n.p, c.p_mem_off.p, c.p_fac_off.p,
                       c.mem_pool.p, c.dx.p, fac, c.pstatus.p, true, a, 0, 0, c.s);
    return 0;
  }
  std::vector<int32_t> nn(l.size());
  for (size_t i = 0; i < l.size(); i++) nn[i] = c.h_n[l[i]] - (method_aug(m) ? c.h_m[l[i]] : 0);
  // an earlier tiled launch may still read the task buffer
  HIPCHK(hipStreamSynchronize(c.s));
  tile_tasks(l, nn, c.h_tile_tasks, c.h_tile_off);
  HIPCHK(c.tile_tasks.reserve(c.h_tile_tasks.size()));
  HIPCHK(hipMemcpyAsync(c.tile_tasks.p, c.h_tile_tasks.data(), sizeof(int2) * c.h_tile_tasks.size(),
                        hipMemcpyHostToDevice, c.s));
  launch_factor_tiled(c.g, c.ktype, method_fp32(m), c.tile_tasks.p, c.h_tile_off, c.p_n.p, c.p_mem_off.p,
                      c.p_fac_off.p, c.mem_pool.p, c.dx.p, fac, c.pstatus.p, a, c.s);
  for (int p : l) c.h_fnb[p] = 16;   // 16-column blocks with stored inverses (blocked_solve.hpp)
  return 0;
}

// Refactor the listed patches in FP64 after an FP32 breakdown or a refinement
// that did not converge (counted in "fallback"): large patches of an FP32
// Cholesky method with the FP64 tiled Cholesky (a Bunch-Kaufman factor of
// thousands of members in one workgroup would take minutes), the others, and any
// whose FP64 Cholesky breaks down, with Bunch-Kaufman.  solve: also solve them.
static constexpr int BK_FALLBACK_MAX = 4096;
static int refactor_fp64(Ctx& c, const std::vector<int32_t>& ps, bool solve) {
  if (ps.empty()) return 0;
  const int P = c.g.P;
  const AugView ap = aug_view(c, true), ae = aug_view(c, false);
  c.fallback += (double)ps.size();
  std::vector<int32_t> big[2], bk;   // big[1]: with a polynomial block
  for (int p : ps) {
    const int mc = c.h_method[p];
    if (c.big_factor && method_big(mc) && method_fp32(mc)) {
      const int a = method_aug(mc) ? 1 : 0;
      big[a].push_back(p);
      c.h_method[p] = (uint8_t)(a ? M_AUG64B : M_CHOL64B);
      c.big_refactor64 += 1;
    } else {
      bk.push_back(p);
      if (method_aug(mc)) c.aug_fallback_bk += 1;
      c.h_method[p] = M_BK64;
    }
  }
  for (int a = 0; a < 2; a++) CHK(alloc_factor_slots(c, big[a], a ? M_AUG64B : M_CHOL64B));
  HIPCHK(hipMemcpyAsync(c.p_fac_off.p, c.h_fac_off.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, c.s));
  for (int a = 0; a < 2; a++) {
    if (big[a].empty()) continue;
    const int m = a ? M_AUG64B : M_CHOL64B;
    CHK(upload_list(c, big[a], m));
    CHK(factor_big(c, m, big[a], c.list.p + (size_t)m * P));
  }
  HIPCHK(hipMemcpyAsync(c.p_fnb.p, c.h_fnb.data(), P, hipMemcpyHostToDevice, c.s));
  if (!big[0].empty() || !big[1].empty()) {
    std::vector<int32_t> st(P);
    HIPCHK(hipMemcpyAsync(st.data(), c.pstatus.p, sizeof(int32_t) * P, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    for (int a = 0; a < 2; a++) {
      std::vector<int32_t> keep;
      for (int p : big[a]) {
        if (st[p] == 0) { keep.push_back(p); continue; }
        bk.push_back(p);
        if (a) c.aug_fallback_bk += 1;
        c.h_method[p] = M_BK64;
      }
      big[a].swap(keep);
    }
  }
  if (!bk.empty()) {
    // one workgroup factors a Bunch-Kaufman patch: beyond a few thousand rows that
    // takes minutes, so such a patch is reported instead
    for (int p : bk)
      if (c.h_n[p] + c.h_m[p] > BK_FALLBACK_MAX)
        return set_err(E_SUP, "PartitionFactor", "patch %d (%d members) is not positive definite in FP64 and exceeds "
                       "the %d rows of the Bunch-Kaufman fallback", p, c.h_n[p], BK_FALLBACK_MAX);
    CHK(alloc_factor_slots(c, bk, M_BK64));
    HIPCHK(hipMemcpyAsync(c.p_fac_off.p, c.h_fac_off.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, c.s));
    CHK(upload_list(c, bk, 0));
    launch_factor_bk(c.g, c.ktype, c.list.p, (int)bk.size(), c.p_n.p, c.p_mem_off.p, c.p_fac_off.p, c.mem_pool.p,
                     c.dx.p, c.fac_pool.p, c.piv_pool.p, c.info.p, ap, c.s);
    if (solve)
      launch_solve_bk(c.list.p, (int)bk.size(), c.p_n.p, c.p_mem_off.p, c.p_fac_off.p, c.mem_pool.p, c.fac_pool.p,
                      c.piv_pool.p, c.dval.p, c.nf, c.coef_pool.p, ap, c.s);
  }
  if (solve) {
    for (int a = 0; a < 2; a++) {
      if (big[a].empty()) continue;
      const int m = a ? M_AUG64B : M_CHOL64B;
      int nmax = 1;
      for (int p : big[a]) nmax = std::max(nmax, (int)c.h_n[p]);
      CHK(upload_list(c, big[a], m));
      if (a)
        launch_solve_aug(c.g, c.ktype, false, c.list.p + (size_t)m * P, (int)big[a].size(), c.p_n.p, c.p_mem_off.p,
                         c.p_fac_off.p, c.mem_pool.p, c.dx.p, c.fac_pool.p, c.dval.p, c.nf, c.coef_pool.p,
                         c.refine_max_iter, c.refine_tol, c.pstatus.p, c.piters.p, c.solve_ws.p, nmax, ap, c.s);
      else
        launch_solve_chol(c.g, c.ktype, false, c.list.p + (size_t)m * P, (int)big[a].size(), c.p_n.p, c.p_mem_off.p,
                          c.p_fac_off.p, c.mem_pool.p, c.dx.p, c.fac_pool.p, c.dval.p, c.nf, c.coef_pool.p,
                          c.refine_max_iter, c.refine_tol, c.pstatus.p, c.piters.p, c.solve_ws.p, nmax, ae, c.s);
    }
  }
  HIPCHK(hipStreamSynchronize(c.s));
  return 0;
}

// Factor the new patches of each method; Cholesky breakdowns are refactored in
// FP64 (refactor_fp64, counted in "fallback").
static int factor_lists(Ctx& c, std::vector<int32_t>* fl) {
  const int P = c.g.P;
  const AugView ap = aug_view(c, true), ae = aug_view(c, false);
  // small-patch methods: patches ordered by size bucket, each bucket a launch of its own
  std::vector<int> bstart[N_METHODS];
  for (int m = M_CHOL64; m < N_METHODS; m++) {
    if (fl[m].empty() || method_big(m)) continue;
    auto nn = [&](int p) { return c.h_n[p] - (method_aug(m) ? c.h_m[p] : 0); };
    std::stable_sort(fl[m].begin(), fl[m].end(), [&](int a, int b) { return factor_bucket(nn(a)) < factor_bucket(nn(b)); });
    bstart[m].assign(N_FACTOR_BUCKETS + 1, (int)fl[m].size());
    for (int i = (int)fl[m].size() - 1; i >= 0; i--) bstart[m][factor_bucket(nn(fl[m][i]))] = i;
    for (int b = N_FACTOR_BUCKETS - 1; b >= 0; b--) bstart[m][b] = std::min(bstart[m][b], bstart[m][b + 1]);
  }
  for (int m = 0; m < N_METHODS; m++) if (!fl[m].empty()) CHK(upload_list(c, fl[m], m));
  if (!fl[M_BK64].empty())
    launch_factor_bk(c.g, c.ktype, c.list.p, (int)fl[M_BK64].size(), c.p_n.p, c.p_mem_off.p, c.p_fac_off.p, c.mem_pool.p,
                     c.dx.p, c.fac_pool.p, c.piv_pool.p, c.info.p, ap, c.s);
  // null-space data of the saddle-point patches before their factors (large
  // patches: the O(n^2) product over many workgroups, by row blocks)
  for (int m = M_AUG64; m <= M_AUG32B; m++) {
    if (fl[m].empty()) continue;
    if (method_big(m) && c.big_factor) {
      std::vector<int2> tasks;
      const int R = aug_resid_rows();
      for (int p : fl[m])
        for (int b = 0; (int64_t)b * R < c.h_n[p]; b++) tasks.push_back(make_int2(p, b));
      HIPCHK(hipStreamSynchronize(c.s));   // an earlier use of the task buffer
      HIPCHK(c.resid_tasks.reserve(tasks.size()));
      HIPCHK(hipMemcpyAsync(c.resid_tasks.p, tasks.data(), sizeof(int2) * tasks.size(), hipMemcpyHostToDevice, c.s));
      launch_aug_prep_tasks(c.g, c.ktype, c.list.p + (size_t)m * P, (int)fl[m].size(), c.p_n.p, c.p_mem_off.p,
                            c.mem_pool.p, c.dx.p, ap, c.resid_tasks.p, (int)tasks.size(), c.s);
      HIPCHK(hipStreamSynchronize(c.s));   // the host task list goes out of scope
    } else {
      launch_aug_prep(c.g, c.ktype, c.list.p + (size_t)m * P, (int)fl[m].size(), c.p_n.p, c.p_mem_off.p, c.mem_pool.p,
                      c.dx.p, ap, c.s);
    }
  }
  bool any_chol = false;
  for (int m = M_CHOL64; m < N_METHODS; m++)
    if (!fl[m].empty()) {
      any_chol = true;
      void* fac = method_fp32(m) ? (void*)c.fac32_pool.p : (void*)c.fac_pool.p;
      const AugView& a = method_aug(m) ? ap : ae;
      const int32_t* dl = c.list.p + (size_t)m * P;
      if (method_big(m)) {
        CHK(factor_big(c, m, fl[m], dl));
        continue;
      }
      for (int b = 0; b < N_FACTOR_BUCKETS; b++) {
        const int b0 = bstart[m][b], b1 = bstart[m][b + 1];
        if (b1 <= b0) continue;
        const int nb = factor_nb_for(c, method_fp32(m), b, dl + b0, std::min(b1 - b0, 256), fac, a);
        launch_factor_chol(c.g, c.ktype, method_fp32(m), dl + b0, b1 - b0, c.p_n.p, c.p_mem_off.p, c.p_fac_off.p,
                           c.mem_pool.p, c.dx.p, fac, c.pstatus.p, false, a, nb, FACTOR_BUCKET_N[b], c.s);
        for (int i = b0; i < b1; i++) c.h_fnb[fl[m][i]] = (uint8_t)nb;
      }
    }
  HIPCHK(hipMemcpyAsync(c.p_fnb.p, c.h_fnb.data(), P, hipMemcpyHostToDevice, c.s));
  if (!any_chol) return 0;
  std::vector<int32_t> st(P);
  HIPCHK(hipMemcpyAsync(st.data(), c.pstatus.p, sizeof(int32_t) * P, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  std::vector<int32_t> redo;
  for (int m = M_CHOL64; m < N_METHODS; m++)
    for (int p : fl[m]) if (st[p] != 0) redo.push_back(p);
  return refactor_fp64(c, redo, false);
}

// Task lists of the multi-workgroup B22 solves for the large saddle-point patches l
// (trsv_tiles.hip), uploaded to c.trsv_dtasks; c.aug_hstate sized for n_list list
// positions of the two-phase solve.
static int trsv_prepare(Ctx& c, const std::vector<int32_t>& l, size_t n_list, std::vector<int>& off) {
  std::vector<int> nn(l.size());
  for (size_t q = 0; q < l.size(); q++) nn[q] = c.h_n[l[q]] - c.h_m[l[q]];
  std::vector<int2> tasks;
  trsv_tasks(l, nn, tasks, off);
  HIPCHK(hipStreamSynchronize(c.s));   // an earlier use of the task buffer
  HIPCHK(c.trsv_dtasks.reserve(tasks.size() + 1));
  if (!tasks.empty())
    HIPCHK(hipMemcpyAsync(c.trsv_dtasks.p, tasks.data(), sizeof(int2) * tasks.size(), hipMemcpyHostToDevice, c.s));
  HIPCHK(c.aug_hstate.reserve(n_list * 2 * (size_t)POLY_MAX * 8 + 1));
  HIPCHK(hipStreamSynchronize(c.s));   // tasks is a host temporary
  return 0;
}

// One solve (or refinement step step_it >= 0) of the large saddle-point patches l of
// method m listed at dl: k_solve_aug up to the B22 right-hand side, the B22 solves
// over many workgroups for the patches run (those still refining), the rest.
static int solve_aug_big_tiled(Ctx& c, int m, const std::vector<int32_t>& l, const std::vector<int32_t>& run,
                               const int32_t* dl, int nmax, int step_it, double* dstate) {
  const AugView ap = aug_view(c, true);
  const bool f32 = method_fp32(m);
  const void* fac = f32 ? (const void*)c.fac32_pool.p : (const void*)c.fac_pool.p;
  std::vector<int> off;
  CHK(trsv_prepare(c, run, l.size(), off));
  for (int phase = 1; phase <= 2; phase++) {
    launch_solve_aug_step(c.g, c.ktype, f32, dl, (int)l.size(), c.p_n.p, c.p_mem_off.p, c.p_fac_off.p, c.mem_pool.p,
                          c.dx.p, fac, c.dval.p, c.nf, c.coef_pool.p, c.refine_max_iter, c.refine_tol, c.pstatus.p,
                          c.piters.p, c.solve_ws.p, nmax, ap, step_it, dstate, phase, c.aug_hstate.p, c.s);
    if (phase == 1)
      launch_trsv_tiles(f32, c.trsv_dtasks.p, off, c.p_n.p, ap, c.p_mem_off.p, c.p_fac_off.p, fac, c.solve_ws.p, c.nf,
                        c.s);
  }
  return 0;
}

// FP32 factors of large saddle-point patches: one refinement iteration per launch,
// the O(n^2) matrix-free residual spread over many workgroups (launch_aug_resid_big)
// instead of one workgroup per patch, which took most of the set-up of patches with
// thousands of members.  The same operations in the same order as the one-launch
// solve, so the same coefficients.
static int solve_aug_stepped(Ctx& c, const std::vector<int32_t>& l, int nmax) {
  const int P = c.g.P;
  const AugView ap = aug_view(c, true);
  const int32_t* dl = c.list.p + (size_t)M_AUG32B * P;
  HIPCHK(c.refine_dstate.reserve(P));
  std::vector<int2> tasks;
  const int R = aug_resid_rows();
  for (int p : l)
    for (int b = 0; (int64_t)b * R < c.h_n[p]; b++) tasks.push_back(make_int2(p, b));
  HIPCHK(c.resid_tasks.reserve(tasks.size()));
  HIPCHK(hipMemcpyAsync(c.resid_tasks.p, tasks.data(), sizeof(int2) * tasks.size(), hipMemcpyHostToDevice, c.s));
  std::vector<int32_t> st(P);
  std::vector<int32_t> run(l);   // the patches still refining
  for (int it = 0; it <= c.refine_max_iter; it++) {
    if (it > 0)
      launch_aug_resid_big(c.g, c.ktype, c.resid_tasks.p, (int)tasks.size(), c.p_n.p, c.p_mem_off.p, c.mem_pool.p,
                           c.dval.p, c.nf, c.pstatus.p, c.solve_ws.p, ap, c.s);
    if (c.tiled_solve)
      CHK(solve_aug_big_tiled(c, M_AUG32B, l, run, dl, nmax, it, c.refine_dstate.p));
    else
      launch_solve_aug_step(c.g, c.ktype, true, dl, (int)l.size(), c.p_n.p, c.p_mem_off.p, c.p_fac_off.p, c.mem_pool.p,
                            c.dx.p, c.fac32_pool.p, c.dval.p, c.nf, c.coef_pool.p, c.refine_max_iter, c.refine_tol,
                            c.pstatus.p, c.piters.p, c.solve_ws.p, nmax, ap, it, c.refine_dstate.p, 0, nullptr, c.s);
    HIPCHK(hipMemcpyAsync(st.data(), c.pstatus.p, sizeof(int32_t) * P, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    run.clear();
    for (int p : l)
      if (st[p] == 2) run.push_back(p);
    if (run.empty()) break;
  }
  return 0;
}

// Solve the coefficients of the listed patches with each patch's factor; an FP32
// factor whose refinement does not reach the tolerance is replaced by an FP64
// Bunch-Kaufman factor and solved again (counted in "fallback").
static int solve_list(Ctx& c, const std::vector<int32_t>& sl) {
  const int P = c.g.P;
  const AugView ap = aug_view(c, true), ae = aug_view(c, false);
  std::vector<int32_t> l[N_METHODS];
  for (int p : sl) l[c.h_method[p]].push_back(p);
  // one-workgroup solves size their LDS rows by the launch's largest patch: the
  // patches go by size, in bands launched apart, so that the many small patches
  // are not held to the occupancy of the few large ones.  Patches with a
  // matrix-core factor (inverse diagonal blocks) take the lean solves.
  static constexpr int SOLVE_BANDS[] = {96, 128, 192, 256, 384, 1 << 30};
  struct Band { int b0, b1; bool lean; };
  std::vector<Band> bands[N_METHODS];
  for (int m = M_CHOL64; m < N_METHODS; m++) {
    if (l[m].empty()) continue;
    if (method_big(m)) { bands[m].push_back({0, (int)l[m].size(), false}); continue; }
    auto lean = [&](int p) {
      if (!c.lean_solve || c.h_fnb[p] != 16) return false;
      if (method_aug(m))
        return solve_aug_lean_supported(c.ktype, method_fp32(m)) && c.h_n[p] <= solve_aug_lean_max_n();
      return c.h_n[p] <= solve_llt_max_n();
    };
    std::sort(l[m].begin(), l[m].end(), [&](int a, int b) {
      return std::make_pair(!lean(a), c.h_n[a]) < std::make_pair(!lean(b), c.h_n[b]);
    });
    size_t i = 0;
    while (i < l[m].size()) {
      const bool ln = lean(l[m][i]);
      for (int lim : SOLVE_BANDS) {
        const size_t b0 = i;
        while (i < l[m].size() && lean(l[m][i]) == ln && c.h_n[l[m][i]] <= lim) i++;
        if (i > b0) bands[m].push_back({(int)b0, (int)i, ln});
      }
    }
  }
  for (int m = 0; m < N_METHODS; m++) if (!l[m].empty()) CHK(upload_list(c, l[m], m));
  if (!l[M_BK64].empty())
    launch_solve_bk(c.list.p, (int)l[M_BK64].size(), c.p_n.p, c.p_mem_off.p, c.p_fac_off.p, c.mem_pool.p, c.fac_pool.p,
                    c.piv_pool.p, c.dval.p, c.nf, c.coef_pool.p, ap, c.s);
  bool any_chol = false;
  for (int m = M_CHOL64; m < N_METHODS; m++) any_chol = any_chol || !l[m].empty();
  if (c.nf > 8 && any_chol)
    return set_err(E_SUP, "PartitionSolve", "Cholesky solve handles up to 8 interpolated fields (found %d)", c.nf);
  size_t ws_need = 0;
  if (!l[M_CHOL64B].empty() || !l[M_CHOL32B].empty()) ws_need = (c.mem_used + 1) * 2 * (size_t)std::max(c.nf, 1);
  if (!l[M_AUG64B].empty() || !l[M_AUG32B].empty())
    ws_need = std::max(ws_need, (c.mem_used + 1) * aug_solve_ws_per_member(std::max(c.nf, 1)));
  if (ws_need) HIPCHK(c.solve_ws.reserve(ws_need));
  for (int m = M_CHOL64; m < N_METHODS; m++)
    if (!l[m].empty()) {
      const void* fac = method_fp32(m) ? (const void*)c.fac32_pool.p : (const void*)c.fac_pool.p;
      // one launch per band (large-patch methods: one launch)
      for (const Band& bd : bands[m]) {
        const int b0 = bd.b0, b1 = bd.b1;
        int nmax = 1;
        for (int i = b0; i < b1; i++) nmax = std::max(nmax, (int)c.h_n[l[m][i]]);
        const int32_t* dl = c.list.p + (size_t)m * P + b0;
        if (bd.lean && method_aug(m))
          launch_solve_aug_lean(c.g, c.ktype, method_fp32(m), dl, b1 - b0, c.p_n.p, c.p_mem_off.p, c.p_fac_off.p,
                                c.mem_pool.p, fac, c.dval.p, c.nf, c.coef_pool.p, c.refine_max_iter, c.refine_tol,
                                c.pstatus.p, c.piters.p, nmax, ap, c.s);
        else if (bd.lean)
          launch_solve_llt(c.g, c.ktype, method_fp32(m), dl, b1 - b0, c.p_n.p, c.p_mem_off.p, c.p_fac_off.p,
                           c.mem_pool.p, c.dx.p, fac, c.dval.p, c.nf, c.coef_pool.p, c.refine_max_iter, c.refine_tol,
                           c.pstatus.p, c.piters.p, nmax, ae.eps, c.s);
        else if (m == M_AUG32B && c.big_factor && c.stepped_refine)
          CHK(solve_aug_stepped(c, l[m], nmax));
        else if (m == M_AUG64B && c.big_factor && c.tiled_solve)
          CHK(solve_aug_big_tiled(c, m, l[m], l[m], dl, nmax, -1, nullptr));
        else if (method_aug(m))
          launch_solve_aug(c.g, c.ktype, method_fp32(m), dl, b1 - b0, c.p_n.p, c.p_mem_off.p, c.p_fac_off.p,
                           c.mem_pool.p, c.dx.p, fac, c.dval.p, c.nf, c.coef_pool.p, c.refine_max_iter, c.refine_tol,
                           c.pstatus.p, c.piters.p, method_big(m) ? c.solve_ws.p : nullptr, nmax, ap, c.s);
        else
          launch_solve_chol(c.g, c.ktype, method_fp32(m), dl, b1 - b0, c.p_n.p, c.p_mem_off.p, c.p_fac_off.p,
                            c.mem_pool.p, c.dx.p, fac, c.dval.p, c.nf, c.coef_pool.p, c.refine_max_iter, c.refine_tol,
                            c.pstatus.p, c.piters.p, method_big(m) ? c.solve_ws.p : nullptr, nmax, ae, c.s);
      }
    }
  if (l[M_CHOL32].empty() && l[M_CHOL32B].empty() && l[M_AUG32].empty() && l[M_AUG32B].empty()) return 0;
  std::vector<int32_t> st(P), its(P);
  HIPCHK(hipMemcpyAsync(st.data(), c.pstatus.p, sizeof(int32_t) * P, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipMemcpyAsync(its.data(), c.piters.p, sizeof(int32_t) * P, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  std::vector<int32_t> redo;
  std::vector<int32_t> refined(l[M_CHOL32]);
  for (int m : {M_CHOL32B, M_AUG32, M_AUG32B}) refined.insert(refined.end(), l[m].begin(), l[m].end());
  for (int p : refined) {
    c.refine_iters_max = std::max(c.refine_iters_max, (double)its[p]);
    c.refine_iters_sum += its[p];
    c.refined_patches += 1;
    if (its[p] >= 0 && its[p] < (int)c.refine_hist.size()) c.refine_hist[its[p]] += 1;
    if (st[p] != 0) redo.push_back(p);
  }
  return refactor_fp64(c, redo, true);
}

// LagRBF_PartitionsSetUp (partition.c:417-478): touch the corner patches of every
// particle's cell, set up (gather + assemble + factor) the new ones, solve the
// unsolved ones.  Batched: one launch per phase over the patch list.
int partitions_setup(Ctx& c, bool all_patches) {
  const char* fn = "LagRBF_PartitionsSetUp";
  const int P = c.g.P;
  if (c.factors_released) return set_err(E_STATE, fn, "patch factors were released");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c.s));
  HIPCHK(hipMemsetAsync(c.touched.p, all_patches ? 1 : 0, P, c.s));
  if (!all_patches) launch_touch(c.g, c.pcell.p, c.alive.p, c.Np, c.touched.p, c.s);
  std::vector<uint8_t> t(P);
  HIPCHK(hipMemcpyAsync(t.data(), c.touched.p, P, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  std::vector<int32_t> newl, solvel;
  std::vector<int32_t> fl[N_METHODS];   // new patches per factor method
  size_t need_mem = c.mem_used;
  for (int p = 0; p < P; p++) {
    if (!t[p]) continue;
    if (!c.h_setup[p]) {
      const int n = c.data_filled ? c.h_pcount[p] : 0;
      c.h_n[p] = n;
      c.h_mem_off[p] = (int64_t)need_mem;
      // with a polynomial block a patch's slots in the member-indexed pools are
      // padded by POLY_MAX: its Bunch-Kaufman pivots cover n + m rows
      need_mem += n + (c.poly_deg >= 0 ? POLY_MAX : 0);
      newl.push_back(p);
      c.max_patch = std::max(c.max_patch, n);
      // extensions: polynomial terms and shape parameter of the patch
      c.h_deg[p] = patch_degree(c.g.dim, c.poly_deg, n);
      c.h_m[p] = poly_terms(c.g.dim, c.h_deg[p]);
      c.h_eps[p] = patch_eps(c, p, n);
      const int mp = c.h_m[p];
      int m = M_BK64;
      const bool chol = chol_kernel_supported(c.ktype) || (c.ktype == K_PHS3 && mp > 0);
      if (c.factor_mode != 1 && chol) {
        // the mixed path factors in FP32 and refines in FP64 unless told to factor
        // in FP64 (the FP32 evaluation then reads the FP64 solution all the same)
        const bool f32 = c.precision && !c.mixed_factor64;
        if (mp > 0) {
          if (n - mp <= c.chol_big_min) m = f32 ? M_AUG32 : M_AUG64;
          else m = f32 ? M_AUG32B : M_AUG64B;
          c.h_aux_off[p] = (int64_t)c.aux_used;
          c.aux_used += aug_aux_size(n, mp);
        } else if (n <= c.chol_big_min) {
          m = f32 ? M_CHOL32 : M_CHOL64;
        } else {
          m = f32 ? M_CHOL32B : M_CHOL64B;
        }
      }
      c.h_method[p] = (uint8_t)m;
      if (n > 0) fl[m].push_back(p);
    }
  }
  if (!newl.empty()) {
    const auto ta = std::chrono::steady_clock::now();
    HIPCHK(c.mem_pool.grow(need_mem + 1, c.mem_used, c.s));
    HIPCHK(c.piv_pool.grow(need_mem + 1, c.mem_used, c.s));
    if (c.aux_used > 0) {
      HIPCHK(c.aux_pool.grow(c.aux_used + 1, c.aux_pool.cap, c.s));
    }
    HIPCHK(hipMemcpyAsync(c.p_m.p, c.h_m.data(), sizeof(int32_t) * P, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.p_deg.p, c.h_deg.data(), sizeof(int32_t) * P, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.p_eps.p, c.h_eps.data(), sizeof(double) * P, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.p_aux_off.p, c.h_aux_off.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, c.s));
    HIPCHK(c.coef_pool.grow((need_mem + 1) * std::max(c.nf, 1), c.mem_used * std::max(c.nf, 1), c.s));
    if (c.precision) HIPCHK(c.coef32_pool.grow((need_mem + 1) * std::max(c.nf, 1), c.mem_used * std::max(c.nf, 1), c.s));
    c.mem_used = need_mem;
    for (int m = 0; m < N_METHODS; m++) CHK(alloc_factor_slots(c, fl[m], m));
    c.setup_alloc_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
    HIPCHK(hipMemcpyAsync(c.p_n.p, c.h_n.data(), sizeof(int32_t) * P, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.p_mem_off.p, c.h_mem_off.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.p_fac_off.p, c.h_fac_off.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.list.p, newl.data(), sizeof(int32_t) * newl.size(), hipMemcpyHostToDevice, c.s));
    launch_patch_gather(c.g, c.list.p, (int)newl.size(), c.cell_start.p, c.cell_nodes.p, c.p_mem_off.p, c.mem_pool.p, c.s);
    hipEvent_t f0, f1;
    HIPCHK(hipEventCreate(&f0)); HIPCHK(hipEventCreate(&f1));
    HIPCHK(hipEventRecord(f0, c.s));
    CHK(factor_lists(c, fl));
    HIPCHK(hipEventRecord(f1, c.s));
    HIPCHK(hipEventSynchronize(f1));
    float fms = 0.f;
    HIPCHK(hipEventElapsedTime(&fms, f0, f1));
    c.setup_factor_ms += fms;
    HIPCHK(hipEventDestroy(f0)); HIPCHK(hipEventDestroy(f1));
    for (int p : newl) c.h_setup[p] = 1;
    if (c.h_cell_start_prev.empty() && c.data_filled) {
      // first patches of this generation: remember the cell lists they were built from
      c.h_cell_start_prev.resize(c.g.C + 1);
      c.h_cell_nodes_prev.resize(c.N);
      HIPCHK(hipMemcpyAsync(c.h_cell_start_prev.data(), c.cell_start.p, sizeof(int32_t) * (c.g.C + 1), hipMemcpyDeviceToHost, c.s));
      if (c.N > 0) HIPCHK(hipMemcpyAsync(c.h_cell_nodes_prev.data(), c.cell_nodes.p, sizeof(int32_t) * c.N, hipMemcpyDeviceToHost, c.s));
    }
    c.patches_setup += (double)newl.size();
    c.setup_N = c.N;
    HIPCHK(hipStreamSynchronize(c.s));   // list buffer is reused below
  }
  for (int p = 0; p < P; p++)
    if (t[p] && c.h_setup[p] && !c.h_solved[p] && c.h_n[p] > 0) solvel.push_back(p);
  if (!solvel.empty()) {
    hipEvent_t s0, s1;
    HIPCHK(hipEventCreate(&s0)); HIPCHK(hipEventCreate(&s1));
    HIPCHK(hipEventRecord(s0, c.s));
    CHK(solve_list(c, solvel));
    HIPCHK(hipEventRecord(s1, c.s));
    HIPCHK(hipEventSynchronize(s1));
    float sms = 0.f;
    HIPCHK(hipEventElapsedTime(&sms, s0, s1));
    c.setup_solve_ms += sms;
    HIPCHK(hipEventDestroy(s0)); HIPCHK(hipEventDestroy(s1));
    if (c.precision) {
      // FP32 copy of the refined coefficients for the FP32 evaluation
      CHK(upload_list(c, solvel, 0));
      launch_coef_to_f32(c.list.p, (int)solvel.size(), c.p_n.p, c.p_mem_off.p, c.nf, c.coef_pool.p, c.coef32_pool.p, c.s);
      HIPCHK(hipStreamSynchronize(c.s));
    }
  }
  for (int p : solvel) c.h_solved[p] = 1;
  for (int p = 0; p < P; p++) if (t[p] && c.h_setup[p] && c.h_n[p] == 0) c.h_solved[p] = 1;
  bool changed = false;
  for (int p = 0; p < P; p++) {
    uint8_t a = (c.h_setup[p] && c.h_n[p] > 0) ? 1 : 0;
    if (a != c.h_active[p]) { c.h_active[p] = a; changed = true; }
  }
  if (changed) HIPCHK(hipMemcpyAsync(c.p_active.p, c.h_active.data(), P, hipMemcpyHostToDevice, c.s));
  if (changed || !solvel.empty()) {
    c.blk_valid = false;
    c.table_gen++;
  }
  HIPCHK(hipEventRecord(e1, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  c.setup_ms += ms;
  HIPCHK(hipEventDestroy(e0)); HIPCHK(hipEventDestroy(e1));
  (void)fn;
  return 0;
}

void release_factors(Ctx& c) {
  (void)hipStreamSynchronize(c.s);
  c.fac_pool.release(); c.fac32_pool.release(); c.piv_pool.release(); c.aux_pool.release(); c.solve_ws.release();
  c.factors_released = true;
}

// LagRBF_PartitionsReset, partition.c:494-524
static int partitions_reset(Ctx& c) {
  const int P = c.g.P;
  c.blk_valid = false; c.table_gen++;
  if (c.solve_method == 2) {
    std::fill(c.h_setup.begin(), c.h_setup.end(), 0);
    std::fill(c.h_solved.begin(), c.h_solved.end(), 0);
    std::fill(c.h_n.begin(), c.h_n.end(), 0);
    std::fill(c.h_active.begin(), c.h_active.end(), 0);
    c.mem_used = 0; c.fac_used = 0; c.fac32_used = 0; c.aux_used = 0; c.setup_N = -1;
    c.h_cell_start_prev.clear(); c.h_cell_nodes_prev.clear();
    c.memb_current = true;
    HIPCHK(hipMemsetAsync(c.p_active.p, 0, P, c.s));
    HIPCHK(hipMemsetAsync(c.p_n.p, 0, sizeof(int32_t) * P, c.s));
  } else {
    std::fill(c.h_solved.begin(), c.h_solved.end(), 0);
  }
  return 0;
}

// r^3 tables carry the coefficients times eps_r^3: the evaluation then forms phi as
// r^2 sqrt(r^2) (two multiplications fewer per node than (eps r)^3).
static double table_coef_scale(const Ctx& c) { return c.ktype == K_PHS3 ? c.g.eps_r * c.g.eps_r * c.g.eps_r : 1.0; }

// RBFInterpolation(ctx, swi, swd, name), interpolation.c:178-272
// Cell-major node table for the evaluation, rebuilt after the data or any patch changed.
int ensure_blk(Ctx& c, bool f32) {
  if (c.blk_valid && c.blk_f32 == f32) return 0;
  const size_t n = ((size_t)c.N + BLK_PAD_ROWS) * blk_row(c.g.dim, c.nf);   // N rows + padding rows
  if (f32) HIPCHK(c.blk32.reserve(n));
  else HIPCHK(c.blk64.reserve(n));
  if (c.N > 0)
    launch_build_blk(c.g, f32, c.dx.p, c.N, c.cell_nodes.p, c.dcell.p, c.cell_start.p, c.p_active.p, c.p_mem_off.p,
                     c.coef_pool.p, c.nf, c.nf, 0, table_coef_scale(c), f32 ? (void*)c.blk32.p : (void*)c.blk64.p, c.s);
  if (c.poly_deg >= 0) {
    // polynomial part of every active patch, zero for the others
    HIPCHK(c.ptab.reserve((size_t)c.g.P * c.poly_mm * std::max(c.nf, 1)));
    launch_build_ptab(c.g, c.p_active.p, aug_view(c, true), c.poly_mm, c.nf, c.nf, 0, c.ptab.p, c.s);
  }
  c.blk_valid = true;
  c.blk_f32 = f32;
  return 0;
}

// Evaluation timing without a host wait per launch: event pairs are recorded
// around each evaluation and resolved at the next synchronisation point.
static int event_take(Ctx& c, hipEvent_t& e) {
  if (!c.ev_pool.empty()) {
    e = c.ev_pool.back();
    c.ev_pool.pop_back();
    return 0;
  }
  HIPCHK(hipEventCreate(&e));
  return 0;
}

static int flush_eval_timing(Ctx& c) {
  for (auto& pr : c.ev_pending) {
    HIPCHK(hipEventSynchronize(pr.second));
    float ms = 0.f;
    // a pair that no dispatch recorded (nothing was launched) is not counted
    if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      c.eval_ms += ms;
      c.eval_count += 1;
    } else {
      (void)hipGetLastError();
    }
    c.ev_pool.push_back(pr.first);
    c.ev_pool.push_back(pr.second);
  }
  c.ev_pending.clear();
  return 0;
}

static void release_events(Ctx& c) {
  (void)flush_eval_timing(c);
  for (hipEvent_t e : c.ev_pool) (void)hipEventDestroy(e);
  c.ev_pool.clear();
}

// Table of field group g (fields g*dim .. g*dim+dim-1): the two snapshots of a
// time-blended velocity, each evaluated with the dim-field kernels.
static int ensure_blk_group(Ctx& c, bool f32, int g) {
  if (c.blkg_gen[g] == c.table_gen && c.blkg_f32[g] == f32) return 0;
  const int dim = c.g.dim;
  const size_t n = ((size_t)c.N + BLK_PAD_ROWS) * blk_row(dim, dim);
  if (f32) HIPCHK(c.blkg32[g].reserve(n));
  else HIPCHK(c.blkg64[g].reserve(n));
  if (c.N > 0)
    launch_build_blk(c.g, f32, c.dx.p, c.N, c.cell_nodes.p, c.dcell.p, c.cell_start.p, c.p_active.p, c.p_mem_off.p,
                     c.coef_pool.p, dim, c.nf, g * dim, table_coef_scale(c), f32 ? (void*)c.blkg32[g].p : (void*)c.blkg64[g].p,
                     c.s);
  if (c.poly_deg >= 0) {
    HIPCHK(c.ptabg[g].reserve((size_t)c.g.P * c.poly_mm * dim));
    launch_build_ptab(c.g, c.p_active.p, aug_view(c, true), c.poly_mm, dim, c.nf, g * dim, c.ptabg[g].p, c.s);
  }
  c.blkg_gen[g] = c.table_gen;
  c.blkg_f32[g] = f32;
  return 0;
}

// group < 0: fields 0 .. nfields-1 of the n_interpolation_fields table; group
// 0 / 1: the dim fields of that snapshot (cell-table evaluation only).
// Density bucket of the autotuner: particles per cell of the swarm (the whole
// swarm's while the adaptive loop evaluates its unfinished prefix).
static int eval_bucket(const Ctx& c) { return density_bucket(c.bucket_np >= 0 ? c.bucket_np : c.Np, c.g.C); }

static int rbf_interpolation(Ctx& c, double* out, int nfields, int group = -1) {
  if (nfields <= 0 || c.Np == 0) return 0;
  PatchView pv{c.p_n.p, c.p_active.p, c.p_mem_off.p, c.mem_pool.p, c.coef_pool.p, c.nf};
  if (!c.mem_pool.p) {
    // no patch was ever set up: every patch is inactive, nothing to evaluate but the
    // kernel still needs valid (unused) pool pointers
    HIPCHK(c.mem_pool.reserve(1)); HIPCHK(c.coef_pool.reserve(1));
    pv.members = c.mem_pool.p; pv.coef = c.coef_pool.p;
  }
  const bool cell_order = c.eval_mode == 1 || (c.eval_mode == 0 && c.memb_current);
  if (c.eval_mode == 1 && !c.memb_current)
    return set_err(E_STATE, "RBFInterpolation", "-lagrbf_eval cell needs patches built from the current data");
  if (cell_order && !c.sorted_valid) {
    launch_sort_particles(c.pcell.p, c.alive.p, c.Np, c.g.C, c.pstart.p, c.perm.p, c.pscratch.p, c.s);
    c.sorted_valid = true;
  }
  if (!c.zeros.p) { HIPCHK(c.zeros.reserve(1)); HIPCHK(hipMemset(c.zeros.p, 0, sizeof(double))); }
  const bool blk32 = eval_f32(c);
  // the per-cell fallback kernels know neither the r^3 kernel nor the extensions:
  // those go through the reference-order kernel
  const bool ext = c.poly_deg >= 0 || c.local_eps > 0.0 || c.ktype == K_PHS3;
  const bool use_blk = cell_order && eval_blk_supported(group >= 0 ? c.g.dim : c.nf, c.g.dim, blk32);
  if (group >= 0 && !use_blk)
    return set_err(E_SUP, "RBFInterpolation", "snapshot fields need the cell-table evaluation (patches from the current data)");
  if (use_blk) CHK(group >= 0 ? ensure_blk_group(c, blk32, group) : ensure_blk(c, blk32));
  const int nft = group >= 0 ? c.g.dim : c.nf;
  const void* table = group >= 0 ? (blk32 ? (const void*)c.blkg32[group].p : (const void*)c.blkg64[group].p)
                                 : (blk32 ? (const void*)c.blk32.p : (const void*)c.blk64.p);
  // timing: the cell-table evaluation dispatch records its own start and end
  // (once its launch shape is chosen); the other paths are bracketed by stream
  // events and resolved at once
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed_dispatch = use_blk && c.tuner.planned(eval_bucket(c)) >= 0;
  if (!c.capturing && (timed_dispatch || !use_blk)) {
    CHK(event_take(c, e0));
    CHK(event_take(c, e1));
    if (!use_blk) HIPCHK(hipEventRecord(e0, c.s));
  }
  if (use_blk) {
    const int per_cell = (c.Np + c.g.C - 1) / std::max(c.g.C, 1);
    const int block = std::min(256, std::max(64, (per_cell + 63) / 64 * 64));
    const int bucket = eval_bucket(c);
    const int planned = c.tuner.planned(bucket);
    if (planned < 0 || eval_variant_listed(planned))
      launch_nonempty_cells(c.pstart.p, c.g.C, c.cell_list.p, c.list_count.p, c.s);
    auto launch = [&](int v) {
      bool listed;
      int lanes, pipe, waves;
      eval_variant(v, listed, lanes, pipe, waves);
      launch_eval_blk(c.g, eval_ktype(c), blk32, c.px.p, c.ld, c.perm.p, c.pstart.p, c.w.p, table, nft, c.cell_start.p,
                      nfields, out, block, c.Np, listed ? c.cell_list.p : nullptr, listed ? c.list_count.p : nullptr,
                      lanes, pipe, waves, eval_ext(c, group), c.s);
    };
    set_eval_timing_events(e0, e1);
    c.last_variant =
        c.tuner.run(bucket, nft == c.g.dim && nfields == nft, blk32 && nft == c.g.dim, c.cell_ordered || !c.reorder,
                    !c.use_graph, c.s, launch);
    set_eval_timing_events(nullptr, nullptr);
  } else if (cell_order && !ext) {
    const int per_cell = (c.Np + c.g.C - 1) / std::max(c.g.C, 1);
    const int block = std::min(256, std::max(64, (per_cell + 63) / 64 * 64));
    const bool f32 = c.precision && eval_f32_supported(c.ktype) && c.xs32.p && c.coef32_pool.p;
    if (f32)
      launch_eval_cell(c.g, c.ktype, true, c.px.p, c.ld, c.perm.p, c.pstart.p, c.w.p, c.p_active.p, c.p_mem_off.p,
                       c.coef32_pool.p, c.nf, c.zeros32.p, c.xs32.p, c.N, c.cell_start.p, nfields, out, block, c.s);
    else
      launch_eval_cell(c.g, c.ktype, false, c.px.p, c.ld, c.perm.p, c.pstart.p, c.w.p, c.p_active.p, c.p_mem_off.p,
                       c.coef_pool.p, c.nf, c.zeros.p, c.xs_sorted.p ? c.xs_sorted.p : c.pux.p, c.N, c.cell_start.p,
                       nfields, out, block, c.s);
  } else {
    launch_eval_ref(c.g, c.ktype, c.px.p, c.ld, c.Np, c.pcell.p, c.alive.p, c.w.p, pv, c.dx.p ? c.dx.p : c.pux.p,
                    c.cell_start.p, nfields, out, aug_view(c, true), c.s);
  }
  if (e0) {
    if (!use_blk) HIPCHK(hipEventRecord(e1, c.s));
    c.ev_pending.emplace_back(e0, e1);
    if (!use_blk || c.ev_pending.size() >= 256) CHK(flush_eval_timing(c));
  }
  return 0;
}

static int interpolation(Ctx& c) {
  const char* fn = "Interface_Interpolation";
  if (!c.points_filled) return set_err(E_NULL, fn, "Interpolation swarm (ctx->swi) is empty: call lagrbf_fill_swarm first");
  CHK(partitions_setup(c, c.setup_all));
  CHK(rbf_interpolation(c, c.fields.p, c.nf));
  return 0;
}

// ---------------------------------------------------------------------------
// Advection
static int eval_callback(Ctx& c) {
  const char* fn = "AdvectionEvaluateVelocityFunction";
  const int dim = c.g.dim, Np = c.Np;
  if (!c.vfn && !c.vfn_batch) return set_err(E_NULL, fn, "AdvectionCtx->velocity_fn is NULL, please provide a valid function");
  std::vector<double> xs((size_t)dim * c.ld), vs((size_t)dim * c.ld, 0.0);
  std::vector<int32_t> al(Np);
  if (Np == 0) return 0;
  HIPCHK(hipMemcpyAsync(xs.data(), c.px.p, sizeof(double) * dim * c.ld, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipMemcpyAsync(vs.data(), c.vel.p, sizeof(double) * dim * c.ld, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipMemcpyAsync(al.data(), c.alive.p, sizeof(int32_t) * Np, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  std::vector<double> X, V;
  std::vector<int> idx;
  for (int i = 0; i < Np; i++) if (al[i]) idx.push_back(i);
  const int m = (int)idx.size();
  X.resize((size_t)m * dim + 1); V.resize((size_t)m * dim + 1, 0.0);
  for (int q = 0; q < m; q++) for (int d = 0; d < dim; d++) X[(size_t)q * dim + d] = xs[(size_t)d * c.ld + idx[q]];
  int err = 0;
  if (c.vfn_batch) err = c.vfn_batch(m, dim, X.data(), V.data(), c.vuser);
  else for (int q = 0; q < m && !err; q++) err = c.vfn(&X[(size_t)q * dim], &V[(size_t)q * dim], c.vuser);
  if (err) return set_err(err, fn, "velocity callback returned %d", err);
  for (int q = 0; q < m; q++) for (int d = 0; d < dim; d++) vs[(size_t)d * c.ld + idx[q]] = V[(size_t)q * dim + d];
  HIPCHK(hipMemcpyAsync(c.vel.p, vs.data(), sizeof(double) * dim * c.ld, hipMemcpyHostToDevice, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  return 0;
}

// AdvectionEvaluateVelocity, advection.c:391-414
static int eval_velocity(Ctx& c) {
  if (c.analytical) return eval_callback(c);
  return rbf_interpolation(c, c.vel.p, c.g.dim);
}

// Velocity at the stage point into an RK stage array (v_adv_s): interpolated
// straight into it, or through the velocity field for a callback.
static int eval_velocity_into(Ctx& c, double* dst) {
  if (!c.analytical) return rbf_interpolation(c, dst, c.g.dim);
  CHK(eval_callback(c));
  HIPCHK(hipMemcpyAsync(dst, c.vel.p, sizeof(double) * c.g.dim * c.ld, hipMemcpyDeviceToDevice, c.s));
  return 0;
}

// AdvectionInitializeTimestepVelocity, advection.c:353-373
static int init_step_velocity(Ctx& c) {
  if (c.analytical) return eval_callback(c);
  launch_copy_rows(c.fields.p, c.ld, c.vel.p, c.ld, c.g.dim, c.Np, c.alive.p, c.s);
  return 0;
}

static int cfl_dt(Ctx& c, double& dt) {
  unsigned long long init = (unsigned long long)0x7fefffffffffffffull;   // DBL_MAX bits
  HIPCHK(hipMemcpyAsync(c.counters.p + 1, &init, sizeof(init), hipMemcpyHostToDevice, c.s));
  launch_cfl(c.vel.p, c.ld, c.Np, c.g.dim, c.alive.p, c.cfl_fac * (1.0 / c.g.eps_w), c.counters.p + 1, c.s);
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, c.counters.p + 1, sizeof(bits), hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  std::memcpy(&dt, &bits, sizeof(dt));
  return 0;
}

static StageArgs stage_args(int nv, const double* const* v, const double* a, double scale, int project, int proj_in_loop,
                            int write_vel, const Ctx& c) {
  StageArgs s{};
  s.nv = nv;
  for (int j = 0; j < nv; j++) { s.v[j] = v[j]; s.a[j] = a[j]; }
  s.scale = scale; s.project = project; s.proj_in_loop = proj_in_loop; s.write_vel = write_vel;
  s.rmin = c.rmin; s.rmax = c.rmax;
  return s;
}

static void stage(Ctx& c, const double* x0, const StageArgs& a) {
  c.sorted_valid = false;
  launch_stage(c.g, c.pux.p, c.puy.p, c.puz.p, c.px.p, x0, c.vel.p, c.ld, c.Np, c.pcell.p, c.alive.p, c.counters.p,
               c.w.p, a, c.s);
}


// The reordered swarm becomes the current one (the other set is the next target).
static void swap_particle_buffers(Ctx& c) {
  swap_buf(c.px, c.px_alt); swap_buf(c.vel, c.vel_alt); swap_buf(c.fields, c.fields_alt); swap_buf(c.w, c.w_alt);
  swap_buf(c.pcell, c.pcell_alt); swap_buf(c.alive, c.alive_alt); swap_buf(c.porig, c.porig_alt);
  c.parity ^= 1;
}

static bool setup_complete(const Ctx& c);

// Store the swarm in the cell order of the last sort (alive particles grouped by
// cell, dead ones at the tail): the evaluation's particle reads and writes become
// contiguous.  Positions, the step's start velocity, interpolated fields, weights,
// cell ids, alive flags and the slot -> fill index map move together; nothing a
// particle computes depends on its slot.
static int reorder_particles(Ctx& c) {
  if (!c.reorder || !c.sorted_valid || c.Np == 0) return 0;
  const size_t ld = (size_t)c.ld;
  const int dim = c.g.dim, nf = std::max(c.nf, 1);
  HIPCHK(c.px_alt.reserve(ld * dim));
  HIPCHK(c.vel_alt.reserve(ld * dim));
  HIPCHK(c.fields_alt.reserve(ld * nf));
  HIPCHK(c.w_alt.reserve(ld * c.g.nen));
  HIPCHK(c.pcell_alt.reserve(ld));
  HIPCHK(c.alive_alt.reserve(ld));
  HIPCHK(c.porig_alt.reserve(ld));
  ReorderSet r{};
  r.src[0] = c.px.p; r.dst[0] = c.px_alt.p; r.rows[0] = dim;
  r.src[1] = c.vel.p; r.dst[1] = c.vel_alt.p; r.rows[1] = dim;
  r.src[2] = c.fields.p; r.dst[2] = c.fields_alt.p; r.rows[2] = c.nf;
  r.src[3] = c.w.p; r.dst[3] = c.w_alt.p; r.rows[3] = c.g.nen;
  r.nd = 4;
  r.isrc[0] = c.pcell.p; r.idst[0] = c.pcell_alt.p;
  r.isrc[1] = c.alive.p; r.idst[1] = c.alive_alt.p;
  r.isrc[2] = c.porig.p; r.idst[2] = c.porig_alt.p;
  r.ni = 3;
  HIPCHK(launch_reorder(r, c.perm.p, c.Np, c.ld, c.s));
  swap_particle_buffers(c);
  c.cell_ordered = true;
  // the cell groups now sit at pstart[c] .. pstart[c+1] in slot order
  launch_iota(c.perm.p, c.Np, c.s);
  return 0;
}

static int advect_swarm(Ctx& c, double dt) {
  const int dim = c.g.dim;
  CHK(reorder_particles(c));
  switch (c.scheme) {
    case 0: {  // runge_kutta_1.c:37-75
      const double* v[1] = {c.vel.p};
      const double a[1] = {1.0};
      stage(c, c.px.p, stage_args(1, v, a, dt, 0, 0, 0, c));
      return 0;
    }
    case 1: {  // runge_kutta_2.c:38-91
      launch_save(c.px.p, c.x0.p, c.vel.p, nullptr, c.ld, c.Np, dim, c.alive.p, c.s);
      const double* v[1] = {c.vel.p};
      const double a[1] = {1.0};
      stage(c, c.x0.p, stage_args(1, v, a, 0.5 * dt, 0, 0, 0, c));
      CHK(eval_velocity(c));
      stage(c, c.x0.p, stage_args(1, v, a, dt, 0, 0, 0, c));
      return 0;
    }
    case 2: {  // runge_kutta_4.c:45-156
      static const double cc[3] = {0.5, 0.5, 1.0};
      static const double b[4] = {1.0, 2.0, 2.0, 1.0};
      launch_save(c.px.p, c.x0.p, c.vel.p, c.vadv[0].p, c.ld, c.Np, dim, c.alive.p, c.s);
      for (int s = 0; s < 3; s++) {
        const double* v[1] = {c.vadv[s].p};
        const double a[1] = {1.0};
        stage(c, c.x0.p, stage_args(1, v, a, cc[s] * dt, 1, 0, 1, c));
        CHK(eval_velocity_into(c, c.vadv[s + 1].p));
      }
      const double* v[4] = {c.vadv[0].p, c.vadv[1].p, c.vadv[2].p, c.vadv[3].p};
      stage(c, c.x0.p, stage_args(4, v, b, dt / 6.0, 1, 1, 0, c));
      return 0;
    }
    default: {  // dormand_prince.c:216-310
      static const double A[6][6] = {
          {1.0 / 5.0, 0, 0, 0, 0, 0},
          {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
          {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
          {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
          {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
          {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
      const double dtd = c.adv_dt;
      launch_save(c.px.p, c.x0.p, c.vel.p, c.vadv[0].p, c.ld, c.Np, dim, c.alive.p, c.s);
      for (int s = 0; s < 6; s++) {
        const double* v[7];
        for (int j = 0; j <= s; j++) v[j] = c.vadv[j].p;
        stage(c, c.x0.p, stage_args(s + 1, v, A[s], dtd, 1, 0, 1, c));
        CHK(eval_velocity_into(c, c.vadv[s + 1].p));
      }
      const double* vv[7];
      for (int j = 0; j < 7; j++) vv[j] = c.vadv[j].p;
      unsigned long long zero[2] = {0ull, 0ull};
      HIPCHK(hipMemcpyAsync(c.counters.p + 1, zero, sizeof(zero), hipMemcpyHostToDevice, c.s));
      launch_dopri_err(c.px.p, c.x0.p, vv, c.ld, c.Np, dim, c.alive.p, dtd, c.dopri_atol, c.dopri_rtol, c.counters.p + 1,
                       reinterpret_cast<unsigned int*>(c.counters.p + 2), c.s);
      unsigned long long res[2];
      HIPCHK(hipMemcpyAsync(res, c.counters.p + 1, sizeof(res), hipMemcpyDeviceToHost, c.s));
      HIPCHK(hipStreamSynchronize(c.s));
      double max_err;
      if ((unsigned int)res[1] == 0) max_err = -DBL_MAX;
      else std::memcpy(&max_err, &res[0], sizeof(double));
      c.dopri_err = max_err;
      // AdvectionComputeAdaptiveTimeStep_DOPRI, dormand_prince.c:179-214 (PetscMin/PetscMax macros)
      const double fac_max = max_err <= 1.0 ? c.dopri_fac_max : 1.0;
      const double g = c.dopri_fac * std::pow(1.0 / max_err, 1.0 / 5.0);
      const double inner = (c.dopri_fac_min < g) ? g : c.dopri_fac_min;
      const double f = (fac_max < inner) ? fac_max : inner;
      c.adv_dt = dtd * f;
      return 0;
    }
  }
}

static int update_removed(Ctx& c) {
  unsigned long long r = 0;
  HIPCHK(hipMemcpyAsync(&r, c.counters.p, sizeof(r), hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  c.removed = (double)r;
  return 0;
}

// InternalForwardAdvectionLoop, time_loop.c:39-100
static int advect_interpolate(Ctx& c, double t, double t_next) {
  const char* fn = "InternalForwardAdvectionLoop";
  if (!c.advection) return set_err(E_NULL, fn, "Advection context is NULL (set -advection)");
  if (!c.points_filled) return set_err(E_NULL, fn, "Interpolation swarm (ctx->swi) is NULL");
  if (!c.analytical && c.nf < c.g.dim) return set_err(E_SIZ, fn, "n_interpolation_fields (%d) < dim (%d)", c.nf, c.g.dim);
  CHK(init_step_velocity(c));
  while (t < t_next) {
    double dt, dtc;
    CHK(cfl_dt(c, dtc));
    if (c.scheme == 3) dt = (c.adv_dt < dtc) ? c.adv_dt : dtc;
    else dt = dtc;
    if (t + dt > t_next) dt = t_next - t;
    c.adv_dt = dt;
    CHK(advect_swarm(c, dt));
    if (!c.analytical) {
      if (!setup_complete(c)) CHK(partitions_setup(c, c.setup_all));
      CHK(rbf_interpolation(c, c.fields.p, c.nf));
    }
    CHK(init_step_velocity(c));
    t += dt;
  }
  CHK(update_removed(c));
  if (!c.analytical) {
    CHK(partitions_reset(c));
    // CellInfoReset + DMSwarmSetLocalSizes(swarm_data,0,-1), time_loop.c:94-98
    c.N = 0;
    c.data_filled = false;
    c.blk_valid = false; c.table_gen++;
    HIPCHK(hipMemsetAsync(c.cell_start.p, 0, sizeof(int32_t) * (c.g.C + 1), c.s));
    HIPCHK(hipStreamSynchronize(c.s));
  }
  return 0;
}

// ForwardAnalyticalAdvection, time_loop.c:176-221
static int advect_analytical(Ctx& c, double t, double t_next) {
  const char* fn = "ForwardAnalyticalAdvection";
  if (!c.advection) return set_err(E_NULL, fn, "Advection context is NULL (set -advection)");
  if (!c.points_filled) return set_err(E_NULL, fn, "Interpolation swarm (ctx->swi) is NULL");
  if (!c.vfn && !c.vfn_batch) return set_err(E_NULL, fn, "AdvectionCtx->velocity_fn is NULL, please provide a valid function");
  c.analytical = true;
  CHK(init_step_velocity(c));
  while (t < t_next) {
    double dt, dtc;
    CHK(cfl_dt(c, dtc));
    if (c.scheme == 3) dt = (c.adv_dt < dtc) ? c.adv_dt : dtc;
    else dt = dtc;
    if (t + dt > t_next) dt = t_next - t;
    c.adv_dt = dt;
    CHK(advect_swarm(c, dt));
    CHK(init_step_velocity(c));
    t += dt;
  }
  return update_removed(c);
}

// Every patch set up and solved: a lazy set-up pass can change nothing (whatever
// cells the particles touch), so the per-step pass and its host wait are skipped.
static bool setup_complete(const Ctx& c) {
  if (c.factors_released) return true;
  for (int p = 0; p < c.g.P; p++)
    if (!c.h_setup[p] || !c.h_solved[p]) return false;
  return true;
}

static void destroy_graphs(Ctx& c) {
  for (auto& g : c.gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// One fixed-dt step: RK stages with the velocity interpolated at each stage point,
// the lazy set-up, the fields at the new positions and the next step's velocity
// (InternalForwardAdvectionLoop body, time_loop.c:67-90, at a fixed dt).
static int fixed_step(Ctx& c, double dt) {
  c.adv_dt = dt;
  CHK(advect_swarm(c, dt));
  if (!c.analytical) {
    if (!setup_complete(c)) CHK(partitions_setup(c, c.setup_all));
    CHK(rbf_interpolation(c, c.fields.p, c.nf));
  }
  return init_step_velocity(c);
}

// A step can be captured once its launches are fixed: RK1/2/4 with interpolated
// velocities through the cell-table evaluation, every patch ready, the evaluation
// shape chosen, and the swarm stored in cell order (so every step swaps the
// particle buffers and two graphs, one per parity, cover the loop).
static bool graph_eligible(const Ctx& c) {
  if (!c.use_graph || c.analytical || c.scheme > 2 || !c.reorder || c.Np == 0) return false;
  if (c.eval_mode == 2 || (c.eval_mode == 0 && !c.memb_current) || !eval_blk_supported(c.nf, c.g.dim, eval_f32(c)) || !c.blk_valid) return false;
  if (!c.sorted_valid || !c.mem_pool.p) return false;
  if (c.tuner.planned(eval_bucket(c)) < 0) return false;
  return setup_complete(c);
}

static int capture_step(Ctx& c, double dt, hipGraphExec_t& exec) {
  const char* fn = "lagrbf_advect_steps";
  HIPCHK(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
  c.capturing = true;
  const int r = fixed_step(c, dt);
  c.capturing = false;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c.s, &graph);
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess || !graph) return set_err(E_LIB, fn, "graph capture of the step failed: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiateWithFlags(&exec, graph, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return set_err(E_LIB, fn, "graph instantiation failed: %s", hipGetErrorString(ei));
  return 0;
}

static int advect_steps(Ctx& c, int nsteps, double dt) {
  const char* fn = "lagrbf_advect_steps";
  if (!c.advection) return set_err(E_NULL, fn, "Advection context is NULL (set -advection)");
  if (!c.points_filled) return set_err(E_NULL, fn, "Interpolation swarm (ctx->swi) is NULL");
  if (!c.analytical && c.nf < c.g.dim) return set_err(E_SIZ, fn, "n_interpolation_fields (%d) < dim (%d)", c.nf, c.g.dim);
  if (nsteps < 0) return set_err(E_RANGE, fn, "nsteps must be >= 0");
  CHK(init_step_velocity(c));
  // graphs live for this call only: nothing outside it can move their buffers
  hipEvent_t g0 = nullptr, g1 = nullptr;
  int graph_run = 0;
  int rc = 0;
  for (int s = 0; s < nsteps && !rc; s++) {
    // the call's first step runs eagerly: it allocates what the reorder and the
    // evaluation need and fixes the evaluation shape before anything is captured
    if (s == 0 || !graph_eligible(c)) {
      rc = fixed_step(c, dt);
      continue;
    }
    // the step's host bookkeeping runs while capturing; the graph then runs the work
    hipGraphExec_t& ex = c.gexec[c.parity];
    if (!ex) {
      rc = capture_step(c, dt, ex);
      if (rc) break;
    } else {
      // replay: on the host a step only swaps the particle buffers (the evaluation
      // leaves the swarm grouped, as after the captured step)
      swap_particle_buffers(c);
      c.sorted_valid = true;
      c.adv_dt = dt;
    }
    if (!g0) {
      if ((rc = event_take(c, g0)) || (rc = event_take(c, g1))) break;
      if (hipEventRecord(g0, c.s) != hipSuccess) { rc = set_err(E_LIB, fn, "hipEventRecord failed"); break; }
    }
    const hipError_t e = hipGraphLaunch(ex, c.s);
    if (e != hipSuccess) { rc = set_err(E_LIB, fn, "hipGraphLaunch failed: %s", hipGetErrorString(e)); break; }
    graph_run++;
  }
  if (g0) {
    if (!rc && hipEventRecord(g1, c.s) == hipSuccess && hipEventSynchronize(g1) == hipSuccess) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, g0, g1) == hipSuccess) {
        c.graph_ms += ms;
        c.graph_steps += graph_run;
      }
    }
    c.ev_pool.push_back(g0);
    c.ev_pool.push_back(g1);
  }
  HIPCHK(hipStreamSynchronize(c.s));
  destroy_graphs(c);
  CHK(rc);
  CHK(flush_eval_timing(c));
  return update_removed(c);
}

// Velocity at the current stage points for stage fraction cst (time t + cst dt of
// each particle): the interpolated field, or with -lagrbf_time_blend the blend
// (1 - tau) vA + tau vB of the two snapshots.
static int adaptive_velocity(Ctx& c, double* dst, double cst) {
  const int dim = c.g.dim;
  if (!c.blend) return rbf_interpolation(c, dst, dim);
  // -1 (default) and 1: fused where the cell table evaluates 2d fields, which is
  // the FP32 evaluation (the FP64 one would need 96 accumulator registers per lane)
  const bool f32 = eval_f32(c);
  const bool fused = c.blend_fused != 0 && eval_blk_supported(2 * dim, dim, f32);
  if (fused) {
    // both snapshots in one evaluation (the 2d fields of the data): phi once per
    // particle and node; each field's sums are those of the per-snapshot passes
    CHK(rbf_interpolation(c, c.vtmp.p, 2 * dim));
    launch_blend(dst, c.vtmp.p, c.vtmp.p + (size_t)dim * c.ld, c.ld, c.Np, dim, c.alive.p, c.ptime.p, c.pdt.p, cst,
                 c.blend_t[0], c.blend_t[1], c.s);
    return 0;
  }
  CHK(rbf_interpolation(c, dst, dim, 0));
  CHK(rbf_interpolation(c, c.vtmp.p, dim, 1));
  launch_blend(dst, dst, c.vtmp.p, c.ld, c.Np, dim, c.alive.p, c.ptime.p, c.pdt.p, cst, c.blend_t[0], c.blend_t[1],
               c.s);
  return 0;
}

// Per-particle adaptive Dormand-Prince 5(4) from t0 to t1 (no reference
// counterpart: dormand_prince.c steps the whole swarm with one dt).  Every
// particle starts at t0 with dt = min(-dt, t1 - t0); each attempt runs the six
// stages for all unfinished particles with their own dt, then accepts or rejects
// per particle and adapts its dt (adaptive.hip).  Particles that reach t1 stop;
// particles that leave the domain are removed as in the fixed-dt loops.  The
// lazy set-up pass runs between attempts while patches are missing.
// Adaptive loop: store the swarm in cell order, finished and removed particles
// last, moving the per-particle integration state with it (start position x0,
// first-same-as-last stage k1, time, step size, finished flag; the self index
// stays the identity).  The per-stage grouping then finds long same-cell runs
// (one atomic per run) and the evaluation reads contiguous particles.
static int reorder_adaptive(Ctx& c) {
  if (!c.reorder || c.Np == 0) return 0;
  if (!c.sorted_valid) {
    launch_sort_particles(c.pcell.p, c.alive.p, c.Np, c.g.C, c.pstart.p, c.perm.p, c.pscratch.p, c.s);
    c.sorted_valid = true;
  }
  const size_t ld = (size_t)c.ld;
  const int dim = c.g.dim;
  HIPCHK(c.ptime_alt.reserve(ld));
  HIPCHK(c.pdt_alt.reserve(ld));
  HIPCHK(c.pfin_alt.reserve(ld));
  HIPCHK(c.psusp_alt.reserve(ld));
  // vadv[1] and vadv[2] are recomputed by every attempt: they take k1 and x0
  ReorderSet r{};
  r.src[0] = c.x0.p; r.dst[0] = c.vadv[2].p; r.rows[0] = dim;
  r.src[1] = c.vadv[0].p; r.dst[1] = c.vadv[1].p; r.rows[1] = dim;
  r.src[2] = c.ptime.p; r.dst[2] = c.ptime_alt.p; r.rows[2] = 1;
  r.src[3] = c.pdt.p; r.dst[3] = c.pdt_alt.p; r.rows[3] = 1;
  r.nd = 4;
  r.isrc[0] = c.pfin.p; r.idst[0] = c.pfin_alt.p;
  r.isrc[1] = c.psusp.p; r.idst[1] = c.psusp_alt.p;
  r.ni = 2;
  HIPCHK(launch_reorder(r, c.perm.p, c.Np, c.ld, c.s));
  swap_buf(c.x0, c.vadv[2]);
  swap_buf(c.vadv[0], c.vadv[1]);
  swap_buf(c.ptime, c.ptime_alt);
  swap_buf(c.pdt, c.pdt_alt);
  swap_buf(c.pfin, c.pfin_alt);
  swap_buf(c.psusp, c.psusp_alt);
  return reorder_particles(c);
}

static int advect_adaptive(Ctx& c, double t0, double t1, int max_attempts) {
  const char* fn = "lagrbf_advect_adaptive";
  if (!c.advection) return set_err(E_NULL, fn, "Advection context is NULL (set -advection)");
  if (!c.points_filled) return set_err(E_NULL, fn, "Interpolation swarm (ctx->swi) is NULL");
  if (!c.data_filled) return set_err(E_NULL, fn, "Data have not been provided");
  const int dim = c.g.dim, Np = c.Np;
  if (c.nf < dim) return set_err(E_SIZ, fn, "n_interpolation_fields (%d) < dim (%d)", c.nf, dim);
  if (!(t1 > t0)) return set_err(E_RANGE, fn, "t_end (%g) must exceed t_start (%g)", t1, t0);
  if (max_attempts < 0) return set_err(E_RANGE, fn, "max_attempts must be >= 0");
  if (!(c.adv_dt > 0.0)) return set_err(E_RANGE, fn, "-dt (initial step) must be positive");
  c.analytical = false;
  const size_t ld = (size_t)c.ld;
  HIPCHK(c.x0.reserve(ld * dim));
  for (int j = 0; j < 7; j++) HIPCHK(c.vadv[j].reserve(ld * dim));
  if (c.blend) HIPCHK(c.vtmp.reserve(ld * 2 * dim));
  HIPCHK(c.ptime.reserve(ld));
  HIPCHK(c.pdt.reserve(ld));
  HIPCHK(c.pfin.reserve(ld));
  HIPCHK(c.pself.reserve(ld));
  HIPCHK(c.psusp.reserve(ld));
  launch_fill_i32(c.psusp.p, (int)ld, 0, c.s);
  CHK(partitions_setup(c, c.setup_all));
  const double dt0 = std::min(c.adv_dt, t1 - t0);
  launch_dp45_init(Np, c.alive.p, c.ptime.p, c.pdt.p, c.pfin.p, c.pself.p, t0, dt0, c.s);
  launch_save(c.px.p, c.x0.p, c.vel.p, nullptr, c.ld, Np, dim, c.alive.p, c.s);
  CHK(adaptive_velocity(c, c.vadv[0].p, 0.0));
  static const double A[6][6] = {
      {1.0 / 5.0, 0, 0, 0, 0, 0},
      {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
      {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
      {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
      {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
      {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
  static const double Cst[6] = {1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
  DBuf<unsigned long long> cnt;
  HIPCHK(cnt.reserve(3));
  // Once the swarm is reordered, the unfinished particles fill slots [0, live):
  // the attempt's launches cover that prefix only (the reorder itself covers the
  // whole swarm, so the finished tail moves with it).  The autotuner keeps the
  // whole swarm's density bucket.
  // The evaluations shrink from the whole swarm to a few stragglers, and the full
  // swarm is evaluated only in the first attempts: timing every launch shape there
  // costs more than the fastest shape saves (at 16.8M particles the shapes run the
  // loop within 1% of each other, and the trial took 0.2 s).  The loop runs the
  // tuner's default shape untimed unless a shape is forced.
  struct Prefix {
    Ctx& c;
    int np;
    ~Prefix() { c.Np = np; c.bucket_np = -1; c.tuner.untuned = false; }
  } restore{c, Np};
  c.tuner.untuned = true;
  int live = Np;
  for (int a = 0; a < max_attempts; a++) {
    CHK(reorder_adaptive(c));
    const int n = c.reorder ? live : Np;
    c.Np = n;
    c.bucket_np = Np;
    for (int s = 0; s < 6; s++) {
      const double* v[7];
      for (int j = 0; j <= s; j++) v[j] = c.vadv[j].p;
      StageArgs sa = stage_args(s + 1, v, A[s], 0.0, 1, 0, 0, c);
      sa.dt_req = c.pdt.p;
      sa.req = c.pself.p;
      sa.mul = 1.0;
      sa.div = 1.0;
      // trial points of stages 2-6 outside the domain reject the attempt; the
      // step's end point (stage 7) outside removes the particle
      sa.suspend = s < 5 ? c.psusp.p : nullptr;
      stage(c, c.x0.p, sa);
      CHK(adaptive_velocity(c, c.vadv[s + 1].p, Cst[s]));
    }
    HIPCHK(hipMemsetAsync(cnt.p, 0, 3 * sizeof(unsigned long long), c.s));
    const double* k[7];
    for (int j = 0; j < 7; j++) k[j] = c.vadv[j].p;
    launch_dp45_control(c.g, c.pcell.p, c.px.p, c.x0.p, k, c.vadv[0].p, c.ld, n, dim, c.alive.p, c.pfin.p, c.psusp.p, c.counters.p, c.ptime.p, c.pdt.p, t1,
                        c.dopri_atol, c.dopri_rtol, c.dopri_fac, c.dopri_fac_min, c.dopri_fac_max, cnt.p, c.s);
    c.Np = Np;
    c.bucket_np = -1;
    unsigned long long h[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    live = (int)h[2];
    c.dp_attempts += 1;
    c.dp_accepted += (double)h[0];
    c.dp_rejected += (double)h[1];
    c.dp_active = (double)h[2];
    c.sorted_valid = false;
    if (h[2] == 0) break;
    if (!setup_complete(c)) CHK(partitions_setup(c, c.setup_all));
  }
  launch_dp45_finish(Np, c.alive.p, c.pfin.p, c.s);
  c.sorted_valid = false;
  CHK(flush_eval_timing(c));
  return update_removed(c);
}

// Particle swarm field by name: "fields", the velocity field, "weights" (SoA, leading dimension ld)
static int particle_field(Ctx& c, const char* name, const double*& src, int& nfl, const char* fn) {
  if (!name || !strcmp(name, "fields")) { src = c.fields.p; nfl = c.nf; return 0; }
  if (!strcmp(name, kVelocity) || (c.advection && c.velocity_name == name)) { src = c.vel.p; nfl = c.g.dim; return 0; }
  if (!strcmp(name, "weights")) { src = c.w.p; nfl = c.g.nen; return 0; }
  return set_err(E_WRONG, fn, "unknown field '%s'", name);
}

// lagrbf_output: rows by pid - 1 (AssembleArray_Allreduce, interface_utils.c:188-263)
static int output(Ctx& c, bool is_data, int n, double* coords, double* fields, int64_t* pidx, const char* fieldname) {
  const char* fn = "Interface_Output_Allreduce";
  const int dim = c.g.dim;
  HIPCHK(hipStreamSynchronize(c.s));
  if (is_data) {
    const int N = c.N;
    std::vector<double> xs((size_t)N * dim), vs((size_t)N * std::max(c.nf, 1));
    if (N > 0 && coords) HIPCHK(hipMemcpy(xs.data(), c.dx.p, sizeof(double) * N * dim, hipMemcpyDeviceToHost));
    if (N > 0 && fields && c.nf > 0) HIPCHK(hipMemcpy(vs.data(), c.dval.p, sizeof(double) * N * c.nf, hipMemcpyDeviceToHost));
    if (fields && fieldname && strcmp(fieldname, "fields")) return set_err(E_WRONG, fn, "data swarm has no field '%s'", fieldname);
    if (coords) memset(coords, 0, sizeof(double) * (size_t)n * dim);
    if (fields) memset(fields, 0, sizeof(double) * (size_t)n * c.nf);
    if (pidx) memset(pidx, 0, sizeof(int64_t) * (size_t)n);
    for (int p = 0; p < N; p++) {
      int64_t g = c.h_dpid[p] - 1;
      if (g < 0 || g >= n) return set_err(E_RANGE, fn, "pid %lld out of range [0, %d)", (long long)g, n);
      if (coords) for (int d = 0; d < dim; d++) coords[g * dim + d] = xs[(size_t)p * dim + d];
      if (fields) for (int f = 0; f < c.nf; f++) fields[g * c.nf + f] = vs[(size_t)p * c.nf + f];
      if (pidx) pidx[g] = c.h_dpid[p];
    }
    return 0;
  }
  const int Np = c.Np;
  const double* src = nullptr;
  int nfl = 0;
  if (fields) CHK(particle_field(c, fieldname, src, nfl, fn));
  std::vector<double> xs((size_t)dim * c.ld), fs((size_t)std::max(nfl, 1) * c.ld);
  std::vector<int32_t> al(std::max(Np, 1)), orig(std::max(Np, 1));
  if (Np > 0) {
    if (coords) HIPCHK(hipMemcpy(xs.data(), c.px.p, sizeof(double) * dim * c.ld, hipMemcpyDeviceToHost));
    if (fields && nfl > 0) HIPCHK(hipMemcpy(fs.data(), src, sizeof(double) * nfl * c.ld, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(al.data(), c.alive.p, sizeof(int32_t) * Np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(orig.data(), c.porig.p, sizeof(int32_t) * Np, hipMemcpyDeviceToHost));
  }
  if (coords) memset(coords, 0, sizeof(double) * (size_t)n * dim);
  if (fields) memset(fields, 0, sizeof(double) * (size_t)n * nfl);
  if (pidx) memset(pidx, 0, sizeof(int64_t) * (size_t)n);
  for (int p = 0; p < Np; p++) {
    if (!al[p]) continue;
    const int64_t pid = c.h_ppid[orig[p]];
    int64_t g = pid - 1;
    if (g < 0 || g >= n) return set_err(E_RANGE, fn, "pid %lld out of range [0, %d)", (long long)g, n);
    if (coords) for (int d = 0; d < dim; d++) coords[g * dim + d] = xs[(size_t)d * c.ld + p];
    if (fields) for (int f = 0; f < nfl; f++) fields[g * nfl + f] = fs[(size_t)f * c.ld + p];
    if (pidx) pidx[g] = pid;
  }
  return 0;
}

// Live particles in fill order: coordinates [n*dim] (when xs), rows of the field
// src [n*nfl] (when src) and pids (when pid), all AoS on the host.
static int gather_live(Ctx& c, const double* src, int nfl, std::vector<double>* xs, std::vector<double>* fs,
                       std::vector<int64_t>* pid) {
  const int dim = c.g.dim, Np = c.Np;
  HIPCHK(hipStreamSynchronize(c.s));
  std::vector<double> hx, hf;
  std::vector<int32_t> al(std::max(Np, 1)), orig(std::max(Np, 1));
  if (Np > 0) {
    if (xs) { hx.resize((size_t)dim * c.ld); HIPCHK(hipMemcpy(hx.data(), c.px.p, sizeof(double) * hx.size(), hipMemcpyDeviceToHost)); }
    if (src && nfl > 0) { hf.resize((size_t)nfl * c.ld); HIPCHK(hipMemcpy(hf.data(), src, sizeof(double) * hf.size(), hipMemcpyDeviceToHost)); }
    HIPCHK(hipMemcpy(al.data(), c.alive.p, sizeof(int32_t) * Np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(orig.data(), c.porig.p, sizeof(int32_t) * Np, hipMemcpyDeviceToHost));
  }
  std::vector<int32_t> slot(std::max(Np, 1), -1);
  for (int p = 0; p < Np; p++)
    if (al[p]) slot[orig[p]] = p;
  if (xs) xs->clear();
  if (fs) fs->clear();
  if (pid) pid->clear();
  for (int i = 0; i < Np; i++) {
    const int p = slot[i];
    if (p < 0) continue;
    if (xs) for (int d = 0; d < dim; d++) xs->push_back(hx[(size_t)d * c.ld + p]);
    if (fs) for (int f = 0; f < nfl; f++) fs->push_back(hf[(size_t)f * c.ld + p]);
    if (pid) pid->push_back(c.h_ppid[i]);
  }
  return 0;
}

static int make_dirs(const std::string& path, const char* fn) {
  std::string cur;
  size_t b = 0;
  while (b <= path.size()) {
    size_t e = path.find('/', b);
    if (e == std::string::npos) e = path.size();
    cur = path.substr(0, e);
    if (!cur.empty()) {
      struct stat st;
      if (stat(cur.c_str(), &st) != 0) {
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
          return set_err(io::E_FILE_OPEN, fn, "cannot create directory %s: %s", cur.c_str(), strerror(errno));
      } else if (!S_ISDIR(st.st_mode)) {
        return set_err(io::E_FILE_OPEN, fn, "%s exists and is not a directory", cur.c_str());
      }
    }
    b = e + 1;
  }
  return 0;
}

// "step%06d", with "_forward" / "_backward" for direction 1 / 2 (swarm_output.c:196-197)
static std::string step_prefix(int64_t step, int direction) {
  char sp[64];
  snprintf(sp, sizeof(sp), "step%06lld", (long long)step);
  std::string s = sp;
  if (direction == 1) s += "_forward";
  if (direction == 2) s += "_backward";
  return s;
}

static const char* kCoorField = "DMSwarmPIC_coor";
static const char* kPidField = "DMSwarm_pid";
static const char* kInterpName = "swarm_interp";

// SwarmOutput, swarm_output.c:169-234: <root>/step<k>/<prefix>.swarm_interp_<field>.pbvec
// (coordinates first, then each requested field; "DMSwarm_pid" as an IS .pbis),
// plus an entry of <root>/timeseries_X.pvd.  Rows are the live particles in fill
// order.  permute reproduces SwarmOutput_VecCreate's reordering
// (swarm_output.c:6-43), which reads the row-major field as if it were stored
// component-major.
static int swarm_output(Ctx& c, const char* root, int64_t step, double time, int nfields, const char** fields,
                        int direction, bool permute) {
  const char* fn = "SwarmOutput";
  if (!root) return set_err(E_NULL, fn, "output directory is NULL");
  if (nfields < 0 || (nfields > 0 && !fields)) return set_err(E_NULL, fn, "field list is NULL");
  if (direction < 0 || direction > 2) return set_err(E_RANGE, fn, "direction must be 0 (none), 1 (forward) or 2 (backward)");
  const std::string stepdir = "step" + std::to_string((long long)step);
  const std::string outdir = std::string(root) + "/" + stepdir;
  CHK(make_dirs(outdir, fn));
  char sp[64];
  snprintf(sp, sizeof(sp), "step%06lld", (long long)step);
  const std::string pvd = std::string(root) + "/timeseries_X.pvd";
  const std::string entry = std::string(sp) + "_coords.pbvec";
  CHK(io::pvd_open_append(c.pvd_written, step, pvd.c_str(), time, entry.c_str(), stepdir.c_str()));
  c.pvd_written = true;
  const std::string prefix = step_prefix(step, direction);
  auto fname = [&](const char* field, const char* ext) {
    return outdir + "/" + prefix + "." + kInterpName + "_" + field + ext;
  };
  const int dim = c.g.dim;
  std::vector<double> xs, fs;
  std::vector<int64_t> pid;
  CHK(gather_live(c, nullptr, 0, &xs, nullptr, &pid));
  const int64_t n = (int64_t)pid.size();
  CHK(io::vec_write(fname(kCoorField, ".pbvec").c_str(), n * dim, xs.data(), dim, 4));
  for (int f = 0; f < nfields; f++) {
    const char* name = fields[f] ? fields[f] : "";
    if (!strcmp(name, kPidField)) {
      CHK(io::is_write(fname(name, ".pbis").c_str(), n, pid.data(), 4));
      continue;
    }
    const double* src = nullptr;
    int nfl = 0;
    CHK(particle_field(c, name, src, nfl, fn));
    CHK(gather_live(c, src, nfl, nullptr, &fs, nullptr));
    if (permute && nfl > 1) {
      std::vector<double> pv(fs.size());
      for (int64_t p = 0; p < n; p++)
        for (int k = 0; k < nfl; k++) pv[(size_t)(p * nfl + k)] = fs[(size_t)(k * n + p)];
      fs.swap(pv);
    }
    CHK(io::vec_write(fname(name, ".pbvec").c_str(), n * nfl, fs.data(), nfl, 4));
  }
  return 0;
}

// SwarmLoadFromFiles, swarm_load.c:133-201, into the particle swarm: coordinates
// from <dir>/<prefix>.swarm_interp_DMSwarmPIC_coor.pbvec, then "fields" (.pbvec)
// and "DMSwarm_pid" (.pbis) when listed; the swarm is refilled with them.
static int swarm_load(Ctx& c, const char* dirname, int64_t step, int direction, int nfields, const char** fields) {
  const char* fn = "SwarmLoadFromFiles";
  if (!dirname) return set_err(E_NULL, fn, "directory is NULL");
  if (nfields < 0 || (nfields > 0 && !fields)) return set_err(E_NULL, fn, "field list is NULL");
  if (direction < 0 || direction > 2) return set_err(E_RANGE, fn, "direction must be 0 (none), 1 (forward) or 2 (backward)");
  const std::string base = std::string(dirname) + "/" + step_prefix(step, direction) + "." + kInterpName + "_";
  const int dim = c.g.dim;
  std::vector<double> xs, vals;
  std::vector<int64_t> pid;
  int bs = 1;
  const std::string cf = base + kCoorField + ".pbvec";
  CHK(io::vec_read(cf.c_str(), 4, xs, bs));
  if (bs != dim) return set_err(E_SIZ, fn, "%s: block size %d differs from the dimension %d", cf.c_str(), bs, dim);
  const int64_t np = (int64_t)xs.size() / dim;
  if (np > INT32_MAX) return set_err(E_RANGE, fn, "%lld points exceed the 32-bit swarm size", (long long)np);
  bool have_vals = false, have_pid = false;
  for (int f = 0; f < nfields; f++) {
    const char* name = fields[f] ? fields[f] : "";
    if (!strcmp(name, kPidField)) {
      const std::string pf = base + name + ".pbis";
      CHK(io::is_read(pf.c_str(), 4, pid));
      if ((int64_t)pid.size() != np)
        return set_err(E_SIZ, fn, "Loaded point index size %zu does not match swarm size %lld", pid.size(), (long long)np);
      have_pid = true;
    } else if (!strcmp(name, "fields")) {
      const std::string vf = base + name + ".pbvec";
      CHK(io::vec_read(vf.c_str(), 4, vals, bs));
      if ((int64_t)vals.size() / bs != np)
        return set_err(E_SIZ, fn, "Loaded vector size %lld does not match swarm size %lld", (long long)(vals.size() / bs),
                       (long long)np);
      if (bs != c.nf) return set_err(E_SIZ, fn, "%s: block size %d differs from n_interpolation_fields %d", vf.c_str(), bs, c.nf);
      have_vals = true;
    } else {
      return set_err(E_SUP, fn, "loading field '%s' is not supported (coordinates, \"fields\" and \"%s\" are)", name, kPidField);
    }
  }
  return fill_points(c, (int)np, xs.data(), have_vals ? vals.data() : nullptr, have_pid ? pid.data() : nullptr);
}

// fill a swarm from a loaded dataset (StagYY / PETSc binary ingest)
static int fill_from_dataset(Ctx& c, bool is_data, const io::Dataset& ds) {
  const char* fn = is_data ? "FillDataSwarm" : "FillInterpSwarm";
  if (ds.dim != c.g.dim) return set_err(E_SIZ, fn, "dataset dimension %d differs from the context's %d", ds.dim, c.g.dim);
  if (ds.nfields != c.nf)
    return set_err(E_SIZ, fn, "dataset has %d fields, n_interpolation_fields is %d", ds.nfields, c.nf);
  if (ds.n() > INT32_MAX) return set_err(E_RANGE, fn, "%lld points exceed the 32-bit swarm size", (long long)ds.n());
  const int n = (int)ds.n();
  const double* v = ds.fields.empty() ? nullptr : ds.fields.data();
  return is_data ? fill_data(c, n, ds.coords.data(), v, nullptr) : fill_points(c, n, ds.coords.data(), v, nullptr);
}

Ctx* create_from_view(const OptView& o) {
  Ctx* c = new Ctx();
  if (ctx_from_options(*c, o) || ctx_init_device(*c)) {
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return nullptr;
  }
  // the no-allocation entry point's scratch on this device (its own error is kept
  // for that entry point, which refuses to run without it)
  (void)noalloc_init_device(c->device);
  return c;
}

OptView current_options() {
  OptView v;
  Options env;
  if (const char* e = std::getenv("LAGRBF_OPTIONS")) env.insert_string(e);
  v.db = env.snapshot();
  for (auto& kv : global_options().snapshot()) v.db[kv.first] = kv.second;
  return v;
}

}  // namespace lrb

using namespace lrb;

#define CTX_OR_NULL(fnname)                                                       \
  Ctx* c = static_cast<Ctx*>(ctx);                                                \
  if (!c) return set_err(E_NULL, fnname, "LagRBFCtx is NULL");                    \
  if (hipSetDevice(c->device) != hipSuccess) return set_err(E_LIB, fnname, "hipSetDevice failed")

extern "C" {

void* lagrbf_create(void) {
  OptView o = current_options();
  return create_from_view(o);
}

void lagrbf_destroy(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->s);
  destroy_graphs(*c);
  release_events(*c);
  (void)hipStreamDestroy(c->s);
  delete c;
}

int lagrbf_fill_swarm(void* ctx, bool is_data, int32_t npoints, double* coords, double* values, int64_t* pidx) {
  CTX_OR_NULL(is_data ? "FillDataSwarm" : "FillInterpSwarm");
  if (npoints < 0) return set_err(E_RANGE, "lagrbf_fill_swarm", "npoints must be >= 0");
  if (npoints > 0 && !coords) return set_err(E_NULL, "lagrbf_fill_swarm", "coords is NULL");
  return is_data ? fill_data(*c, npoints, coords, values, pidx) : fill_points(*c, npoints, coords, values, pidx);
}

int lagrbf_interpolation(void* ctx) {
  CTX_OR_NULL("Interface_Interpolation");
  return interpolation(*c);
}

int lagrbf_output(void* ctx, bool is_data, int32_t npoints, double* coords, double* fields, int64_t* pidx,
                  const char* fieldname) {
  CTX_OR_NULL("Interface_Output_Allreduce");
  if (npoints < 0) return set_err(E_RANGE, "lagrbf_output", "npoints must be >= 0");
  return output(*c, is_data, npoints, coords, fields, pidx, fieldname);
}

int lagrbf_avect_analytical(void* ctx, double time, double time_next) {
  CTX_OR_NULL("ForwardAnalyticalAdvection");
  return advect_analytical(*c, time, time_next);
}

int lagrbf_advect_interpolate(void* ctx, double time, double time_next) {
  CTX_OR_NULL("InternalForwardAdvectionLoop");
  return advect_interpolate(*c, time, time_next);
}

int lagrbf_advection_function(void* ctx, int (*velocity_fn)(double*, double*, void*), void* fn_ctx) {
  CTX_OR_NULL("AdvectionSetVelocityFunction");
  if (!c->advection) return set_err(E_NULL, "AdvectionSetVelocityFunction", "Advection context is NULL");
  c->vfn = velocity_fn; c->vuser = fn_ctx; c->analytical = true;
  return 0;
}

int lagrbf_advection_function_batch(void* ctx, int (*fnb)(int32_t, int32_t, double*, double*, void*), void* fn_ctx) {
  CTX_OR_NULL("AdvectionSetVelocityFunctionBatch");
  if (!c->advection) return set_err(E_NULL, "AdvectionSetVelocityFunctionBatch", "Advection context is NULL");
  c->vfn_batch = fnb; c->vuser = fn_ctx; c->analytical = true;
  return 0;
}

int lagrbf_reset_coefficients(void* ctx) {
  CTX_OR_NULL("LagRBF_PartitionsReset");
  return partitions_reset(*c);
}

void lagrbf_interpolate_driver(int32_t ndata, double* dcoords, double* dvalues, int32_t npts, double* pcoords,
                               double* pvalues, int32_t nglobal, char* options_file) {
  OptView o = current_options();
  if (options_file && options_file[0]) {
    Options f;
    int e = f.insert_file(options_file);
    if (e) { set_err(e, "lagrbf_interpolate", "cannot read options file %s", options_file); return; }
    for (auto& kv : f.snapshot()) o.db[kv.first] = kv.second;
  }
  Ctx* c = create_from_view(o);
  if (!c) return;
  int e = fill_data(*c, ndata, dcoords, dvalues, nullptr);
  if (!e) e = fill_points(*c, npts, pcoords, nullptr, nullptr);
  if (!e) e = interpolation(*c);
  if (!e) e = output(*c, false, nglobal, nullptr, pvalues, nullptr, "fields");
  lagrbf_destroy(c);
}

int lagrbf_set_option(const char* name, const char* value) {
  if (!name) return set_err(E_NULL, "lagrbf_set_option", "name is NULL");
  global_options().set(name, value ? value : "");
  return 0;
}

int lagrbf_clear_options(void) {
  global_options().clear();
  return 0;
}

int lagrbf_options_insert_file(const char* path) {
  if (!path) return set_err(E_NULL, "lagrbf_options_insert_file", "path is NULL");
  int e = global_options().insert_file(path);
  if (e) return set_err(e, "lagrbf_options_insert_file", "cannot read %s", path);
  return 0;
}

int lagrbf_advect_steps(void* ctx, int32_t nsteps, double dt) {
  CTX_OR_NULL("lagrbf_advect_steps");
  return advect_steps(*c, nsteps, dt);
}

int lagrbf_advect_adaptive(void* ctx, double t_start, double t_end, int32_t max_attempts) {
  CTX_OR_NULL("lagrbf_advect_adaptive");
  return advect_adaptive(*c, t_start, t_end, max_attempts);
}

int lagrbf_get_info(void* ctx, const char* key, double* value) {
  CTX_OR_NULL("lagrbf_get_info");
  if (!key || !value) return set_err(E_NULL, "lagrbf_get_info", "key/value is NULL");
  std::string k(key);
  if (k == "removed") { int e = update_removed(*c); if (e) return e; *value = c->removed; }
  else if (k == "fallback") *value = c->fallback;
  else if (k == "refine_iters") *value = c->refine_iters_max;
  else if (k == "refine_iters_mean") *value = c->refined_patches > 0 ? c->refine_iters_sum / c->refined_patches : 0.0;
  else if (k == "precision") *value = c->precision;
  else if (k == "patches_setup") *value = c->patches_setup;
  else if (k == "dt") *value = c->adv_dt;
  else if (k == "dopri_err") *value = c->dopri_err;
  else if (k == "eval_ms" || k == "eval_count") {
    if (int e = flush_eval_timing(*c)) return e;
    *value = k == "eval_ms" ? c->eval_ms : c->eval_count;
  }
  else if (k == "dp_attempts") *value = c->dp_attempts;
  else if (k == "dp_accepted") *value = c->dp_accepted;
  else if (k == "dp_rejected") *value = c->dp_rejected;
  else if (k == "dp_active") *value = c->dp_active;
  else if (k == "eval_variant") *value = c->last_variant;
  else if (k == "graph_steps") *value = c->graph_steps;
  else if (k == "graph_ms") *value = c->graph_ms;
  else if (k == "setup_ms") *value = c->setup_ms;
  else if (k == "n_particles") *value = c->Np;
  else if (k == "n_nodes") *value = c->N;
  else if (k == "max_patch") *value = c->max_patch;
  else if (k == "eps_w") *value = c->g.eps_w;
  else if (k == "eps_r") *value = c->g.eps_r;
  else if (k == "aug_fallback_bk") *value = c->aug_fallback_bk;
  else if (k == "big_refactor64") *value = c->big_refactor64;
  else if (k.rfind("refine_hist_", 0) == 0) {
    // refined patches that took i iterations (FP32 factors)
    const int i = std::atoi(k.c_str() + 12);
    *value = (i >= 0 && i < (int)c->refine_hist.size()) ? c->refine_hist[i] : 0.0;
  }
  else if (k == "poly_terms") *value = c->poly_mm;
  else if (k == "factor_trial_ms") *value = c->factor_trial_ms;
  else if (k == "setup_alloc_ms") *value = c->setup_alloc_ms;
  else if (k == "device_cache_bytes") *value = (double)dev_cached_bytes();
  else if (k == "setup_factor_ms") *value = c->setup_factor_ms;
  else if (k == "setup_solve_ms") *value = c->setup_solve_ms;
  else return set_err(E_WRONG, "lagrbf_get_info", "unknown key '%s'", key);
  return 0;
}

int64_t lagrbf_debug_get(void* ctx, const char* what, int32_t index, void* out, int64_t cap) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !what) return -set_err(E_NULL, "lagrbf_debug_get", "NULL argument");
  if (hipSetDevice(c->device) != hipSuccess || hipStreamSynchronize(c->s) != hipSuccess) return -E_LIB;
  std::string w(what);
  const void* src = nullptr;
  int64_t n = 0, esz = 4;
  if (w == "cellid_points") { src = c->pcell.p; n = c->Np; }
  else if (w == "cellid_data") { src = c->dcell.p; n = c->N; }
  else if (w == "cell_start") { src = c->cell_start.p; n = c->g.C + 1; }
  else if (w == "cell_nodes") { src = c->cell_nodes.p; n = c->N; }
  else if (w == "alive") { src = c->alive.p; n = c->Np; }
  else if (w == "patch_members" || w == "patch_coeffs") {
    if (index < 0 || index >= c->g.P) return -set_err(E_RANGE, "lagrbf_debug_get", "patch %d out of range", index);
    if (!c->h_setup[index]) return 0;
    const int64_t off = c->h_mem_off[index];
    n = c->h_n[index];
    if (w == "patch_members") src = c->mem_pool.p + off;
    else { src = c->coef_pool.p + off * c->nf; n *= c->nf; esz = 8; }
  } else if (w == "patch_beta") {
    // polynomial coefficients of a patch, row-major [m][nf] (extension)
    if (index < 0 || index >= c->g.P) return -set_err(E_RANGE, "lagrbf_debug_get", "patch %d out of range", index);
    if (!c->h_setup[index] || c->poly_deg < 0) return 0;
    src = c->beta_pool.p + (size_t)index * c->poly_mm * std::max(c->nf, 1);
    n = (int64_t)c->h_m[index] * c->nf;
    esz = 8;
  } else if (w == "patch_eps") {
    if (index < 0 || index >= c->g.P) return -set_err(E_RANGE, "lagrbf_debug_get", "patch %d out of range", index);
    if (out && cap >= 1) *static_cast<double*>(out) = c->h_eps[index];
    return 1;
  } else if (w == "patch_flags") {
    if (index < 0 || index >= c->g.P) return -set_err(E_RANGE, "lagrbf_debug_get", "patch %d out of range", index);
    if (out && cap >= 3) {
      int32_t* o = static_cast<int32_t*>(out);
      o[0] = c->h_setup[index]; o[1] = c->h_solved[index]; o[2] = c->h_n[index];
    }
    return 3;
  } else return -set_err(E_WRONG, "lagrbf_debug_get", "unknown item '%s'", what);
  if (out && n > 0 && src) {
    if (cap < n) return -set_err(E_SIZ, "lagrbf_debug_get", "buffer too small (%lld < %lld)", (long long)cap, (long long)n);
    if (w == "cellid_points" || w == "alive") {
      // per-particle arrays are returned in fill order (the swarm is stored in cell order)
      std::vector<int32_t> v((size_t)n), orig((size_t)n);
      if (hipMemcpy(v.data(), src, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) return -E_LIB;
      if (hipMemcpy(orig.data(), c->porig.p, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) return -E_LIB;
      int32_t* o = static_cast<int32_t*>(out);
      for (int64_t q = 0; q < n; q++) o[orig[q]] = v[q];
    } else if (hipMemcpy(out, src, (size_t)(n * esz), hipMemcpyDeviceToHost) != hipSuccess) return -E_LIB;
  }
  return n;
}

int lagrbf_swarm_output(void* ctx, const char* root, int64_t step, double time, int32_t nfields, const char** fields,
                        int32_t direction, bool permute) {
  CTX_OR_NULL("SwarmOutput");
  return swarm_output(*c, root, step, time, nfields, fields, direction, permute);
}

int lagrbf_swarm_load(void* ctx, const char* dirname, int64_t step, int32_t direction, int32_t nfields,
                      const char** fields) {
  CTX_OR_NULL("SwarmLoadFromFiles");
  return swarm_load(*c, dirname, step, direction, nfields, fields);
}

int lagrbf_load_data(void* ctx, int64_t step) {
  CTX_OR_NULL("SwarmDataLoad");
  io::Dataset ds;
  CHK(io::load_data_from_options(current_options(), step, c->g.dim, c->nf, ds));
  return fill_from_dataset(*c, true, ds);
}

int lagrbf_fill_swarm_from_dataset(void* ctx, bool is_data, void* dataset) {
  CTX_OR_NULL(is_data ? "FillDataSwarm" : "FillInterpSwarm");
  if (!dataset) return set_err(E_NULL, "lagrbf_fill_swarm_from_dataset", "dataset is NULL");
  return fill_from_dataset(*c, is_data, *static_cast<io::Dataset*>(dataset));
}

int lagrbf_last_error(char* buf, int32_t len) {
  if (!buf || len <= 0) return E_NULL;
  snprintf(buf, (size_t)len, "%s", g_last_error.c_str());
  return 0;
}

}  // extern "C"
