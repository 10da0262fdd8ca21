// capi.hip -- extern "C" entry points (include/spgrid_topo_opt_amd.h).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/spgrid_topo_opt_amd.h"
#include "fem_wire.hpp"
#include "layout.hpp"
#include "solver.hpp"

using namespace spg;

// What a kept fem_solve state carries to the next call besides the solver:
// the last displacement per input block (the reference keeps its SPGrid u
// channel, SPGrid_CPU_Solver.cpp:389-404) and the zero-guess residual of
// the first solve (state.initial_residual_linf, :318-326).
struct WarmState {
  std::vector<int32_t> base;  // [nb*3] input blocks of the last solve
  NodeArray<double> u;        // [nb*64*3] their displacement
  double initial_linf = 0.0;
  bool valid = false;
};

struct spg_solver {
  std::unique_ptr<Solver> s;
  // inputs kept for reuse checks and density updates
  std::vector<int32_t> base;
  std::vector<int32_t> dcoord, daxis;
  std::vector<double> dvalue;
  spg_config cfg;
  WarmState warm;
};

static thread_local std::string g_err;

// largest bottom (unknowns) fem_solve factors as a dense inverse
constexpr int32_t kFemSolveDenseBottomDofs = 4096;

// a displacement block of the FEMOutputs stream: int32 base[3], 64 x 4 doubles
constexpr size_t kRecordBytes = 12 + 64 * 32;

#define API_TRY try {
#define API_CATCH                          \
  }                                        \
  catch (SpgError & e) {                   \
    g_err = e.what();                      \
    return e.code;                         \
  }                                        \
  catch (LayoutError & e) {                \
    g_err = e.what();                      \
    return e.code;                         \
  }                                        \
  catch (WireError & e) {                  \
    g_err = e.what();                      \
    return e.code;                         \
  }                                        \
  catch (std::exception & e) {             \
    g_err = e.what();                      \
    return SPG_E_INVALID;                  \
  }                                        \
  catch (...) {                            \
    g_err = "unknown error";               \
    return SPG_E_INVALID;                  \
  }

static ProblemDesc make_desc(const spg_config* cfg, int64_t nblocks, const int32_t* base, const double* density,
                             int64_t ndir, const int32_t* dc, const int32_t* da, const double* dv) {
  ProblemDesc p;
  for (int d = 0; d < 3; ++d) p.res[d] = cfg->res[d];
  p.dx = cfg->dx;
  p.mu = cfg->mu;
  p.lambda = cfg->lambda;
  p.penalty = cfg->penalty;
  p.min_stiffness = cfg->min_stiffness;
  p.min_density = cfg->min_density;
  p.mg_levels = cfg->mg_levels;
  p.smooth_iters = cfg->smooth_iters;
  p.boundary_iters = cfg->boundary_iters;
  p.nblocks = nblocks;
  p.block_base = base;
  p.density = density;
  p.ndir = ndir;
  p.dir_coord = dc;
  p.dir_axis = da;
  p.dir_value = dv;
  return p;
}

namespace spg {
// error text for entry points defined in other translation units
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace spg

extern "C" {

const char* spg_last_error(void) { return g_err.c_str(); }
const char* spg_version(void) { return "spgrid_topo_opt_amd 0.1 (gfx950)"; }

int spg_device_count(int32_t* count) {
  API_TRY
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return SPG_OK;
  API_CATCH
}

int spg_create(const spg_config* cfg, int64_t nblocks, const int32_t* block_base, const double* density,
               int64_t ndir, const int32_t* dir_coord, const int32_t* dir_axis, const double* dir_value,
               spg_solver** out) {
  API_TRY
  if (!cfg || !out || (nblocks > 0 && (!block_base || !density))) throw SpgError(SPG_E_INVALID, "null argument");
  if (nblocks < 0 || ndir < 0) throw SpgError(SPG_E_INVALID, "negative block or Dirichlet count");
  if (ndir > 0 && (!dir_coord || !dir_axis || !dir_value)) throw SpgError(SPG_E_INVALID, "null Dirichlet arrays");
  *out = nullptr;
  auto h = std::make_unique<spg_solver>();
  h->cfg = *cfg;
  h->base.assign(block_base, block_base + nblocks * 3);
  h->dcoord.assign(dir_coord, dir_coord + ndir * 3);
  h->daxis.assign(dir_axis, dir_axis + ndir);
  h->dvalue.assign(dir_value, dir_value + ndir);
  ProblemDesc pd = make_desc(cfg, nblocks, block_base, density, ndir, dir_coord, dir_axis, dir_value);
  h->s = std::make_unique<Solver>(*cfg, pd);
  *out = h.release();
  return SPG_OK;
  API_CATCH
}

void spg_destroy(spg_solver* s) { delete s; }

int spg_update_density(spg_solver* s, const double* density) {
  API_TRY
  if (!s || !density) throw SpgError(SPG_E_INVALID, "null argument");
  s->s->update_density(density);
  return SPG_OK;
  API_CATCH
}

int spg_update_density_device(spg_solver* s, const double* density_dev) {
  API_TRY
  if (!s || !density_dev) throw SpgError(SPG_E_INVALID, "null argument");
  s->s->update_density_device(density_dev);
  return SPG_OK;
  API_CATCH
}

int spg_set_forces(spg_solver* s, int64_t n, const int32_t* coord, const double* value) {
  API_TRY
  if (!s || (n > 0 && (!coord || !value))) throw SpgError(SPG_E_INVALID, "null argument");
  if (n < 0) throw SpgError(SPG_E_INVALID, "negative force count");
  s->s->set_forces(n, coord, value);
  return SPG_OK;
  API_CATCH
}

int spg_solve(spg_solver* s, const spg_solve_params* prm, double* u_out4, spg_info* info) {
  API_TRY
  if (!s || !prm) throw SpgError(SPG_E_INVALID, "null argument");
  spg_info tmp{};
  const int rc = s->s->solve(*prm, info ? info : &tmp);
  if (u_out4) s->s->pack_output(u_out4);
  if (rc != SPG_OK) {
    const spg_info& i = info ? *info : tmp;
    g_err = "solve ended with code " + std::to_string(rc) + " after " + std::to_string(i.iterations) +
            " iterations, residual " + std::to_string(i.r_linf) + " of " + std::to_string(i.r0_linf);
  }
  return rc;
  API_CATCH
}

int spg_get_displacement_device(spg_solver* s, double* u_dev) {
  API_TRY
  if (!s || !u_dev) throw SpgError(SPG_E_INVALID, "null argument");
  s->s->displacement_in3(u_dev);
  return SPG_OK;
  API_CATCH
}

int spg_set_initial_guess_device(spg_solver* s, const double* u_dev) {
  API_TRY
  if (!s || !u_dev) throw SpgError(SPG_E_INVALID, "null argument");
  s->s->set_initial_guess_in3(u_dev);
  return SPG_OK;
  API_CATCH
}

int spg_solve_adjoint_device(spg_solver* s, const spg_solve_params* prm, const double* g_dev, double* lam_dev,
                             spg_info* info) {
  API_TRY
  if (!s || !prm || !g_dev || !lam_dev) throw SpgError(SPG_E_INVALID, "null argument");
  spg_info tmp{};
  return s->s->solve_adjoint(*prm, g_dev, lam_dev, info ? info : &tmp);
  API_CATCH
}

int spg_sensitivity_device(spg_solver* s, const double* u_dev, const double* lam_dev, double* dc_dev,
                           double* objective) {
  API_TRY
  if (!s || !dc_dev) throw SpgError(SPG_E_INVALID, "null argument");
  const double obj = s->s->sensitivity(u_dev, lam_dev, dc_dev);
  if (objective) *objective = obj;
  return SPG_OK;
  API_CATCH
}

int spg_num_levels(spg_solver* s, int32_t* levels) {
  API_TRY
  if (!s || !levels) throw SpgError(SPG_E_INVALID, "null argument");
  *levels = s->s->levels();
  return SPG_OK;
  API_CATCH
}

int spg_get_info(spg_solver* s, spg_info* info) {
  API_TRY
  if (!s || !info) throw SpgError(SPG_E_INVALID, "null argument");
  *info = spg_info{};
  s->s->static_info(info);
  return SPG_OK;
  API_CATCH
}

int spg_level_size(spg_solver* s, int32_t level, int64_t* nblocks) {
  API_TRY
  if (!s || !nblocks) throw SpgError(SPG_E_INVALID, "null argument");
  if (level < 0 || level >= s->s->levels()) throw SpgError(SPG_E_INVALID, "level out of range");
  *nblocks = s->s->layout().lv[level].nb;
  return SPG_OK;
  API_CATCH
}

int spg_level_nodes(spg_solver* s, int32_t level, int32_t* xyz, uint8_t* flags) {
  API_TRY
  if (!s) throw SpgError(SPG_E_INVALID, "null argument");
  if (level < 0 || level >= s->s->levels()) throw SpgError(SPG_E_INVALID, "level out of range");
  const LevelDesc& D = s->s->layout().lv[level];
  for (int64_t b = 0; b < D.nb; ++b)
    for (int l = 0; l < 64; ++l) {
      if (xyz) {
        xyz[(b * 64 + l) * 3 + 0] = D.base[3 * b] + (l >> 4);
        xyz[(b * 64 + l) * 3 + 1] = D.base[3 * b + 1] + ((l >> 2) & 3);
        xyz[(b * 64 + l) * 3 + 2] = D.base[3 * b + 2] + (l & 3);
      }
      if (flags) flags[b * 64 + l] = D.flags[b * 64 + l];
    }
  return SPG_OK;
  API_CATCH
}

int spg_kernel(spg_solver* s, int32_t op, int32_t level, int32_t precision, int32_t forward, const double* in,
               const double* b, double* out) {
  API_TRY
  if (!s || !out) throw SpgError(SPG_E_INVALID, "null argument");
  s->s->run_kernel(op, level, precision, forward, in, b, out);
  return SPG_OK;
  API_CATCH
}

int spg_level_stencils(spg_solver* s, int32_t level, double* out) {
  API_TRY
  if (!s || !out) throw SpgError(SPG_E_INVALID, "null argument");
  s->s->level_stencils(level, out);
  return SPG_OK;
  API_CATCH
}

int spg_bench_apply(spg_solver* s, int32_t precision, int32_t reps, double* seconds) {
  API_TRY
  if (!s || !seconds || reps <= 0) throw SpgError(SPG_E_INVALID, "bad argument");
  *seconds = s->s->bench_apply(precision, reps);
  return SPG_OK;
  API_CATCH
}

int spg_bench_kernel(spg_solver* s, int32_t op, int32_t level, int32_t precision, int32_t arg, int32_t reps,
                     double* seconds) {
  API_TRY
  if (!s || !seconds || reps <= 0) throw SpgError(SPG_E_INVALID, "bad argument");
  *seconds = s->s->bench_kernel(op, level, precision, arg, reps);
  return SPG_OK;
  API_CATCH
}

// ------------------------------------------------------------------ wire
static void emit(const FemOutputs& o, void* output) {
  WireWriter w;
  write_outputs(w, o);
  std::vector<uint8_t>& buf = w.finalize();
  char* data = new char[buf.size()];
  std::memcpy(data, buf.data(), buf.size());
  *reinterpret_cast<void**>(output) = data;
}

static bool same_topology(const spg_solver* h, const FemInputs& in, const std::vector<int32_t>& base) {
  if (h->base != base) return false;
  if (h->daxis.size() != in.dirichlet_nodes.size()) return false;
  for (size_t i = 0; i < in.dirichlet_nodes.size(); ++i) {
    const auto& d = in.dirichlet_nodes[i];
    if (h->daxis[i] != d.axis || h->dvalue[i] != d.value) return false;
    for (int c = 0; c < 3; ++c)
      if (h->dcoord[3 * i + c] != d.coord[c]) return false;
  }
  const spg_config& c = h->cfg;
  return c.res[0] == in.resolution[0] && c.res[1] == in.resolution[1] && c.res[2] == in.resolution[2] &&
         c.dx == in.dx && c.mu == in.global_mu && c.lambda == in.global_lambda &&
         c.penalty == in.get_double("penalty", 0.0) && c.min_stiffness == in.get_double("minimum_stiffness", 0.0) &&
         c.min_density == in.get_double("minimum_density", 0.0) && c.mg_levels == in.mg_level &&
         c.smooth_iters == in.pre_and_post_smoothing_iter && c.boundary_iters == in.get_int("boundary_smoothing", 5) &&
         c.device == in.get_int("device", 0);
}

// The previous displacement mapped onto a new block list by block
// coordinates (blocks new to the list start at zero).
static std::vector<double> carry_over(const WarmState& w, const std::vector<int32_t>& base) {
  std::unordered_map<uint64_t, size_t> prev;
  const size_t np = w.base.size() / 3;
  prev.reserve(np * 2);
  for (size_t b = 0; b < np; ++b) prev.emplace(bkey(w.base[3 * b], w.base[3 * b + 1], w.base[3 * b + 2]), b);
  const size_t nb = base.size() / 3;
  std::vector<double> u(nb * 192, 0.0);
  for (size_t b = 0; b < nb; ++b) {
    auto it = prev.find(bkey(base[3 * b], base[3 * b + 1], base[3 * b + 2]));
    if (it != prev.end()) std::memcpy(&u[b * 192], &w.u[it->second * 192], 192 * sizeof(double));
  }
  return u;
}

namespace {
// SPG_SETUP_TRACE=2: seconds per stage of fem_solve on stderr
struct StageTrace {
  bool on = false;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  StageTrace() {
    const char* e = std::getenv("SPG_SETUP_TRACE");
    on = e && std::atoi(e) >= 2;
  }
  void operator()(const char* what) {
    if (!on) return;
    const auto n = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[spg] fem_solve %-24s %.3f s\n", what, std::chrono::duration<double>(n - t).count());
    t = n;
  }
};
}  // namespace

// Drop-in for `fem_solve` (topo_opt/fem_interface.h:11; solver/SPGrid_CPU_Solver.cpp:26-409).
void fem_solve(void* input, void* output) {
  StageTrace trace;
  FemOutputs out;
  std::unique_ptr<char[]> emitted;  // the finished stream of a solve that ran
  out.success = false;
  if (!output) return;
  // the kept state: returned to the caller when keep_state is set (also
  // after a failed solve, so the caller can still release it), freed
  // otherwise (SPGrid_CPU_Solver.cpp:389-408)
  std::unique_ptr<spg_solver> h;
  bool keep = false;
  try {
    FemInputs in;
    WireReader r(input);
    read_inputs(r, in);
    trace("read inputs");
    h.reset(static_cast<spg_solver*>(in.get_state_ptr()));
    keep = in.get_int("keep_state", 0) != 0;
    for (int d = 0; d < 3; ++d) out.displacements.resolution[d] = in.resolution[d];
    if (!in.use_density_only) {
      // SPGrid_CPU_Solver.cpp:131-135
      out.info = "Unsupported mode! (mu/lambda grids are not supported; use_density_only must be true)";
      out.error_code = SPG_E_INVALID;
      throw SpgError(SPG_E_INVALID, out.info);
    }
    const size_t nb = in.density.blocks.size();
    std::vector<int32_t> base(nb * 3);
    NodeArray<double> dens(nb * 64);
    parallel_ranges(
        (int64_t)nb,
        [&](int64_t b0, int64_t b1) {
          for (int64_t b = b0; b < b1; ++b) {
            for (int c = 0; c < 3; ++c) base[3 * b + c] = in.density.blocks[b].base[c];
            std::memcpy(&dens[64 * b], in.density.blocks[b].data, 64 * sizeof(double));
          }
        },
        1024);
    std::vector<int32_t> dc(in.dirichlet_nodes.size() * 3), da(in.dirichlet_nodes.size());
    std::vector<double> dv(in.dirichlet_nodes.size());
    for (size_t i = 0; i < in.dirichlet_nodes.size(); ++i) {
      for (int c = 0; c < 3; ++c) dc[3 * i + c] = in.dirichlet_nodes[i].coord[c];
      da[i] = in.dirichlet_nodes[i].axis;
      dv[i] = in.dirichlet_nodes[i].value;
    }
    spg_config cfg{};
    for (int d = 0; d < 3; ++d) cfg.res[d] = in.resolution[d];
    cfg.dx = in.dx;
    cfg.mu = in.global_mu;
    cfg.lambda = in.global_lambda;
    cfg.penalty = in.get_double("penalty", 0.0);
    cfg.min_stiffness = in.get_double("minimum_stiffness", 0.0);
    cfg.min_density = in.get_double("minimum_density", 0.0);
    cfg.mg_levels = in.mg_level;
    cfg.smooth_iters = in.pre_and_post_smoothing_iter;
    cfg.boundary_iters = in.get_int("boundary_smoothing", 5);
    cfg.device = in.get_int("device", 0);
    // a call factors the bottom once and solves once: the nested-dissection
    // Cholesky factors faster than the dense inverse above about 4 K unknowns
    cfg.bottom_dense_max = kFemSolveDenseBottomDofs;
    // solver state reuse (fem_interface.h:193-209, SPGrid_CPU_Solver.cpp:58-75):
    // a state pointer means a warm start.  The kept hierarchy serves when the
    // blocks, Dirichlet data and settings are the same and no cell switches
    // on or off; otherwise the solver is rebuilt (the reference re-flags
    // every call, :102-145) and the previous u is carried over by block.
    trace("unpack");
    const bool warm = h != nullptr;
    WarmState prev;
    if (h) prev = std::move(h->warm);
    bool rebuilt = true;
    if (h && same_topology(h.get(), in, base) && h->s->active_set_matches(dens.data())) {
      trace("reuse checks");
      h->s->update_density(dens.data());
      trace("update density");
      rebuilt = false;
    } else {
      h.reset();
      spg_solver* nh = nullptr;
      int rc = spg_create(&cfg, (int64_t)nb, base.data(), dens.data(), (int64_t)da.size(), dc.data(), da.data(),
                          dv.data(), &nh);
      if (rc != SPG_OK) throw SpgError(rc, g_err);
      h.reset(nh);
      trace("create solver");
    }
    if (warm && rebuilt && prev.valid) {
      const std::vector<double> g = carry_over(prev, base);
      double* gd = nullptr;
      if (hipMalloc(&gd, g.size() * sizeof(double)) != hipSuccess) throw SpgError(SPG_E_HIP, "hipMalloc failed");
      std::unique_ptr<double, void (*)(double*)> guard(gd, [](double* p) { (void)hipFree(p); });
      if (hipMemcpy(gd, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        throw SpgError(SPG_E_HIP, "hipMemcpy failed");
      h->s->set_initial_guess_in3(gd);
      trace("carry over");
    }
    std::vector<int32_t> fc(in.forces.size() * 3);
    std::vector<double> fv(in.forces.size() * 3);
    for (size_t i = 0; i < in.forces.size(); ++i)
      for (int c = 0; c < 3; ++c) {
        fc[3 * i + c] = in.forces[i].coord[c];
        fv[3 * i + c] = in.forces[i].f[c];
      }
    h->s->set_forces((int64_t)in.forces.size(), fc.data(), fv.data());
    trace("forces");
    spg_solve_params prm{};
    prm.tolerance = in.krylov.tolerance;
    // the reference's CG runs for at most 500 iterations whatever
    // krylov.max_iterations says (SPGrid_CPU_Solver.cpp:326; its caller
    // sends cg_max_iterations, default 50, spgrid_topo_opt.cpp:1000);
    // extra_ints["pcg_max_iterations"] sets another cap per pass
    const int cap = in.get_int("pcg_max_iterations", 0);
    prm.max_iterations = cap > 0 ? cap : 500;
    const int dci = in.get_int("defect_correction_iter", 0);
    prm.max_passes = dci > 0 ? dci : 1;
    prm.precision = in.get_int("fp64_solve", 0) ? 1 : 0;
    prm.inner_tolerance = in.get_double("inner_tolerance", 0.0);
    prm.print_residuals = in.krylov.print_residuals ? 1 : 0;
    // warm start: reuse u if forced or if it beats the stored first residual
    prm.warm_start = warm ? (in.get_int("forced_reuse", 0) ? 2 : 1) : 0;
    prm.reference_linf = warm && prev.valid ? prev.initial_linf : 0.0;
    spg_info info{};
    const int rc = h->s->solve(prm, &info);
    trace("solve");
    out.residuals.push_back(info.r_linf / (info.r0_linf > 0 ? info.r0_linf : 1.0));
    // a solve that ran to its end succeeds, converged or not, as the
    // reference's does (SPGrid_CPU_Solver.cpp:326,368): its caller stops the
    // optimization on success = 0 (objective/minimal_compliance.cpp:64-67).
    // error_code and info still say when the tolerance was not met.
    out.error_code = rc;
    out.success = (rc == SPG_OK || rc == SPG_E_NOT_CONVERGED);
    out.info = "iterations=" + std::to_string(info.iterations) + " passes=" + std::to_string(info.passes) +
               " r0_linf=" + std::to_string(info.r0_linf) + " r_linf=" + std::to_string(info.r_linf) +
               (warm ? (info.warm_reused ? " warm_start=reused" : " warm_start=discarded") : "");
    if (rc == SPG_E_NOT_CONVERGED) out.info += " (tolerance not reached)";
    else if (rc != SPG_OK) out.info += " (failed)";
    if (keep) out.set_state_ptr(h.get());
    // the stream in one allocation: head, the displacement block records
    // (packed on the device and copied into place once), tail
    WireWriter head, tail(false);
    write_outputs_head(head, out, nb);
    write_outputs_tail(tail, out);
    const size_t hb = head.bytes().size(), rb = nb * kRecordBytes, tb = tail.bytes().size();
    std::unique_ptr<char[]> data(new char[hb + rb + tb]);
    uint8_t* recs = reinterpret_cast<uint8_t*>(data.get()) + hb;
    h->s->pack_output_records(base.data(), recs);
    const uint64_t total = hb + rb + tb;
    std::memcpy(data.get(), head.bytes().data(), hb);
    std::memcpy(data.get(), &total, sizeof(total));
    std::memcpy(recs + rb, tail.bytes().data(), tb);
    trace("pack output");
    // what the next warm start begins from: the displacement of a finished
    // solve, and the first solve's zero-guess residual
    if (rc == SPG_OK || rc == SPG_E_NOT_CONVERGED) {
      WarmState& w = h->warm;
      w.base = base;
      w.u.resize(nb * 192);
      parallel_ranges(
          (int64_t)nb,
          [&](int64_t b0, int64_t b1) {
            for (int64_t b = b0; b < b1; ++b)
              for (int l = 0; l < 64; ++l)
                std::memcpy(&w.u[(b * 64 + l) * 3], recs + b * kRecordBytes + 12 + l * 32, 3 * sizeof(double));
          },
          256);
      w.initial_linf = prev.valid ? prev.initial_linf : info.r0_linf;
      w.valid = true;
    } else {
      h->warm = std::move(prev);
    }
    emitted = std::move(data);
    trace("keep warm state");
  } catch (SpgError& e) {
    out.success = false;
    out.error_code = e.code;
    out.info = e.what();
  } catch (LayoutError& e) {
    out.success = false;
    out.error_code = e.code;
    out.info = e.what();
  } catch (WireError& e) {
    out.success = false;
    out.error_code = e.code == -7 ? SPG_E_VERSION : SPG_E_INVALID;
    out.info = e.what();
  } catch (std::exception& e) {
    out.success = false;
    out.error_code = SPG_E_INVALID;
    out.info = e.what();
  } catch (...) {
    out.success = false;
    out.error_code = SPG_E_INVALID;
    out.info = "unknown error";
  }
  if (keep && h) out.set_state_ptr(h.release());
  h.reset();
  trace("release");
  if (emitted) {
    *reinterpret_cast<void**>(output) = emitted.release();
    return;
  }
  try {
    emit(out, output);
  } catch (...) {
    *reinterpret_cast<void**>(output) = nullptr;
  }
  trace("write outputs");
}

// The reference interface test's fake solver (topo_opt/interface_test.cpp:7-34):
// u_p = density + p for every node of every density block.
void fem_solve_test(void* input, void* output) {
  FemOutputs out;
  if (!output) return;
  try {
    FemInputs in;
    WireReader r(input);
    read_inputs(r, in);
    for (int d = 0; d < 3; ++d) out.displacements.resolution[d] = in.resolution[d];
    for (auto& db : in.density.blocks) {
      VectorBlock vb;
      for (int c = 0; c < 3; ++c) vb.base[c] = db.base[c];
      for (int l = 0; l < 64; ++l)
        for (int p = 0; p < 4; ++p) vb.data[l][p] = p < 3 ? db.data[l] + p : 0.0;
      out.displacements.blocks.push_back(vb);
    }
    out.success = true;
  } catch (std::exception& e) {
    out.success = false;
    out.error_code = SPG_E_INVALID;
    out.info = e.what();
  }
  emit(out, output);
}

}  // extern "C"
