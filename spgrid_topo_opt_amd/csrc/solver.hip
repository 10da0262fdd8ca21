// solver.hip -- see solver.hpp.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>

#include "solver.hpp"

namespace spg {

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw SpgError(SPG_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define RBCHK(x)                                                                                           \
  do {                                                                                                     \
    rocblas_status s_ = (x);                                                                               \
    if (s_ != rocblas_status_success)                                                                      \
      throw SpgError(SPG_E_HIP, std::string(#x) + ": rocblas status " + std::to_string((int)s_));          \
  } while (0)

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// boundary phases of explicit levels with at most this many unknowns run as
// one dense map (2 maps x 3n x 6n entries in FP64 and FP32)
static constexpr int kMaxDenseBoundary = 6144;
// pcg() result: a level-0 boundary phase reported a barrier time-out; the
// pass is repeated with the per-colour list launches
static constexpr int kCoopRetry = -100;
// byte offset of the barrier error word in the pinned scalar buffer
static constexpr int kCoopWordOffset = 192;
// relative residual reduction one FP32 MG-PCG pass is asked for at most
static constexpr double kFp32PassReduction = 1e-6;
// bottom solve: the dense FP64 inverse up to this many unknowns (one
// bandwidth-bound GEMV per V-cycle), the nested-dissection Cholesky
// (bottom_nd.hip) above it, unless spg_config.bottom_dense_max sets another
// limit; SPG_BOTTOM=dense|nd overrides both (dense stays limited to
// kMaxDenseBottomDofs: its inverse is O(n^2)).  Per density update the
// dense inverse costs 7, 22, 70 and 192 ms at 1.9 K, 4.6 K, 9 K and 15 K
// unknowns, the nested dissection 8, 19, 37 and 62 ms, which then costs
// about 0.08 ms more per V-cycle at 9 K: the dense inverse pays when many
// solves share a factorization.
static constexpr int64_t kDenseBottomDofs = 16384;
static constexpr int64_t kMaxDenseBottomDofs = 60000;

DBuf::~DBuf() {
  if (p) (void)hipFree(p);
}
void DBuf::alloc(size_t n) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
  bytes = n;
  if (n) HIPCHK(hipMalloc(&p, n));
}

template <class V>
static void upload(DBuf& b, const V& v) {
  using T = typename V::value_type;
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

// the size rule of the colour-major copy of b (Solver::colour_major_b): u,
// b, E and the neighbour table of level 0 in FP32 above 256 MiB
static bool colour_major_by_size(int64_t nb) {
  return nb * (int64_t)(2 * 192 * 4 + 64 * 4 + 27 * 4 + 64) > ((int64_t)256 << 20);
}

template <> float* LevelRT::u<float>() const { return u32.as<float>(); }
template <> double* LevelRT::u<double>() const { return u64.as<double>(); }
template <> float* LevelRT::b<float>() const { return b32.as<float>(); }
template <> double* LevelRT::b<double>() const { return b64.as<double>(); }
template <> float* LevelRT::r<float>() const { return r32.as<float>(); }
template <> double* LevelRT::r<double>() const { return r64.as<double>(); }
template <> const float* LevelRT::st<float>() const { return st32.as<float>(); }
template <> const double* LevelRT::st<double>() const { return st64.as<double>(); }
template <> const float* Solver::E<float>() const { return E32_.as<float>(); }
template <> const double* Solver::E<double>() const { return E64_.as<double>(); }

Solver::Solver(const spg_config& cfg, const ProblemDesc& pd) : cfg_(cfg) {
  const double t0 = now_s();
  if (const char* e = std::getenv("SPG_BOTTOM_FP64")) bottom64_ = std::atoi(e) != 0;
  if (const char* e = std::getenv("SPG_COARSE_FP64")) coarse64_ = std::atoi(e) != 0;
  // flags, hierarchy and the index-range check on the host, before any
  // device call or allocation
  lay_ = build_layout(pd);
  const double t_layout = now_s();
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw SpgError(SPG_E_HIP, "no HIP device available");
  if (cfg.device < 0 || cfg.device >= ndev) throw SpgError(SPG_E_INVALID, "device ordinal out of range");
  HIPCHK(hipSetDevice(cfg.device));
  size_t free0 = 0, total0 = 0;
  HIPCHK(hipMemGetInfo(&free0, &total0));
  HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&st2_, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&ev_a_, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ev_b_, hipEventDisableTiming));
  if (const char* e = std::getenv("SPG_OVERLAP")) overlap_ = std::atoi(e) != 0;
  if (const char* e = std::getenv("SPG_GS_CMB")) cmb_mode_ = std::atoi(e) != 0 ? 1 : 0;
  RBCHK(rocblas_create_handle(&rb_));
  RBCHK(rocblas_set_stream(rb_, st_));
  double kmu[576], kla[576];
  element_tables(kmu, kla);
  for (int e = 0; e < 576; ++e) K0_[e] = cfg.mu * kmu[e] + cfg.lambda * kla[e];
  activate();
  const double t_init = now_s();
  upload_layout();
  sync();
  const double t_upload = now_s();
  build_operators();
  setup_colour_major();
  sync();
  size_t free1 = 0;
  HIPCHK(hipMemGetInfo(&free1, &total0));
  dev_bytes_ = (double)free0 - (double)free1;
  setup_s_ = now_s() - t0;
  if (const char* e = std::getenv("SPG_SETUP_TRACE"); e && std::atoi(e) > 0)
    std::fprintf(stderr, "[spg] setup %.3f s: host layout %.3f, device init %.3f, upload %.3f, operators %.3f\n",
                 setup_s_, t_layout - t0, t_init - t_layout, t_upload - t_init, now_s() - t_upload);
}

Solver::~Solver() {
  if (rb_) rocblas_destroy_handle(rb_);
  if (hscal_) (void)hipHostFree(hscal_);
  if (stage_) (void)hipHostFree(stage_);
  if (ev_a_) (void)hipEventDestroy(ev_a_);
  if (ev_b_) (void)hipEventDestroy(ev_b_);
  if (st2_) (void)hipStreamDestroy(st2_);
  if (st_) (void)hipStreamDestroy(st_);
}

void Solver::sync() { HIPCHK(hipStreamSynchronize(st_)); }

// The element constants live in device constant memory, one set per device,
// shared by every solver on it.  The last material uploaded is recorded per
// device; a solver whose material differs waits for the device (work of the
// other solver may still read the constants) and uploads its own.  Calls
// into solvers of different materials on one device from several threads
// at once are not supported.
void Solver::activate() {
  struct Key {
    bool set = false;
    double mu = 0, lambda = 0;
  };
  static std::mutex mu;
  static std::vector<Key> keys;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)keys.size() <= cfg_.device) keys.resize(cfg_.device + 1);
  Key& k = keys[cfg_.device];
  if (k.set && k.mu == cfg_.mu && k.lambda == cfg_.lambda) return;
  int prev = 0;
  HIPCHK(hipGetDevice(&prev));
  HIPCHK(hipSetDevice(cfg_.device));
  HIPCHK(hipDeviceSynchronize());
  const hipError_t e = upload_k0(K0_);
  HIPCHK(hipSetDevice(prev));
  if (e != hipSuccess) throw SpgError(SPG_E_HIP, std::string("element constants: ") + hipGetErrorString(e));
  k.set = true;
  k.mu = cfg_.mu;
  k.lambda = cfg_.lambda;
}

void Solver::check_coop() {
  if (lv_.empty() || !lv_[0].coop.p) return;
  unsigned w[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(w, lv_[0].coop.p, sizeof(w), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  if (w[1]) {
    // reported once: the next phase starts from a clear word
    HIPCHK(hipMemsetAsync(static_cast<unsigned*>(lv_[0].coop.p) + 1, 0, sizeof(unsigned), st_));
    HIPCHK(hipStreamSynchronize(st_));
    throw SpgError(SPG_E_HIP,
                   "level-0 boundary phase: grid barrier wait timed out (workgroups not resident together; "
                   "SPG_COOP_LAUNCH=1 uses the cooperative launch API, SPG_BOUNDARY_PATH=2 per-colour launches)");
  }
}

void Solver::upload_layout() {
  const int L = lay_.nlev;
  lv_.clear();
  lv_.resize(L);
  for (int l = 0; l < L; ++l) {
    const LevelDesc& D = lay_.lv[l];
    LevelRT& R = lv_[l];
    R.d.nb = (int)D.nb;
    R.d.cell = cell_consts(cfg_.mu, cfg_.lambda);
    R.n3 = D.nb * 192;
    upload(R.nbr, D.nbr);
    upload(R.flags, D.flags);
    upload(R.base, D.base);
    if (!D.restrict_src.empty()) upload(R.rsrc, D.restrict_src);
    if (!D.prolong_src.empty()) upload(R.psrc, D.prolong_src);
    std::vector<int32_t> bnd;
    R.d.bnd_max = 0;
    for (int c = 0; c < 8; ++c) {
      R.d.bnd_off[c] = (int)bnd.size();
      bnd.insert(bnd.end(), D.bnd_list[c].begin(), D.bnd_list[c].end());
      R.d.bnd_max = std::max(R.d.bnd_max, (int)D.bnd_list[c].size());
    }
    R.d.bnd_off[8] = (int)bnd.size();
    if (!bnd.empty()) upload(R.bnd, bnd);
    R.d.bnd = R.bnd.as<int>();
    if (!bnd.empty()) {
      upload(R.bci, D.bnd_ci);
      upload(R.bslot, D.bnd_slot);
      R.bA64.alloc(bnd.size() * 288 * 8);
      R.bA32.alloc(bnd.size() * 288 * 4);
      R.bt64.alloc(bnd.size() * 4 * 8);
      R.bt32.alloc(bnd.size() * 4 * 4);
    }
    if (!D.bnd_cslot.empty()) {
      upload(R.b0cslot, D.bnd_cslot);
      upload(R.b0node, D.bnd_cnode);
      R.b0E64.alloc(D.bnd_cslot.size() * 8);
      R.b0E32.alloc(D.bnd_cslot.size() * 4);
      R.b0D64.alloc(bnd.size() * 8 * 8);
      R.b0D32.alloc(bnd.size() * 8 * 4);
      R.d.b0node = R.b0node.as<int16_t>();
      R.d.b0E64 = R.b0E64.as<double>();
      R.d.b0E32 = R.b0E32.as<float>();
      R.d.b0D64 = R.b0D64.as<double>();
      R.d.b0D32 = R.b0D32.as<float>();
      R.b0ub.alloc((bnd.size() + 1) * 3 * 8);
      R.coop.alloc(2 * sizeof(unsigned));
      HIPCHK(hipMemset(R.coop.p, 0, 2 * sizeof(unsigned)));
      R.d.b0ub = R.b0ub.p;
      R.d.coop = R.coop.as<unsigned>();
    }
    R.d.bci = R.bci.as<int>();
    R.d.bslot = R.bslot.as<int>();
    R.d.bA64 = R.bA64.as<double>();
    R.d.bA32 = R.bA32.as<float>();
    R.d.bt64 = R.bt64.as<double>();
    R.d.bt32 = R.bt32.as<float>();
    R.d.nbr = R.nbr.as<int>();
    if (l == 0 && !bnd.empty()) {
      R.bnslot.alloc(bnd.size() * 27 * 4);
      launch_bnd_nslot(st_, R.d, R.bnslot.as<int>());
      R.d.bnslot = R.bnslot.as<int>();
    }
    if (l == 0 && !bnd.empty() && !D.gs_dist.empty()) {
      // near-band step lists of both sweep directions, and their neighbour slots
      std::vector<int32_t> all;
      for (int i = 0; i < 16; ++i) {
        const std::vector<int32_t>& L = D.gs_near[i / 8][i % 8];
        R.gsnear_off[i] = (int64_t)all.size();
        if (!L.empty()) {
          // the layout's step colours must be the sweep's
          const int32_t s0 = L[0] & 63;
          const int colour = (((s0 >> 4) & 1) << 2) | (((s0 >> 2) & 1) << 1) | (s0 & 1);
          if (colour != sweep_colour(true, i < 8, i % 8))
            throw SpgError(SPG_E_INVALID, "near-band step list in the wrong colour order");
        }
        all.insert(all.end(), L.begin(), L.end());
      }
      R.gsnear_off[16] = (int64_t)all.size();
      upload(R.gsdist, D.gs_dist);
      upload(R.gsnear, all);
      R.gsnslot.alloc(std::max<size_t>(all.size(), 1) * 27 * 4);
      launch_list_nslot(st_, R.d, (int)all.size(), R.gsnear.as<int>(), R.gsnslot.as<int>());
    }
    R.d.flags = R.flags.as<uint8_t>();
    R.d.base = R.base.as<int>();
    R.d.rsrc = R.rsrc.as<int>();
    R.d.psrc = R.psrc.as<int>();
    if (R.psrc.p) {
      R.psrc27.alloc((size_t)D.nb * 27 * 4);
      launch_prolong_table(st_, R.d, R.psrc27.as<int>());
      R.d.psrc27 = R.psrc27.as<int>();
    }
    R.u32.alloc(R.n3 * 4);
    R.b32.alloc(R.n3 * 4);
    R.r32.alloc(R.n3 * 4);
    R.u64.alloc(R.n3 * 8);
    R.b64.alloc(R.n3 * 8);
    R.r64.alloc(R.n3 * 8);
  }
  // level-0 cell data
  upload(E64_, lay_.E);
  E32_.alloc(lay_.E.size() * 4);  // rounded on the device (round to nearest, as a host cast)
  launch_cast_d2f(st_, E64_.as<double>(), E32_.as<float>(), (int64_t)lay_.E.size());
  upload(rho_, lay_.rho);
  upload(in_slot_, lay_.in_node_slot);
  upload(ud_idx_, lay_.ud_slot);
  upload(ud_val_, lay_.ud_value);
  n_ud_ = (int64_t)lay_.ud_slot.size();
  const int64_t n3 = lv_[0].n3;
  for (DBuf* b : {&x32_, &rr32_, &s32_, &q32_, &z32_}) b->alloc(n3 * 4);
  for (DBuf* b : {&x64_, &rr64_, &s64_, &q64_, &z64_, &u_, &f_, &res_, &tmp_}) b->alloc(n3 * 8);
  scal_.alloc(sizeof(CgScalars) + reduce_partials(0) * sizeof(double));
  part_.alloc((reduce_partials(0) + 16) * sizeof(double));
  apart_.alloc(apply_dot_partials(lv_[0].d.nb) * sizeof(double));
  HIPCHK(hipHostMalloc((void**)&hscal_, 256));
}

// Galerkin position matrices of one depth for this solver's K0, on its
// device: built and uploaded once per process and kept until it ends (the
// deepest usual one, depth 5, is 576 x 8^5 doubles = 151 MB), so solvers
// rebuilt for new block sets or densities with the same material reuse them
const double* Solver::posmat_device(int depth) {
  struct Entry {
    int dev, depth;
    std::vector<double> k0;
    double* ptr;
  };
  static std::mutex mu;
  static std::vector<Entry>* cache = new std::vector<Entry>();  // never freed: device memory held to the end
  const std::vector<double> k0(K0_, K0_ + 576);
  std::lock_guard<std::mutex> lock(mu);
  for (const Entry& e : *cache)
    if (e.dev == cfg_.device && e.depth == depth && e.k0 == k0) return e.ptr;
  const std::vector<double> M = galerkin_position_matrices(depth, K0_);
  double* p = nullptr;
  HIPCHK(hipMalloc((void**)&p, M.size() * sizeof(double)));
  HIPCHK(hipMemcpy(p, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
  cache->push_back(Entry{cfg_.device, depth, k0, p});
  return p;
}

void Solver::build_operators() {
  const int L = lay_.nlev;
  if (!lay_.use_mg) return;
  const char* tr = std::getenv("SPG_SETUP_TRACE");
  const bool trace = tr && std::atoi(tr) >= 2;
  double tm = now_s();
  auto mark = [&](const char* what) {
    if (!trace) return;
    sync();
    const double t = now_s();
    std::fprintf(stderr, "[spg] operators %-22s %.3f s\n", what, t - tm);
    tm = t;
  };
  const double one = 1.0, zero = 0.0;
  for (int l = 0; l < L; ++l) {
    const LevelDesc& D = lay_.lv[l];
    if (!D.is_explicit) continue;
    LevelRT& R = lv_[l];
    const int depth = (L == 1) ? 0 : l;
    const int64_t npos = (int64_t)1 << (3 * depth);
    upload(R.cell_row, D.cell_row);
    if (!D.cell_fine_blocks.empty()) upload(R.cfb, D.cell_fine_blocks);
    upload(R.cgi, D.cell_gpu_index);
    const double* dM = posmat_device(depth);
    DBuf G, Cm;
    const int64_t ncell = D.ncell;
    R.st64.alloc((size_t)D.nb * 243 * 64 * 8);
    R.st32.alloc((size_t)D.nb * 243 * 64 * 4);
    if (ncell > 0) {
      G.alloc((size_t)(ncell * npos) * 8);
      Cm.alloc((size_t)ncell * 576 * 8);
      launch_galerkin_gather(st_, ncell, depth, R.cfb.as<int>(), R.cgi.as<int>(), E64_.as<double>(), G.as<double>());
      // Cm^T (576 x ncell) = M^T (576 x npos) * G^T (npos x ncell), column-major views
      RBCHK(rocblas_dgemm(rb_, rocblas_operation_none, rocblas_operation_none, 576, (rocblas_int)ncell,
                          (rocblas_int)npos, &one, dM, 576, G.as<double>(), (rocblas_int)npos, &zero,
                          Cm.as<double>(), 576));
    }
    mark("  gather + product");
    launch_galerkin_assemble(st_, R.d, R.cell_row.as<int>(), Cm.as<double>(), cfg_.dx, R.st64.as<double>());
    launch_stencil_to_float(st_, R.st64.as<double>(), R.st32.as<float>(), (int64_t)D.nb * 243 * 64);
    sync();
    mark("  assemble");
  }
  mark("Galerkin levels");
  build_boundary_rows();
  mark("boundary rows");
  // bottom: FP64 Cholesky of the Dirichlet-deleted stencil matrix, then its inverse
  // (Multigrid_Hierarchy.h:332-341; Elasticity_Matrix.cpp:15-233 factorizes with PARDISO)
  const int B = L - 1;
  const LevelDesc& D = lay_.lv[B];
  std::vector<int32_t> var(D.nb * 64, -1);
  int64_t nv = 0;
  for (int64_t s = 0; s < D.nb * 64; ++s)
    if (D.flags[s] & D_ACTIVE) {
      var[s] = (int32_t)nv;
      nv += 3;
    }
  nvar_ = nv;
  if (trace) std::fprintf(stderr, "[spg] operators bottom unknowns        %lld\n", (long long)nv);
  const char* bsel = std::getenv("SPG_BOTTOM");
  const int64_t dense_max =
      std::min(cfg_.bottom_dense_max > 0 ? (int64_t)cfg_.bottom_dense_max : kDenseBottomDofs, kMaxDenseBottomDofs);
  const bool want_nd = bsel ? std::string(bsel) == "nd" : nv > dense_max;
  if (!want_nd && nv > kMaxDenseBottomDofs)
    throw SpgError(SPG_E_INVALID, "bottom level has " + std::to_string(nv) + " unknowns, more than the dense "
                                      "inverse takes (" + std::to_string(kMaxDenseBottomDofs) + "); SPG_BOTTOM=dense "
                                      "asked for it");
  if (want_nd) {
    // nested-dissection Cholesky: symbolic once per topology, numeric for
    // every new set of stencils
    try {
      if (!nd_) {
        nd_ = std::make_unique<NdBottom>(D.nb, D.base.data(), D.nbr.data(), D.flags.data(), D_ACTIVE);
        nd_->upload(st_);
      }
      nd_->factor(lv_[B].st64.as<double>(), st_, rb_);
      mark("bottom (nested dissection)");
    } catch (const NdNotSpd& e) {
      throw SpgError(SPG_E_SINGULAR, e.what());
    } catch (const SpgError&) {
      throw;
    } catch (const std::exception& e) {
      throw SpgError(SPG_E_HIP, std::string("nested-dissection bottom: ") + e.what());
    }
    return;
  }
  nd_.reset();
  upload(var_of_, var);
  DBuf Ainv;
  Ainv.alloc((size_t)nv * nv * 8);
  info_.alloc(16);
  HIPCHK(hipMemsetAsync(Ainv.p, 0, (size_t)nv * nv * 8, st_));
  launch_dense_assemble(st_, lv_[B].d, lv_[B].st64.as<double>(), var_of_.as<int>(), nv, Ainv.as<double>());
  if (nv * nv <= (int64_t)std::numeric_limits<int32_t>::max()) {
    RBCHK(rocsolver_dpotrf(rb_, rocblas_fill_lower, (rocblas_int)nv, Ainv.as<double>(), (rocblas_int)nv,
                           info_.as<rocblas_int>()));
    int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, info_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
    sync();
    if (hinfo != 0)
      throw SpgError(SPG_E_SINGULAR, "bottom Cholesky failed: leading minor " + std::to_string(hinfo) +
                                         " not positive definite");
    RBCHK(rocsolver_dpotri(rb_, rocblas_fill_lower, (rocblas_int)nv, Ainv.as<double>(), (rocblas_int)nv,
                           info_.as<rocblas_int>()));
  } else {
    // more than 2^31 entries: the 32-bit interfaces overflow their offsets, so
    // factor with the 64-bit one and form A^-1 = L^-T L^-1 = X^T X, X = L^-1
    RBCHK(rocsolver_dpotrf_64(rb_, rocblas_fill_lower, nv, Ainv.as<double>(), nv, info_.as<int64_t>()));
    int64_t hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, info_.p, sizeof(int64_t), hipMemcpyDeviceToHost, st_));
    sync();
    if (hinfo != 0)
      throw SpgError(SPG_E_SINGULAR, "bottom Cholesky failed: leading minor " + std::to_string(hinfo) +
                                         " not positive definite");
    DBuf X;
    X.alloc((size_t)nv * nv * 8);
    HIPCHK(hipMemsetAsync(X.p, 0, (size_t)nv * nv * 8, st_));
    launch_set_diagonal(st_, X.as<double>(), nv, 1.0);
    RBCHK(rocblas_dtrsm_64(rb_, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                           rocblas_diagonal_non_unit, nv, nv, &one, Ainv.as<double>(), nv, X.as<double>(), nv));
    RBCHK(rocblas_dsyrk_64(rb_, rocblas_fill_lower, rocblas_operation_transpose, nv, nv, &one, X.as<double>(), nv,
                           &zero, Ainv.as<double>(), nv));
    sync();
  }
  launch_symmetrize_lower(st_, Ainv.as<double>(), nv);
  // packed lower-triangle tiles (FP64 for the FP64 cycle, rounded to FP32 for the FP32 cycle)
  nT_ = (int)((nv + 63) / 64);
  const size_t ntile = (size_t)nT_ * (nT_ + 1) / 2;
  const size_t npad = (size_t)nT_ * 64;
  tiles64_.alloc(ntile * 4096 * 8);
  tiles32_.alloc(ntile * 4096 * 4);
  launch_pack_sym_tiles<double>(st_, Ainv.as<double>(), nv, nT_, tiles64_.as<double>());
  launch_pack_sym_tiles<float>(st_, Ainv.as<double>(), nv, nT_, tiles32_.as<float>());
  for (DBuf* d : {&xb64_, &yb64_}) d->alloc(npad * 8);
  for (DBuf* d : {&xb32_, &yb32_}) d->alloc(npad * 4);
  pb64_.alloc((size_t)symv_splits(nT_) * npad * 8);
  pb32_.alloc((size_t)symv_splits(nT_) * npad * 4);
  HIPCHK(hipMemsetAsync(xb64_.p, 0, npad * 8, st_));
  HIPCHK(hipMemsetAsync(xb32_.p, 0, npad * 4, st_));
  sync();
  mark("bottom (dense inverse)");
}

// coefficient rows of the boundary system of every smoothed level
void Solver::build_boundary_rows() {
  for (int l = 0; l < lay_.nlev; ++l) {
    LevelRT& R = lv_[l];
    const int n = R.d.bnd_off[8];
    if (n == 0 || !R.bA64.p) continue;
    if (l == 0 && R.b0cslot.p) {
      launch_bnd0_setup(st_, R.d, R.b0cslot.as<int>(), E64_.as<double>(), cfg_.dx, R.b0E64.as<double>(),
                        R.b0D64.as<double>());
      launch_cast_d2f(st_, R.b0E64.as<double>(), R.b0E32.as<float>(), (int64_t)n * 8);
      launch_cast_d2f(st_, R.b0D64.as<double>(), R.b0D32.as<float>(), (int64_t)n * 8);
    }
    if (l == 0) launch_bnd_coef0(st_, R.d, E64_.as<double>(), cfg_.dx, R.bA64.as<double>());
    else if (lay_.lv[l].is_explicit) launch_bnd_coef_st(st_, R.d, R.st64.as<double>(), R.bA64.as<double>());
    else continue;
    launch_cast_d2f(st_, R.bA64.as<double>(), R.bA32.as<float>(), (int64_t)n * 288);
    // small boundary sets of explicit levels: each phase becomes one dense map
    if (l > 0 && lay_.lv[l].is_explicit && 3 * n <= kMaxDenseBoundary && cfg_.boundary_iters > 0) {
      const size_t entries = (size_t)3 * n * bmap_stride(n);
      for (int d = 0; d < 2; ++d) {
        if (!R.bmap64[d].p) {
          // padding columns stay zero: the build writes columns < 6n only
          R.bmap64[d].alloc(entries * 8);
          R.bmap32[d].alloc(entries * 4);
          HIPCHK(hipMemsetAsync(R.bmap64[d].p, 0, entries * 8, st_));
        }
        launch_bmap_build(st_, R.d, cfg_.boundary_iters, d == 0, R.bmap64[d].as<double>());
        launch_cast_d2f(st_, R.bmap64[d].as<double>(), R.bmap32[d].as<float>(), (int64_t)entries);
        R.d.bmap64[d] = R.bmap64[d].as<double>();
        R.d.bmap32[d] = R.bmap32[d].as<float>();
      }
      if (!R.bmx64.p) {
        const size_t nx = (size_t)bmap_stride(n) + 3 * n;
        R.bmx64.alloc(nx * 8);
        R.bmx32.alloc(nx * 4);
        HIPCHK(hipMemsetAsync(R.bmx64.p, 0, nx * 8, st_));
        HIPCHK(hipMemsetAsync(R.bmx32.p, 0, nx * 4, st_));
      }
      R.d.bmx64 = R.bmx64.as<double>();
      R.d.bmx32 = R.bmx32.as<float>();
      R.d.bmap_sweeps = cfg_.boundary_iters;
    }
  }
  sync();
}

// The lowest i in [0, n) with ev(i) != 0, and ev there (n and 0 if none):
// the answer of one pass in order, found on up to 16 threads
template <class F>
static std::pair<int64_t, int> first_event(int64_t n, F&& ev) {
  std::mutex m;
  int64_t at = n;
  int kind = 0;
  parallel_ranges(
      n,
      [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i)
          if (const int k = ev(i)) {
            std::lock_guard<std::mutex> g(m);
            if (i < at) {
              at = i;
              kind = k;
            }
            return;
          }
      },
      1 << 15);
  return {at, kind};
}

bool Solver::active_set_matches(const double* density) const {
  const Layout& Lc = lay_;
  // 1: NaN (throws), 2: a cell would switch on or off
  const int kind = first_event(Lc.nin_blocks * 64, [&](int64_t i) {
    const int32_t s = Lc.in_node_slot[i];
    const double d = density[i];
    if (std::isnan(d)) return 1;
    // a cell whose min corner has no GPU slot is inactive; it may not turn on
    if (s < 0) return d > 0 ? 2 : 0;
    return (Lc.E[s] != 0.0) != (d > 0) ? 2 : 0;
  }).second;
  if (kind == 1) throw SpgError(SPG_E_NAN, "density contains NaN");
  return kind == 0;
}

void Solver::update_density(const double* density) {
  activate();
  // recompute E (SIMP) for the existing cells, then the coarse operators;
  // nothing changes when the check fails
  const int64_t nin = lay_.nin_blocks * 64;
  const int kind = first_event(nin, [&](int64_t i) {
    const int32_t s = lay_.in_node_slot[i];
    if (s < 0) return 0;
    const double d = density[i];
    if (std::isnan(d)) return 1;
    if (lay_.E[s] != 0.0) return d > 0 ? 0 : 2;
    return d > 0 ? 3 : 0;
  }).second;
  if (kind == 1) throw SpgError(SPG_E_NAN, "density contains NaN");
  if (kind == 2) throw SpgError(SPG_E_INVALID, "density update would deactivate a cell (topology change)");
  if (kind == 3) throw SpgError(SPG_E_INVALID, "density update would activate a cell (topology change)");
  // input nodes have distinct slots: the ranges write disjoint entries
  parallel_ranges(
      nin,
      [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
          const int32_t s = lay_.in_node_slot[i];
          if (s < 0) continue;
          const double d = density[i];
          lay_.rho[s] = d;
          if (lay_.E[s] != 0.0) lay_.E[s] = stiffness_mapping(cfg_.penalty, cfg_.min_stiffness, cfg_.min_density, d);
        }
      },
      1 << 15);
  // both arrays through the kept pinned buffer, filled on host threads: a
  // pageable copy runs at a few GB/s on one thread
  const size_t ne = lay_.E.size() * 8, nr = lay_.rho.size() * 8;
  HIPCHK(hipStreamSynchronize(st_));  // the buffer may still feed an earlier copy
  if (stage_bytes_ < ne + nr) {
    if (stage_) (void)hipHostFree(stage_);
    stage_ = nullptr;
    stage_bytes_ = 0;
    HIPCHK(hipHostMalloc(&stage_, ne + nr));
    stage_bytes_ = ne + nr;
  }
  uint8_t* hs = static_cast<uint8_t*>(stage_);
  const uint8_t* srcE = reinterpret_cast<const uint8_t*>(lay_.E.data());
  const uint8_t* srcR = reinterpret_cast<const uint8_t*>(lay_.rho.data());
  parallel_ranges(
      (int64_t)(ne + nr),
      [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b;) {
          const int64_t end = i < (int64_t)ne ? std::min<int64_t>(b, ne) : b;
          std::memcpy(hs + i, i < (int64_t)ne ? srcE + i : srcR + (i - ne), (size_t)(end - i));
          i = end;
        }
      },
      1 << 22);
  HIPCHK(hipMemcpyAsync(E64_.p, hs, ne, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(rho_.p, hs + ne, nr, hipMemcpyHostToDevice, st_));
  launch_cast_d2f(st_, E64_.as<double>(), E32_.as<float>(), (int64_t)lay_.E.size());
  build_operators();
  // the pinned buffer must not change before the copies have read it
  HIPCHK(hipStreamSynchronize(st_));
}

void Solver::update_density_device(const double* density_dev) {
  activate();
  // same checks and SIMP map as update_density, on the device; nothing
  // changes when the check fails
  const int64_t nin = lay_.nin_blocks * 64;
  if (!derr_.p) derr_.alloc(16);
  int* err = derr_.as<int>();
  HIPCHK(hipMemsetAsync(err, 0, sizeof(int), st_));
  launch_density_check(st_, nin, in_slot_.as<int>(), density_dev, E64_.as<double>(), err);
  int herr = 0;
  HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st_));
  sync();
  if (herr & 1) throw SpgError(SPG_E_NAN, "density contains NaN");
  if (herr & 2) throw SpgError(SPG_E_INVALID, "density update would deactivate a cell (topology change)");
  if (herr & 4) throw SpgError(SPG_E_INVALID, "density update would activate a cell (topology change)");
  launch_density_update(st_, nin, in_slot_.as<int>(), density_dev, cfg_.penalty, cfg_.min_stiffness,
                        cfg_.min_density, E64_.as<double>(), E32_.as<float>(), rho_.as<double>());
  build_operators();
}

void Solver::set_forces(int64_t n, const int32_t* coord, const double* value) {
  // ForceOnNode assignment (SPGrid_CPU_Solver.cpp:289-296): later entries win
  if (gmap0_.empty()) {
    const LevelDesc& D = lay_.lv[0];
    gmap0_.reserve(D.nb * 2);
    for (int64_t b = 0; b < D.nb; ++b)
      gmap0_.emplace(bkey(D.base[3 * b] >> 2, D.base[3 * b + 1] >> 2, D.base[3 * b + 2] >> 2), (int32_t)b);
  }
  std::vector<int32_t> idx;
  std::vector<double> val;
  std::unordered_map<int32_t, size_t> at;  // node slot -> its entry: a repeated node keeps the last value
  for (int64_t i = 0; i < n; ++i) {
    const int x = coord[3 * i], y = coord[3 * i + 1], z = coord[3 * i + 2];
    auto it = gmap0_.find(bkey(x >> 2, y >> 2, z >> 2));
    if (it == gmap0_.end()) continue;  // forces outside the solved blocks are ignored
    const int32_t g = it->second;
    const int q = ((x & 3) << 4) | ((y & 3) << 2) | (z & 3);
    const int32_t slot = g * 64 + q;
    auto [pos, fresh] = at.emplace(slot, idx.size());
    if (!fresh) {
      for (int c = 0; c < 3; ++c) val[pos->second + c] = value[3 * i + c];
      continue;
    }
    for (int c = 0; c < 3; ++c) {
      idx.push_back(slot * 3 + c);
      val.push_back(value[3 * i + c]);
    }
  }
  upload(f_idx_, idx);
  upload(f_val_, val);
  n_f_ = (int64_t)idx.size();
}

double Solver::read_scalar(const double* dev) {
  HIPCHK(hipMemcpyAsync(hscal_, dev, sizeof(double), hipMemcpyDeviceToHost, st_));
  sync();
  return hscal_[0];
}

// ------------------------------------------------------------------ V-cycle
// Per block, a colour launch reads the 5^3 halo of u and E (mostly shared
// with the neighbouring blocks' halos) but, in node-major order, whole cache
// lines of b and the flags for the 8 nodes of its colour.  A colour-major
// copy of b, made once per V-cycle, saves 7/8 of those bytes in each of the
// 16 colour launches.  Measured per V-cycle (FP32 / FP64): the 256^3 shell
// (82,765 blocks) 1.605 / 2.405 ms node-major against 1.675 / 2.426 ms with
// the copy; the 512x256x256 band (245,910 blocks) 4.254 / 5.937 ms against
// 4.100 / 5.809 ms.  By default the copy is used above 256 MiB of u, b, E and
// neighbour table in FP32, i.e. above about 137 K blocks.
// The level-0 flags in colour-major order and room for b in that order,
// allocated after everything the solve needs: without the memory for them
// the colour sweeps keep reading node-major.
void Solver::setup_colour_major() {
  if (!(cmb_mode_ == 1 || (cmb_mode_ < 0 && colour_major_by_size(lv_[0].d.nb)))) return;
  const LevelRT& R0 = lv_[0];
  try {
    fcm_.alloc((size_t)R0.d.nb * 64);
    bcm_.alloc((size_t)R0.n3 * 8);
  } catch (const SpgError&) {
    (void)hipGetLastError();
    fcm_.alloc(0);
    bcm_.alloc(0);
    return;
  }
  launch_colour_major_flags(st_, R0.d.nb, R0.d.flags, fcm_.as<uint8_t>());
  lv_[0].d.flags_cm = fcm_.as<uint8_t>();
}

template <class T>
const T* Solver::colour_major_b(const T* b0) {
  if (!lv_[0].d.flags_cm) return nullptr;
  if (cmb_mode_ < 0 && !colour_major_by_size(lv_[0].d.nb)) return nullptr;
  T* bcm = reinterpret_cast<T*>(bcm_.p);
  launch_colour_major<T>(st_, lv_[0].d, b0, bcm);
  return bcm;
}

template <class T>
void Solver::smooth(int l, T* u, const T* b, bool forward, const T* bcm) {
  const LevelRT& R = lv_[l];
  // Multigrid_Hierarchy.h:577-586 / 605-617: boundary sweeps, full sweeps, boundary sweeps
  auto boundary = [&] {
    if (l == 0) launch_gs0_boundary<T>(st_, R.d, E<T>(), dxT<T>(), u, b, cfg_.boundary_iters, forward);
    else launch_st_boundary<T>(st_, R.d, R.st<T>(), u, b, cfg_.boundary_iters, forward);
  };
  int first = 0;
  if (l == 0 && overlap_ && R.gsdist.p && cfg_.smooth_iters > 0 && cfg_.boundary_iters > 0) {
    // The boundary phase writes only band nodes.  The first sweep's step-c
    // update of a node farther than c + 1 from the band reads nothing that
    // phase writes, directly or through earlier steps (Layout::gs_near), so
    // those updates run on st2_ beside the phase; the near ones follow on
    // st_ once both are done, step by step as the sequential sweep orders
    // them.  The result is the sequential sweep's.  256^3 shell, FP32
    // V-cycle: 1.71 ms against 1.62 ms sequential (the near-band list
    // launches cost what the overlap saves), hence off by default.
    HIPCHK(hipEventRecord(ev_a_, st_));
    boundary();
    HIPCHK(hipStreamWaitEvent(st2_, ev_a_, 0));
    for (int c = 0; c < 8; ++c)
      launch_gs0<T>(st2_, R.d, E<T>(), dxT<T>(), u, b, sweep_colour(true, forward, c), R.gsdist.as<uint8_t>(), c + 2);
    HIPCHK(hipEventRecord(ev_b_, st2_));
    HIPCHK(hipStreamWaitEvent(st_, ev_b_, 0));
    const int dir = forward ? 0 : 1;
    for (int c = 0; c < 8; ++c) {
      const int64_t o = R.gsnear_off[dir * 8 + c], n = R.gsnear_off[dir * 8 + c + 1] - o;
      launch_gs0_list<T>(st_, R.d, R.gsnear.as<int>() + o, R.gsnslot.as<int>() + o * 27, (int)n, E<T>(), dxT<T>(), u,
                         b);
    }
    first = 1;
  } else {
    boundary();
  }
  for (int i = first; i < cfg_.smooth_iters; ++i)
    for (int c = 0; c < 8; ++c) {
      if (l == 0) launch_gs0<T>(st_, R.d, E<T>(), dxT<T>(), u, b, sweep_colour(true, forward, c), nullptr, 0, bcm);
      else launch_st_gs<T>(st_, R.d, R.st<T>(), u, b, sweep_colour(false, forward, c));
    }
  boundary();
}

template <class T>
void Solver::level_residual(int l, const T* u, const T* b, T* r) {
  if (l == 0) launch_apply<T>(st_, lv_[0].d, E<T>(), dxT<T>(), u, b, r, 1);
  else launch_st_residual<T>(st_, lv_[l].d, lv_[l].st<T>(), u, b, r);
}

template <class T>
void Solver::bottom(const T* b, T* u) {
  const LevelRT& R = lv_[lay_.nlev - 1];
  if (nd_) {
    // FP64 factors in both cycles: the FP32 cycle's right-hand side is
    // widened and the solution rounded, as the reference casts the PARDISO
    // result to float (Multigrid_Hierarchy.h:346-372, 439-451)
    nd_->solve<T>(R.d.nb, b, u, st_);
    return;
  }
  const bool f = sizeof(T) == 4;
  T* x = f ? (T*)xb32_.p : (T*)xb64_.p;
  T* y = f ? (T*)yb32_.p : (T*)yb64_.p;
  launch_bottom_gather<T>(st_, R.d, var_of_.as<int>(), b, x);
  if (f && bottom64_) {
    // FP32 cycle with the FP64 inverse: widen, apply, round back
    const int64_t npad = (int64_t)nT_ * 64;
    launch_cast_f2d(st_, (const float*)x, xb64_.as<double>(), npad);
    launch_symv<double>(st_, nT_, tiles64_.as<double>(), xb64_.as<double>(), pb64_.as<double>(), yb64_.as<double>());
    launch_cast_d2f(st_, yb64_.as<double>(), (float*)y, npad);
  } else {
    launch_symv<T>(st_, nT_, f ? (const T*)tiles32_.p : (const T*)tiles64_.p, x, f ? (T*)pb32_.p : (T*)pb64_.p, y);
  }
  launch_bottom_scatter<T>(st_, R.d, var_of_.as<int>(), y, b, u);
}

// FP32 cycle whose explicit levels (and bottom) run in FP64: the level
// handed over is restricted in FP32 and widened, and widened corrections are
// rounded back before the FP32 prolongation
void Solver::vcycle_coarse64(const float* b0, float* z, bool z_is_zero) {
  const int L = lay_.nlev;
  const int ex = lay_.explicit_level;
  if (!z_is_zero) launch_zero<float>(st_, z, lv_[0].n3);
  const float* bcm = colour_major_b<float>(b0);
  for (int l = 0; l < L - 1; ++l) {
    const bool d = l >= ex, dn = l + 1 >= ex;
    const bool sm = (l == 0 || l >= ex);
    if (d) {
      double* u = lv_[l].u<double>();
      const double* b = lv_[l].b<double>();
      smooth<double>(l, u, b, true);
      launch_zero<double>(st_, lv_[l + 1].u<double>(), lv_[l + 1].n3);
      level_residual<double>(l, u, b, lv_[l].r<double>());
      launch_restrict<double>(st_, lv_[l + 1].d, lv_[l].r<double>(), lv_[l + 1].b<double>());
    } else {
      float* u = l == 0 ? z : lv_[l].u<float>();
      const float* b = l == 0 ? b0 : lv_[l].b<float>();
      if (sm) smooth<float>(l, u, b, true, l == 0 ? bcm : nullptr);
      if (dn) launch_zero<double>(st_, lv_[l + 1].u<double>(), lv_[l + 1].n3);
      else launch_zero<float>(st_, lv_[l + 1].u<float>(), lv_[l + 1].n3);
      if (sm) {
        level_residual<float>(l, u, b, lv_[l].r<float>());
        launch_restrict<float>(st_, lv_[l + 1].d, lv_[l].r<float>(), lv_[l + 1].b<float>());
      } else {
        launch_restrict<float>(st_, lv_[l + 1].d, b, lv_[l + 1].b<float>());
      }
      if (dn) launch_cast_f2d(st_, lv_[l + 1].b<float>(), lv_[l + 1].b<double>(), lv_[l + 1].n3);
    }
  }
  if (L - 1 >= ex) bottom<double>(lv_[L - 1].b<double>(), lv_[L - 1].u<double>());
  else bottom<float>(lv_[L - 1].b<float>(), lv_[L - 1].u<float>());
  for (int l = L - 2; l >= 0; --l) {
    const bool d = l >= ex, dn = l + 1 >= ex;
    if (d) {
      launch_prolongate<double>(st_, lv_[l].d, lv_[l + 1].u<double>(), lv_[l].u<double>());
      smooth<double>(l, lv_[l].u<double>(), lv_[l].b<double>(), false);
    } else {
      float* u = l == 0 ? z : lv_[l].u<float>();
      const float* b = l == 0 ? b0 : lv_[l].b<float>();
      if (dn) launch_cast_d2f(st_, lv_[l + 1].u<double>(), lv_[l + 1].u<float>(), lv_[l + 1].n3);
      launch_prolongate<float>(st_, lv_[l].d, lv_[l + 1].u<float>(), u);
      if (l == 0 || l >= ex) smooth<float>(l, u, b, false, l == 0 ? bcm : nullptr);
    }
  }
}

template <class T>
void Solver::vcycle(const T* b0, T* z, bool z_is_zero) {
  const int L = lay_.nlev;
  const int ex = lay_.explicit_level;
  if (sizeof(T) == 4 && coarse64_ && lay_.use_mg && L > ex) {
    vcycle_coarse64(reinterpret_cast<const float*>(b0), reinterpret_cast<float*>(z), z_is_zero);
    return;
  }
  if (!lay_.use_mg) {
    launch_copy<T>(st_, b0, z, lv_[0].n3);
    return;
  }
  if (L == 1) {
    bottom<T>(b0, z);
    return;
  }
  if (!z_is_zero) launch_zero<T>(st_, z, lv_[0].n3);
  const T* bcm = colour_major_b<T>(b0);
  for (int l = 0; l < L - 1; ++l) {
    T* u = l == 0 ? z : lv_[l].u<T>();
    const T* b = l == 0 ? b0 : lv_[l].b<T>();
    const bool sm = (l == 0 || l >= ex);
    if (sm) smooth<T>(l, u, b, true, l == 0 ? bcm : nullptr);
    // the restriction also clears the coarse iterate
    if (sm) {
      level_residual<T>(l, u, b, lv_[l].r<T>());
      launch_restrict<T>(st_, lv_[l + 1].d, lv_[l].r<T>(), lv_[l + 1].b<T>(), lv_[l + 1].u<T>());
    } else {
      launch_restrict<T>(st_, lv_[l + 1].d, b, lv_[l + 1].b<T>(), lv_[l + 1].u<T>());
    }
  }
  bottom<T>(lv_[L - 1].b<T>(), lv_[L - 1].u<T>());
  for (int l = L - 2; l >= 0; --l) {
    T* u = l == 0 ? z : lv_[l].u<T>();
    const T* b = l == 0 ? b0 : lv_[l].b<T>();
    launch_prolongate<T>(st_, lv_[l].d, lv_[l + 1].u<T>(), u);
    if (l == 0 || l >= ex) smooth<T>(l, u, b, false, l == 0 ? bcm : nullptr);
  }
}

// ---------------------------------------------------------------------- PCG
// CONJUGATE_GRADIENT_4Ch::Solve (CG/CONJUGATE_GRADIENT.cpp:21-61): projected
// PCG on x (starting at 0) with the L_inf stop test; scalars stay FP64 on the
// device, the host reads ||r||_inf once per iteration.
template <class T>
int Solver::pcg(T* x, T* r, double tol, int maxit, int* iters, double* last_nrm) {
  const LevelRT& R0 = lv_[0];
  const int64_t n = R0.n3;
  T* s = sizeof(T) == 4 ? (T*)s32_.p : (T*)s64_.p;
  T* q = sizeof(T) == 4 ? (T*)q32_.p : (T*)q64_.p;
  T* z = sizeof(T) == 4 ? (T*)z32_.p : (T*)z64_.p;
  CgScalars* sc = scal_.as<CgScalars>();
  double* part = part_.as<double>();
  launch_zero<T>(st_, x, n);
  launch_zero<T>(st_, s, n);
  launch_zero<T>(st_, z, n);
  launch_project<T>(st_, R0.d, r, true);
  launch_cg_scalars(st_, sc, 3);
  launch_linf<T>(st_, r, n, part, &sc->nrm);
  int it = 0;
  int rc = SPG_E_NOT_CONVERGED;
  unsigned* coop_word = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(hscal_) + kCoopWordOffset);
  const bool coop = lv_[0].coop.p && !lv_[0].d.coop_off;
  for (;; ++it) {
    HIPCHK(hipMemcpyAsync(hscal_, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st_));
    if (coop) HIPCHK(hipMemcpyAsync(coop_word, lv_[0].coop.as<unsigned>() + 1, sizeof(unsigned),
                                    hipMemcpyDeviceToHost, st_));
    sync();
    if (coop && *coop_word) {
      // the last V-cycle's boundary phase ended early: its result is not a
      // Gauss-Seidel sweep.  Switch to the list launches for good and let
      // the caller repeat the pass.
      HIPCHK(hipMemsetAsync(lv_[0].coop.as<unsigned>() + 1, 0, sizeof(unsigned), st_));
      sync();
      lv_[0].d.coop_off = 1;
      std::fprintf(stderr,
                   "spgrid_topo_opt_amd: level-0 boundary phase barrier timed out (workgroups not resident "
                   "together); continuing with per-colour launches\n");
      *iters = it;
      return kCoopRetry;
    }
    const CgScalars* h = reinterpret_cast<const CgScalars*>(hscal_);
    *last_nrm = h->nrm;
    if (print_residuals_) {
      FILE* o = print_residuals_ == 2 ? stderr : stdout;
      std::fprintf(o, "%.9g\n", h->nrm);
      std::fflush(o);
    }
    if (h->singular) {
      rc = SPG_E_SINGULAR;
      break;
    }
    if (std::isnan(h->nrm)) {
      rc = SPG_E_NAN;
      break;
    }
    if (h->nrm <= tol) {
      rc = SPG_OK;
      break;
    }
    if (it == maxit) break;
    vcycle<T>(r, z, true);  // z was cleared before the loop / by the last update
    // the scalar stages (beta, alpha, rho_old) run in the last kernel of
    // the reduction they follow
    launch_project_dot<T>(st_, R0.d, z, r, part, &sc->rho, sc, 0);
    launch_cg_direction<T>(st_, s, z, sc, n);
    launch_apply_dot<T>(st_, R0.d, E<T>(), dxT<T>(), s, q, apart_.as<double>(), part, &sc->sq, sc, 1);
    launch_cg_update<T>(st_, x, r, s, q, sc, n, 2, z);
  }
  *iters = it;
  return rc;
}

void Solver::reset_to_ud(double* u) {
  launch_zero<double>(st_, u, lv_[0].n3);
  launch_set_entries(st_, u, ud_idx_.as<int>(), ud_val_.as<double>(), n_ud_);
}

int Solver::solve_rhs(const spg_solve_params& prm, double* u, const double* f, spg_info* info, bool warm) {
  const LevelRT& R0 = lv_[0];
  const int64_t n = R0.n3;
  const double t0 = now_s();
  double* part = part_.as<double>();
  double* red = part + reduce_partials(0);
  auto residual_linf = [&](const double* x) {
    launch_apply<double>(st_, R0.d, E<double>(), cfg_.dx, x, f, res_.as<double>(), 1);
    launch_linf<double>(st_, res_.as<double>(), n, part, red);
    return read_scalar(red);
  };
  // The target is relative to the zero-guess residual: this problem's own,
  // or under a warm start the one stored from the first solve.  A warm
  // guess is kept when forced or when its residual is below that reference
  // (SPGrid_CPU_Solver.cpp:247-285, 318-326).
  double r0 = 0.0, rstart = 0.0;
  bool reused = false;
  if (warm) {
    r0 = prm.reference_linf;
    if (!(r0 > 0)) {
      reset_to_ud(tmp_.as<double>());
      r0 = residual_linf(tmp_.as<double>());
    }
    rstart = residual_linf(u);
    reused = prm.warm_start == 2 || r0 > rstart;
    if (!reused) {
      reset_to_ud(u);
      rstart = residual_linf(u);
    }
  } else {
    r0 = rstart = residual_linf(u);
  }
  const double target = prm.tolerance * r0;
  print_residuals_ = prm.print_residuals;
  int total = 0, passes = 0, rc = SPG_OK;
  double rn = rstart;
  bool conv = (rstart <= target);
  if (std::isnan(rstart) || std::isnan(r0)) rc = SPG_E_NAN;
  const int maxp = prm.max_passes < 1 ? 1 : prm.max_passes;
  for (int pass = 0; pass < maxp && !conv && rc == SPG_OK; ++pass) {
    // per-pass target: an explicit inner reduction, or (inner_tolerance <= 0)
    // the final target itself, but never asking an FP32 pass for more than
    // kFp32PassReduction of its starting residual
    double tolp = target;
    if (prm.inner_tolerance > 0) tolp = std::max(prm.inner_tolerance * rn, 0.7 * target);
    else if (prm.precision == 0) tolp = std::max(target, kFp32PassReduction * rn);
    int it = 0;
    double last = 0;
    int prc;
    for (int attempt = 0;; ++attempt) {
      if (prm.precision == 1) {
        launch_copy<double>(st_, res_.as<double>(), rr64_.as<double>(), n);
        prc = pcg<double>(x64_.as<double>(), rr64_.as<double>(), tolp, prm.max_iterations, &it, &last);
      } else {
        launch_cast_d2f(st_, res_.as<double>(), rr32_.as<float>(), n);
        prc = pcg<float>(x32_.as<float>(), rr32_.as<float>(), tolp, prm.max_iterations, &it, &last);
      }
      if (prc != kCoopRetry) break;
      total += it;
      if (attempt > 0) throw SpgError(SPG_E_HIP, "boundary phase failed after the switch to list launches");
    }
    if (prm.precision == 1) launch_axpy_d<double>(st_, u, x64_.as<double>(), n);
    else launch_axpy_d<float>(st_, u, x32_.as<float>(), n);
    total += it;
    passes++;
    launch_apply<double>(st_, R0.d, E<double>(), cfg_.dx, u, f, res_.as<double>(), 1);
    launch_linf<double>(st_, res_.as<double>(), n, part, red);
    rn = read_scalar(red);
    if (prc == SPG_E_SINGULAR || prc == SPG_E_NAN) rc = prc;
    if (std::isnan(rn)) rc = SPG_E_NAN;
    if (rn <= target) conv = true;
  }
  launch_sumsq(st_, res_.as<double>(), n, part, red);
  const double l2 = std::sqrt(read_scalar(red));
  if (info) {
    static_info(info);
    info->iterations = total;
    info->passes = passes;
    info->converged = conv ? 1 : 0;
    info->r0_linf = r0;
    info->r_linf = rn;
    info->r_l2 = l2;
    info->solve_seconds = now_s() - t0;
    info->r_start_linf = rstart;
    info->warm_reused = reused ? 1 : 0;
  }
  check_coop();
  if (rc == SPG_OK && !conv) rc = SPG_E_NOT_CONVERGED;
  return rc;
}

void Solver::static_info(spg_info* info) const {
  info->levels = lay_.nlev;
  info->setup_seconds = setup_s_;
  info->active_nodes = lay_.lv[0].n_active;
  info->active_cells = lay_.n_active_cells;
  info->gpu_blocks = lay_.lv[0].nb;
  info->bottom_dofs = nvar_;
  info->hbm_bytes = dev_bytes_;
  info->bottom_kind = nd_ ? 1 : 0;
  info->bottom_bytes = nd_ ? (double)nd_->device_bytes()
                           : (double)(tiles64_.bytes + tiles32_.bytes + pb64_.bytes + pb32_.bytes);
}

int Solver::solve(const spg_solve_params& prm, spg_info* info) {
  activate();
  const LevelRT& R0 = lv_[0];
  const int64_t n = R0.n3;
  if (prm.warm_start < 0 || prm.warm_start > 2) throw SpgError(SPG_E_INVALID, "warm_start must be 0, 1 or 2");
  // a warm start begins at the previous solution (or the given guess) with
  // u_D written over it; otherwise at u_D (SPGrid_CPU_Solver.cpp:241-285)
  const bool warm = prm.warm_start != 0 && (have_solution_ || have_guess_);
  if (warm) launch_set_entries(st_, u_.as<double>(), ud_idx_.as<int>(), ud_val_.as<double>(), n_ud_);
  else reset_to_ud(u_.as<double>());
  // f = P P_null forces (SPGrid_CPU_Solver.cpp:289-299)
  launch_zero<double>(st_, f_.as<double>(), n);
  launch_set_entries(st_, f_.as<double>(), f_idx_.as<int>(), f_val_.as<double>(), n_f_);
  launch_project<double>(st_, R0.d, f_.as<double>(), true);
  have_guess_ = false;
  have_solution_ = false;
  const int rc = solve_rhs(prm, u_.as<double>(), f_.as<double>(), info, warm);
  // a failed solve leaves nothing to warm-start from
  have_solution_ = rc == SPG_OK || rc == SPG_E_NOT_CONVERGED;
  return rc;
}

void Solver::set_initial_guess_in3(const double* dev) {
  const LevelRT& R0 = lv_[0];
  launch_zero<double>(st_, u_.as<double>(), R0.n3);
  launch_scatter_input_vec(st_, in_slot_.as<int>(), lay_.nin_blocks * 64, dev, u_.as<double>());
  launch_project<double>(st_, R0.d, u_.as<double>(), true);
  sync();
  have_guess_ = true;
}

int Solver::solve_adjoint(const spg_solve_params& prm, const double* g_in3, double* lam_in3, spg_info* info) {
  activate();
  const LevelRT& R0 = lv_[0];
  const int64_t n = R0.n3;
  const int64_t nin = lay_.nin_blocks * 64;
  launch_zero<double>(st_, tmp_.as<double>(), n);
  launch_scatter_input_vec(st_, in_slot_.as<int>(), nin, g_in3, tmp_.as<double>());
  launch_project<double>(st_, R0.d, tmp_.as<double>(), true);
  // lambda with homogeneous Dirichlet data; reuse z64 as its storage
  launch_zero<double>(st_, z64_.as<double>(), n);
  DBuf lam;
  lam.alloc(n * 8);
  launch_zero<double>(st_, lam.as<double>(), n);
  const int rc = solve_rhs(prm, lam.as<double>(), tmp_.as<double>(), info, false);
  launch_gather_input_vec(st_, in_slot_.as<int>(), nin, lam.as<double>(), lam_in3);
  sync();
  return rc;
}

void Solver::pack_records(const int32_t* base, int hdr, uint8_t* dst) const {
  const int64_t nb = lay_.nin_blocks, nin = nb * 64;
  const size_t rec = (size_t)hdr * 4 + 64 * 32;
  if (nb > 0) {
    DBuf out, bdev;
    out.alloc((size_t)nb * rec);
    if (hdr > 0) {
      bdev.alloc((size_t)nb * 3 * sizeof(int32_t));
      HIPCHK(hipMemcpyAsync(bdev.p, base, (size_t)nb * 3 * sizeof(int32_t), hipMemcpyHostToDevice, st_));
    }
    launch_pack_out4(st_, in_slot_.as<int>(), nin, u_.as<double>(), res_.as<double>(), bdev.as<int32_t>(), hdr,
                     out.as<uint32_t>());
    // through a kept pinned buffer, copied out on host threads: a pageable
    // copy of this size runs at a few GB/s on one thread
    const size_t bytes = (size_t)nb * rec;
    if (stage_bytes_ < bytes) {
      if (stage_) (void)hipHostFree(stage_);
      stage_ = nullptr;
      stage_bytes_ = 0;
      HIPCHK(hipHostMalloc(&stage_, bytes));
      stage_bytes_ = bytes;
    }
    HIPCHK(hipMemcpyAsync(stage_, out.p, bytes, hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    const uint8_t* src = static_cast<const uint8_t*>(stage_);
    parallel_ranges(
        (int64_t)bytes, [&](int64_t a, int64_t b) { std::memcpy(dst + a, src + a, (size_t)(b - a)); }, 1 << 22);
  }
  // Dirichlet entries of nodes without a slot keep their prescribed value
  for (size_t k = 0; k < lay_.ud_extra.size(); ++k) {
    const int64_t e = lay_.ud_extra[k], n = e / 3;
    std::memcpy(dst + (n >> 6) * rec + hdr * 4 + ((n & 63) * 4 + e % 3) * 8, &lay_.ud_extra_value[k], 8);
  }
}

void Solver::pack_output(double* out4) const { pack_records(nullptr, 0, reinterpret_cast<uint8_t*>(out4)); }

void Solver::pack_output_records(const int32_t* base, uint8_t* dst) const { pack_records(base, 3, dst); }

void Solver::displacement_in3(double* dev) const {
  launch_gather_input_vec(st_, in_slot_.as<int>(), lay_.nin_blocks * 64, u_.as<double>(), dev);
  HIPCHK(hipStreamSynchronize(st_));
}

double Solver::sensitivity(const double* u_in3_dev, const double* lam_in3_dev, double* dc_in_dev) {
  activate();
  const LevelRT& R0 = lv_[0];
  const int64_t nslot = (int64_t)R0.d.nb * 64;
  const int64_t nin = lay_.nin_blocks * 64;
  DBuf dc, obj, uu;
  dc.alloc(nslot * 8);
  obj.alloc(nslot * 8);
  const double* u = u_.as<double>();
  if (u_in3_dev) {
    uu.alloc(R0.n3 * 8);
    launch_zero<double>(st_, uu.as<double>(), R0.n3);
    launch_scatter_input_vec(st_, in_slot_.as<int>(), nin, u_in3_dev, uu.as<double>());
    u = uu.as<double>();
  }
  const double* a = u;
  if (lam_in3_dev) {
    launch_zero<double>(st_, tmp_.as<double>(), R0.n3);
    launch_scatter_input_vec(st_, in_slot_.as<int>(), nin, lam_in3_dev, tmp_.as<double>());
    a = tmp_.as<double>();
  }
  launch_sensitivity(st_, R0.d, rho_.as<double>(), u, a, cfg_.penalty, cfg_.min_stiffness, dc.as<double>(),
                     lam_in3_dev ? nullptr : obj.as<double>());
  launch_gather_input_scalar(st_, in_slot_.as<int>(), nin, dc.as<double>(), dc_in_dev);
  double objective = 0.0;
  if (!lam_in3_dev) {
    // ordered FP64 sum of the per-cell objective terms on the host
    std::vector<double> h(nslot);
    HIPCHK(hipMemcpyAsync(h.data(), obj.p, nslot * 8, hipMemcpyDeviceToHost, st_));
    sync();
    for (double v : h) objective += v;
  }
  sync();
  return objective;
}

// ----------------------------------------------------------- kernel hooks
template <class T>
static std::vector<T> to_dev_layout(const double* in, int64_t nb) {
  std::vector<T> v(nb * 192);
  for (int64_t b = 0; b < nb; ++b)
    for (int l = 0; l < 64; ++l)
      for (int c = 0; c < 3; ++c) v[(b * 3 + c) * 64 + l] = (T)in[(b * 64 + l) * 3 + c];
  return v;
}
template <class T>
static void from_dev_layout(const std::vector<T>& v, int64_t nb, double* out) {
  for (int64_t b = 0; b < nb; ++b)
    for (int l = 0; l < 64; ++l)
      for (int c = 0; c < 3; ++c) out[(b * 64 + l) * 3 + c] = (double)v[(b * 3 + c) * 64 + l];
}

void Solver::run_kernel(int op, int level, int precision, int forward, const double* in, const double* b,
                        double* out) {
  activate();
  if (level < 0 || level >= lay_.nlev) throw SpgError(SPG_E_INVALID, "level out of range");
  auto body = [&](auto tag) {
    using T = decltype(tag);
    auto put = [&](DBuf& d, const double* h, int64_t nb) {
      std::vector<T> v = to_dev_layout<T>(h, nb);
      d.alloc(v.size() * sizeof(T));
      HIPCHK(hipMemcpy(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    };
    auto get = [&](const T* d, int64_t nb) {
      std::vector<T> v(nb * 192);
      sync();
      HIPCHK(hipMemcpy(v.data(), d, v.size() * sizeof(T), hipMemcpyDeviceToHost));
      from_dev_layout<T>(v, nb, out);
    };
    const int64_t nb = lv_[level].d.nb;
    DBuf A, Bv, O;
    switch (op) {
      case SPG_OP_APPLY:
      case SPG_OP_RESIDUAL:
      case SPG_OP_OPERATOR: {
        if (level != 0) throw SpgError(SPG_E_INVALID, "apply is a level-0 operator");
        put(A, in, nb);
        if (op == SPG_OP_RESIDUAL) put(Bv, b, nb);
        O.alloc(nb * 192 * sizeof(T));
        launch_apply<T>(st_, lv_[0].d, E<T>(), dxT<T>(), A.as<T>(), Bv.as<T>(), O.as<T>(),
                        op == SPG_OP_APPLY ? 0 : (op == SPG_OP_RESIDUAL ? 1 : 2));
        get(O.as<T>(), nb);
        break;
      }
      case SPG_OP_GS_SWEEP:
      case SPG_OP_GS_BOUNDARY: {
        if (level != 0 && !lay_.lv[level].is_explicit) throw SpgError(SPG_E_INVALID, "level has no smoother");
        put(A, in, nb);
        put(Bv, b, nb);
        const LevelDev& Ld = lv_[level].d;
        const bool fwd = (forward & 1) != 0;
        const int sweeps = std::max(1, forward >> 8);
        if (op == SPG_OP_GS_BOUNDARY) {
          if (level == 0) launch_gs0_boundary<T>(st_, Ld, E<T>(), dxT<T>(), A.as<T>(), Bv.as<T>(), sweeps, fwd);
          else launch_st_boundary<T>(st_, Ld, lv_[level].st<T>(), A.as<T>(), Bv.as<T>(), sweeps, fwd);
        } else {
          const T* bcm = level == 0 ? colour_major_b<T>(Bv.as<T>()) : nullptr;
          for (int c = 0; c < 8; ++c) {
            if (level == 0)
              launch_gs0<T>(st_, Ld, E<T>(), dxT<T>(), A.as<T>(), Bv.as<T>(), sweep_colour(true, fwd, c), nullptr, 0,
                            bcm);
            else
              launch_st_gs<T>(st_, Ld, lv_[level].st<T>(), A.as<T>(), Bv.as<T>(), sweep_colour(false, fwd, c));
          }
        }
        get(A.as<T>(), nb);
        break;
      }
      case SPG_OP_RESTRICT: {
        if (level < 1) throw SpgError(SPG_E_INVALID, "restriction targets level >= 1");
        put(A, in, lv_[level - 1].d.nb);
        O.alloc(nb * 192 * sizeof(T));
        launch_restrict<T>(st_, lv_[level].d, A.as<T>(), O.as<T>());
        get(O.as<T>(), nb);
        break;
      }
      case SPG_OP_PROLONGATE: {
        if (level + 1 >= lay_.nlev) throw SpgError(SPG_E_INVALID, "prolongation needs a coarser level");
        put(A, in, lv_[level + 1].d.nb);
        put(Bv, b, nb);
        launch_prolongate<T>(st_, lv_[level].d, A.as<T>(), Bv.as<T>());
        get(Bv.as<T>(), nb);
        break;
      }
      case SPG_OP_STENCIL_RES: {
        if (!lay_.lv[level].is_explicit) throw SpgError(SPG_E_INVALID, "level has no stencils");
        put(A, in, nb);
        put(Bv, b, nb);
        O.alloc(nb * 192 * sizeof(T));
        launch_st_residual<T>(st_, lv_[level].d, lv_[level].st<T>(), A.as<T>(), Bv.as<T>(), O.as<T>());
        get(O.as<T>(), nb);
        break;
      }
      case SPG_OP_VCYCLE: {
        put(Bv, b, lv_[0].d.nb);
        O.alloc(lv_[0].n3 * sizeof(T));
        vcycle<T>(Bv.as<T>(), O.as<T>());
        get(O.as<T>(), lv_[0].d.nb);
        break;
      }
      case SPG_OP_BOTTOM: {
        const int B = lay_.nlev - 1;
        if (level != B) throw SpgError(SPG_E_INVALID, "bottom solve lives on the last level");
        put(Bv, b, nb);
        O.alloc(nb * 192 * sizeof(T));
        bottom<T>(Bv.as<T>(), O.as<T>());
        get(O.as<T>(), nb);
        break;
      }
      default:
        throw SpgError(SPG_E_INVALID, "unknown op");
    }
  };
  if (precision == 1) body(double());
  else body(float());
  check_coop();
}

void Solver::level_stencils(int level, double* out) {
  if (level < 0 || level >= lay_.nlev || !lay_.lv[level].is_explicit)
    throw SpgError(SPG_E_INVALID, "level has no stencils");
  const int64_t nb = lv_[level].d.nb;
  std::vector<double> v(nb * 243 * 64);
  sync();
  HIPCHK(hipMemcpy(v.data(), lv_[level].st64.p, v.size() * 8, hipMemcpyDeviceToHost));
  for (int64_t b = 0; b < nb; ++b)
    for (int e = 0; e < 243; ++e)
      for (int l = 0; l < 64; ++l) out[(b * 64 + l) * 243 + e] = v[st_index(b, l, e)];
}

double Solver::bench_apply(int precision, int reps) {
  activate();
  const LevelRT& R0 = lv_[0];
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  float ms = 0.f;
  if (precision == 1) {
    launch_copy<double>(st_, u_.as<double>(), x64_.as<double>(), R0.n3);
    launch_apply<double>(st_, R0.d, E<double>(), cfg_.dx, x64_.as<double>(), nullptr, q64_.as<double>(), 2);
    HIPCHK(hipEventRecord(a, st_));
    for (int i = 0; i < reps; ++i)
      launch_apply<double>(st_, R0.d, E<double>(), cfg_.dx, x64_.as<double>(), nullptr, q64_.as<double>(), 2);
    HIPCHK(hipEventRecord(b, st_));
  } else {
    launch_cast_d2f(st_, u_.as<double>(), x32_.as<float>(), R0.n3);
    launch_apply<float>(st_, R0.d, E<float>(), (float)cfg_.dx, x32_.as<float>(), nullptr, q32_.as<float>(), 2);
    HIPCHK(hipEventRecord(a, st_));
    for (int i = 0; i < reps; ++i)
      launch_apply<float>(st_, R0.d, E<float>(), (float)cfg_.dx, x32_.as<float>(), nullptr, q32_.as<float>(), 2);
    HIPCHK(hipEventRecord(b, st_));
  }
  HIPCHK(hipEventSynchronize(b));
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return (double)ms * 1e-3 / reps;
}

double Solver::bench_kernel(int op, int level, int precision, int arg, int reps) {
  activate();
  if (level < 0 || level >= lay_.nlev) throw SpgError(SPG_E_INVALID, "level out of range");
  const int B = lay_.nlev - 1;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float ms = 0.f;
  auto body = [&](auto tag) {
    using T = decltype(tag);
    const LevelRT& R = lv_[level];
    T* u = R.u<T>();
    T* b = R.b<T>();
    T* r = R.r<T>();
    launch_zero<T>(st_, u, R.n3);
    launch_zero<T>(st_, b, R.n3);
    auto run = [&] {
      switch (op) {
        case SPG_OP_APPLY:
        case SPG_OP_RESIDUAL:
          if (level != 0) throw SpgError(SPG_E_INVALID, "apply is a level-0 operator");
          launch_apply<T>(st_, R.d, E<T>(), dxT<T>(), u, b, r, op == SPG_OP_APPLY ? 2 : 1);
          break;
        case SPG_OP_GS_SWEEP: {
          if (level != 0 && !lay_.lv[level].is_explicit) throw SpgError(SPG_E_INVALID, "level has no smoother");
          const T* bcm = level == 0 ? colour_major_b<T>(b) : nullptr;
          for (int c = 0; c < 8; ++c) {
            if (level == 0) launch_gs0<T>(st_, R.d, E<T>(), dxT<T>(), u, b, sweep_colour(true, true, c), nullptr, 0, bcm);
            else launch_st_gs<T>(st_, R.d, R.st<T>(), u, b, sweep_colour(false, true, c));
          }
          break;
        }
        case SPG_OP_GS_BOUNDARY:
          if (level == 0) launch_gs0_boundary<T>(st_, R.d, E<T>(), dxT<T>(), u, b, arg, true);
          else if (lay_.lv[level].is_explicit) launch_st_boundary<T>(st_, R.d, R.st<T>(), u, b, arg, true);
          else throw SpgError(SPG_E_INVALID, "level has no smoother");
          break;
        case SPG_OP_RESTRICT:
          if (level < 1) throw SpgError(SPG_E_INVALID, "restriction targets level >= 1");
          launch_restrict<T>(st_, R.d, lv_[level - 1].r<T>(), b);
          break;
        case SPG_OP_PROLONGATE:
          if (level >= B) throw SpgError(SPG_E_INVALID, "prolongation needs a coarser level");
          launch_prolongate<T>(st_, R.d, lv_[level + 1].u<T>(), u);
          break;
        case SPG_OP_STENCIL_RES:
          if (!lay_.lv[level].is_explicit) throw SpgError(SPG_E_INVALID, "level has no stencils");
          launch_st_residual<T>(st_, R.d, R.st<T>(), u, b, r);
          break;
        case SPG_OP_VCYCLE:
          vcycle<T>(lv_[0].b<T>(), lv_[0].r<T>());
          break;
        case SPG_OP_BOTTOM:
          if (level != B) throw SpgError(SPG_E_INVALID, "bottom solve lives on the last level");
          bottom<T>(b, u);
          break;
        default:
          throw SpgError(SPG_E_INVALID, "unknown op");
      }
    };
    run();  // warm
    HIPCHK(hipEventRecord(e0, st_));
    for (int i = 0; i < reps; ++i) run();
    HIPCHK(hipEventRecord(e1, st_));
  };
  if (precision == 1) body(double());
  else body(float());
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  check_coop();
  return (double)ms * 1e-3 / reps;
}

template void Solver::vcycle<float>(const float*, float*, bool);
template void Solver::vcycle<double>(const double*, double*, bool);

}  // namespace spg
