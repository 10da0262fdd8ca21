// solver.hpp -- device-resident narrow-band elasticity solver.
//
// Owns the level hierarchy on one GPU, the Galerkin stencils (assembled in
// FP64, used in FP32 by the mixed inner solve), the FP64 bottom Cholesky
// inverse, and the solution state; implements the reference's solve
// (solver/SPGrid_CPU_Solver.cpp:241-341) with an iterated FP64
// defect-correction outer loop around a FP32 (or FP64) MG-PCG.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <memory>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/spgrid_topo_opt_amd.h"
#include "bottom_nd.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace spg {

struct SpgError : std::runtime_error {
  int code;
  SpgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// owning device allocation
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
    return *this;
  }
  ~DBuf();
  void alloc(size_t n);
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

struct LevelRT {
  LevelDev d;
  DBuf nbr, flags, rsrc, psrc, psrc27, base, bnd;
  DBuf st64, st32, cell_row, cfb, cgi;
  DBuf bci, bslot, bA64, bA32, bt64, bt32;  // boundary system
  DBuf b0cslot, b0node, b0E64, b0E32, b0D64, b0D32;  // level-0 matrix-free boundary system
  DBuf b0ub, coop;  // cooperative level-0 phase: iterate and {barrier counter, error word}
  DBuf bmap64[2], bmap32[2], bmx64, bmx32;           // dense boundary maps (explicit levels)
  DBuf bnslot;                                       // level 0: neighbour slots of boundary nodes
  DBuf gsdist, gsnear, gsnslot;                      // level 0: band distances, near-band step lists, their slots
  int64_t gsnear_off[17] = {};                       // list (dir * 8 + step) = gsnear[off[i] .. off[i+1])
  DBuf u32, b32, r32, u64, b64, r64;
  int64_t n3 = 0;  // vector length nb*192
  template <class T> T* u() const;
  template <class T> T* b() const;
  template <class T> T* r() const;
  template <class T> const T* st() const;
};

class Solver {
 public:
  Solver(const spg_config& cfg, const ProblemDesc& pd);
  ~Solver();

  void set_forces(int64_t n, const int32_t* coord, const double* value);
  // false if the densities would switch a cell on or off (a topology change
  // that needs a new solver); throws on NaN
  bool active_set_matches(const double* density) const;
  void update_density(const double* density);
  void update_density_device(const double* density_dev);  // input block order, device
  // returns SPG_OK or error code; fills info
  int solve(const spg_solve_params& prm, spg_info* info);
  int solve_adjoint(const spg_solve_params& prm, const double* g_dev_in3, double* lam_dev_in3, spg_info* info);
  void pack_output(double* out4) const;  // host, [nin_blocks * 64][4]
  // host, per input block the FEMOutputs displacement record: int32
  // base[3] (from base[3b..3b+2]) then 64 x 4 doubles (2060 bytes)
  void pack_output_records(const int32_t* base, uint8_t* dst) const;
  void displacement_in3(double* dev) const;
  // starting guess of the next warm-started solve (input order, device)
  void set_initial_guess_in3(const double* dev);
  double sensitivity(const double* u_in3_dev, const double* lam_in3_dev, double* dc_in_dev);

  // kernel-level hooks
  int levels() const { return lay_.nlev; }
  const Layout& layout() const { return lay_; }
  void run_kernel(int op, int level, int precision, int forward, const double* in, const double* b, double* out);
  void level_stencils(int level, double* out);
  double bench_apply(int precision, int reps);
  double bench_kernel(int op, int level, int precision, int arg, int reps);
  // sizes, levels, device bytes and setup time (the solve-independent spg_info fields)
  void static_info(spg_info* info) const;
  const spg_config& config() const { return cfg_; }
  double setup_seconds() const { return setup_s_; }
  double device_bytes() const { return dev_bytes_; }

 private:
  // z = M b0; `z_is_zero` says the caller has already cleared z
  template <class T> void vcycle(const T* b0, T* z, bool z_is_zero = false);
  void vcycle_coarse64(const float* b0, float* z, bool z_is_zero);
  void pack_records(const int32_t* base, int hdr, uint8_t* dst) const;
  template <class T> void smooth(int l, T* u, const T* b, bool forward, const T* bcm = nullptr);
  // the level-0 colour sweeps read b and the flags from colour-major copies
  // when the level's working set exceeds the memory-side cache (or
  // SPG_GS_CMB=1); returns the copy of b0 to pass, or null
  template <class T> const T* colour_major_b(const T* b0);
  void setup_colour_major();
  int cmb_mode_ = -1;  // SPG_GS_CMB: 0 never, 1 always, unset by size
  template <class T> void level_residual(int l, const T* u, const T* b, T* r);
  template <class T> void bottom(const T* b, T* u);
  template <class T> int pcg(T* x, T* r, double tol, int maxit, int* iters, double* last_nrm);
  int print_residuals_ = 0;  // 0 off, 1 stdout, 2 stderr (per PCG iteration)
  DBuf derr_;                // error flags of the device density check
  template <class T> const T* E() const;
  template <class T> T dxT() const { return (T)cfg_.dx; }
  // u holds the starting guess (u_D set); with `warm` it may be replaced by
  // u_D when its residual is not below the reference (the reference's
  // reuse test)
  int solve_rhs(const spg_solve_params& prm, double* u, const double* f, spg_info* info, bool warm);
  void reset_to_ud(double* u);
  void build_operators();
  // makes the device's element constants this solver's material (uploaded
  // only when another material was the last one used on the device)
  void activate();
  void build_boundary_rows();
  void upload_layout();
  double read_scalar(const double* dev);
  void sync();
  // throws (once per event) if a cooperative boundary phase reported a
  // barrier time-out
  void check_coop();

  spg_config cfg_;
  Layout lay_;
  std::vector<LevelRT> lv_;
  double K0_[576];
  hipStream_t st_ = nullptr;
  // SPG_OVERLAP=1: the far part of a level-0 sweep runs on st2_ beside the
  // boundary phase on st_ (off by default: measured slower)
  hipStream_t st2_ = nullptr;
  hipEvent_t ev_a_ = nullptr, ev_b_ = nullptr;
  bool overlap_ = false;
  rocblas_handle rb_ = nullptr;
  // level-0 data
  DBuf E64_, E32_, rho_, in_slot_, ud_idx_, ud_val_, f_idx_, f_val_;
  int64_t n_ud_ = 0, n_f_ = 0;
  // bottom
  int64_t nvar_ = 0;
  int nT_ = 0;  // 64-row tiles of the bottom inverse
  bool bottom64_ = false;  // FP32 cycles apply the FP64 bottom inverse
  bool coarse64_ = false;  // FP32 cycles run the explicit levels in FP64
  DBuf var_of_, info_, tiles64_, tiles32_, xb64_, yb64_, xb32_, yb32_, pb64_, pb32_;
  std::unique_ptr<NdBottom> nd_;  // nested-dissection Cholesky bottom (large bottoms)
  // PCG / outer vectors (level 0)
  DBuf x32_, rr32_, s32_, q32_, z32_, x64_, rr64_, s64_, q64_, z64_;
  DBuf u_, f_, res_, tmp_;
  DBuf scal_, part_, pinned_;
  DBuf apart_;  // per-workgroup partials of the fused operator apply + dot
  DBuf fcm_, bcm_;  // level-0 flags and b in colour-major order (either precision)
  double* hscal_ = nullptr;
  mutable void* stage_ = nullptr;  // pinned staging of the output records
  mutable size_t stage_bytes_ = 0;
  double setup_s_ = 0.0;
  double dev_bytes_ = 0.0;
  const double* posmat_device(int depth);
  std::unordered_map<uint64_t, int32_t> gmap0_;
  bool have_solution_ = false;  // u_ holds a converged (or at least finite) solution
  bool have_guess_ = false;     // u_ holds a guess from set_initial_guess_in3
};

}  // namespace spg
