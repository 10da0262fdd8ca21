// bottom_nd.hpp -- exact FP64 bottom solve by a nested-dissection
// multifrontal Cholesky factorization on the GPU (see bottom_nd.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace spg {

// a front that is not positive definite (the solver reports SPG_E_SINGULAR)
struct NdNotSpd : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// owning device allocation (bytes counted for the solver's footprint)
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem();
  void alloc(size_t n);
};

// one front as the solve kernels see it; dofs are in elimination order
struct NdFrontDev {
  int64_t dof0;      // first pivot dof (the pivots are dof0 .. dof0 + p - 1)
  int64_t mf_off;    // [M; G], m x p row-major, in the forward store
  int64_t mb_off;    // [M^T G^T], p x m row-major, in the backward store
  int64_t w_off;     // m entries of solve workspace
  int64_t u_off;     // b entries: this front's forward update vector
  int64_t bidx_off;  // b entries: dof index of each boundary dof
  int64_t g_off;     // m + 1 entries of the gather list's row pointer
  int32_t p, b;      // pivot and boundary dofs
};

struct NdStats {
  int64_t fronts = 0, depth = 0, factor_entries = 0, max_front = 0;
};

class NdBottom {
 public:
  // symbolic analysis of a level's active nodes (flags & active_bit) from
  // its block origins, 27-neighbour block table and flags (host arrays)
  NdBottom(int64_t nb, const int32_t* base, const int32_t* nbr, const uint8_t* flags, uint8_t active_bit);
  // device copies of the plan and the factor/solve storage
  void upload(hipStream_t st);
  // numeric factorization from the level's FP64 stencils (st_index layout);
  // throws NdNotSpd for a front that is not positive definite
  void factor(const double* ST, hipStream_t st, rocblas_handle rb);
  // u = A^-1 b on the level's vectors (inactive nodes: u = b), FP64 inside
  template <class T>
  void solve(int nb, const T* b, T* u, hipStream_t st);
  int64_t dofs() const { return ndof_; }
  size_t device_bytes() const;
  NdStats stats() const;

 private:
  struct Front {
    int64_t pos0 = 0;  // first pivot node's position
    int np = 0;        // pivot nodes
    int depth = 0, parent = -1;
    std::vector<int> kids;
    std::vector<int64_t> bnd;  // positions of the boundary nodes, ascending
    int64_t ustore_off = 0, emap_off = 0, ent_off = 0, nent = 0;
  };
  void solve_pos(hipStream_t st);

  int64_t nnode_ = 0, ndof_ = 0;
  int maxdepth_ = 0;
  std::vector<Front> fr_;
  std::vector<NdFrontDev> fdev_;
  std::vector<int32_t> var_of_;  // slot -> first dof, -1 if inactive
  std::vector<int64_t> gptr_, gidx_, bidx_;
  std::vector<int32_t> emap_;
  std::vector<int4> ent_;
  std::vector<int2> items_, piv_items_;
  std::vector<int64_t> lvl_off_, piv_off_;
  int64_t mstore_ = 0, wlen_ = 0, ulen_ = 0, ustore_len_ = 0;
  size_t maxm2_ = 0, maxbp_ = 0;
  DevMem d_fr_, d_gptr_, d_gidx_, d_bidx_, d_emap_, d_ent_, d_items_, d_piv_items_, d_var_of_;
  DevMem mfs_, mbs_, w_, u_, bx_, y_, x_;
};

}  // namespace spg
