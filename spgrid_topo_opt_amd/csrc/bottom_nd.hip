// bottom_nd.hip -- exact FP64 bottom solve by a nested-dissection
// multifrontal Cholesky factorization on the GPU.
//
// The reference factorizes the bottom level's Dirichlet-deleted stencil
// matrix with MKL PARDISO (nested-dissection ordering, sparse supernodal
// Cholesky; solver/MG/Elasticity_Matrix.cpp:15-269, called from
// MG/Multigrid_Hierarchy.h:332-372) and solves it once per V-cycle.  This
// file does the same job for bottoms too large for a dense inverse:
//
//   ordering   geometric nested dissection of the bottom nodes: the node box
//              is cut by the median plane normal to its longest axis; a plane
//              of nodes separates the two halves of a 27-point stencil.
//   symbolic   one front per separator (and per leaf box): its pivots are
//              eliminated together, its boundary is the set of later nodes
//              coupled to its subtree (host, once per topology).
//   numeric    per front, in postorder: assemble the stencil entries and the
//              children's update matrices, partial Cholesky (rocSOLVER
//              potrf, rocBLAS trsm and syrk), then the explicit inverse
//              M = L11^-1 (trtri) and G = L21 M (trmm) for the solve.
//   solve      forward (deepest fronts first): y_P = M w_P and
//              u = w_B - G w_P, with w the front's right-hand side plus its
//              children's update vectors; backward (root first):
//              x_P = M^T y_P - G^T x_B.  Every step is a GEMV over a stored
//              row-major copy, one wavefront per row, one launch per tree
//              depth; sums run in a fixed order, so a solve is deterministic.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <functional>
#include <cstdio>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "bottom_nd.hpp"
#include "kernels.hpp"

namespace spg {

namespace {

constexpr int kLeafNodes = 64;  // a box of at most this many nodes is one front

#define NDCHK(x)                                                                              \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define NDRB(x)                                                                                       \
  do {                                                                                                \
    rocblas_status s_ = (x);                                                                          \
    if (s_ != rocblas_status_success)                                                                 \
      throw std::runtime_error(std::string(#x) + ": rocblas status " + std::to_string((int)s_));     \
  } while (0)

// ------------------------------------------------------------------ kernels
// Stencil entries of one front: entry e = (slot of pivot node i, spoke s,
// local dof of i, local dof of neighbour j); F(j*, i*) = st(i, s) (the
// transpose of row i's block, by symmetry), F column-major with leading
// dimension m.
__global__ void k_nd_assemble(int n, const int4* __restrict__ ent, const double* __restrict__ ST, int m,
                              double* __restrict__ F) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * 9) return;
  const int4 e = ent[g / 9];
  const int v = (g % 9) / 3, w = g % 3;
  const int blk = e.x >> 6, l = e.x & 63;
  F[(int64_t)(e.w + w) + (int64_t)(e.z + v) * m] = ST[st_index(blk, l, e.y * 9 + v * 3 + w)];
}

// F(map[a], map[b]) += U(a, b) over the lower triangle of a child's update
// matrix U (b x b, column-major)
__global__ void k_nd_extend_add(int b, const double* __restrict__ U, const int* __restrict__ map, int m,
                                double* __restrict__ F) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)b * b) return;
  const int a = (int)(g % b), c = (int)(g / b);
  if (a < c) return;
  F[(int64_t)map[a] + (int64_t)map[c] * m] += U[g];
}

// The solve's copies of one front: MF (m x p, row-major) = [M; G] and
// MB (p x m, row-major) = [M^T G^T].  M is the lower triangle of F's leading
// p x p block, G (b x p, column-major) sits in T.
__global__ void k_nd_pack(int p, int b, int m, const double* __restrict__ F, const double* __restrict__ T,
                          double* __restrict__ MF, double* __restrict__ MB) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)m * p) return;
  const int r = (int)(g / p), j = (int)(g % p);  // MF[r][j]
  double v;
  if (r < p) v = r >= j ? F[(int64_t)r + (int64_t)j * m] : 0.0;
  else v = T[(int64_t)(r - p) + (int64_t)j * b];
  MF[g] = v;
  MB[(int64_t)j * m + r] = v;
}

// forward gather: w[f] = (f < p ? b[pos0 + f] : 0) + sum of the children's
// update entries mapped to f
__global__ void k_nd_fgather(int n, const int2* __restrict__ item, const NdFrontDev* __restrict__ fr,
                             const int64_t* __restrict__ gptr, const int64_t* __restrict__ gidx,
                             const double* __restrict__ b, const double* __restrict__ ustore,
                             double* __restrict__ w) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int2 it = item[g];
  const NdFrontDev F = fr[it.x];
  const int f = it.y;
  double v = f < F.p ? b[F.dof0 + f] : 0.0;
  const int64_t k0 = gptr[F.g_off + f], k1 = gptr[F.g_off + f + 1];
  for (int64_t k = k0; k < k1; ++k) v += ustore[gidx[k]];
  w[F.w_off + f] = v;
}

__device__ __forceinline__ double wave_sum(double s) {
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  return s;
}

// forward GEMV, one wavefront per row r of a front: rows r < p give
// y = M w_P (lower triangle), rows r >= p the update u = w_B - G w_P
__global__ void __launch_bounds__(256) k_nd_fgemv(int n, const int2* __restrict__ item,
                                                  const NdFrontDev* __restrict__ fr, const double* __restrict__ MF,
                                                  const double* __restrict__ w, double* __restrict__ y,
                                                  double* __restrict__ ustore) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const int2 it = item[row];
  const NdFrontDev F = fr[it.x];
  const int r = it.y;
  const int len = r < F.p ? r + 1 : F.p;
  const double* A = MF + F.mf_off + (int64_t)r * F.p;
  const double* x = w + F.w_off;
  double s = 0.0;
  for (int j = lane; j < len; j += 64) s += A[j] * x[j];
  s = wave_sum(s);
  if (lane == 0) {
    if (r < F.p) y[F.dof0 + r] = s;
    else ustore[F.u_off + (r - F.p)] = x[r] - s;
  }
}

// backward gather: v = [y_P; -x_B]
__global__ void k_nd_bgather(int n, const int2* __restrict__ item, const NdFrontDev* __restrict__ fr,
                             const int64_t* __restrict__ bidx, const double* __restrict__ y,
                             const double* __restrict__ x, double* __restrict__ w) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int2 it = item[g];
  const NdFrontDev F = fr[it.x];
  const int f = it.y;
  w[F.w_off + f] = f < F.p ? y[F.dof0 + f] : -x[bidx[F.bidx_off + (f - F.p)]];
}

// backward GEMV, one wavefront per pivot row i: x_i = sum_{j >= i} MB[i][j] v_j
__global__ void __launch_bounds__(256) k_nd_bgemv(int n, const int2* __restrict__ item,
                                                  const NdFrontDev* __restrict__ fr, const double* __restrict__ MB,
                                                  const double* __restrict__ w, double* __restrict__ x) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const int2 it = item[row];
  const NdFrontDev F = fr[it.x];
  const int i = it.y;
  const int m = F.p + F.b;
  const double* A = MB + F.mb_off + (int64_t)i * m;
  const double* v = w + F.w_off;
  double s = 0.0;
  for (int j = i + lane; j < m; j += 64) s += A[j] * v[j];
  s = wave_sum(s);
  if (lane == 0) x[F.dof0 + i] = s;
}

// level vector (T) -> pos-ordered FP64 right-hand side, and back
template <class T>
__global__ void k_nd_in(int nb, const int* __restrict__ var_of, const T* __restrict__ b, double* __restrict__ x) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int vi = var_of[gid];
  if (vi < 0) return;
  const int blk = gid >> 6, l = gid & 63;
  for (int c = 0; c < 3; ++c) x[vi + c] = (double)b[(blk * 3 + c) * 64 + l];
}
template <class T>
__global__ void k_nd_out(int nb, const int* __restrict__ var_of, const double* __restrict__ x,
                         const T* __restrict__ b, T* __restrict__ u) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int vi = var_of[gid];
  const int blk = gid >> 6, l = gid & 63;
  for (int c = 0; c < 3; ++c) {
    const int k = (blk * 3 + c) * 64 + l;
    u[k] = vi >= 0 ? (T)x[vi + c] : b[k];
  }
}

unsigned grid(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <class V>
void upload_vec(DevMem& d, const std::vector<V>& h, hipStream_t st) {
  d.alloc(h.size() * sizeof(V));
  if (!h.empty()) NDCHK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, st));
}

}  // namespace

DevMem::~DevMem() {
  if (p) (void)hipFree(p);
}
void DevMem::alloc(size_t n) {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = n;
  if (n) NDCHK(hipMalloc(&p, n));
}

// ------------------------------------------------------------------ symbolic
NdBottom::NdBottom(int64_t nb, const int32_t* base, const int32_t* nbr, const uint8_t* flags, uint8_t active_bit) {
  // active nodes, their coordinates and 27 neighbours
  std::vector<int32_t> id_of(nb * 64, -1);
  std::vector<int32_t> slot;
  for (int64_t s = 0; s < nb * 64; ++s)
    if (flags[s] & active_bit) {
      id_of[s] = (int32_t)slot.size();
      slot.push_back((int32_t)s);
    }
  const int64_t n = (int64_t)slot.size();
  nnode_ = n;
  std::vector<int32_t> xyz(n * 3), adj(n * 27, -1);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t blk = slot[i] >> 6;
    const int l = slot[i] & 63, li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
    xyz[i * 3 + 0] = base[blk * 3 + 0] + li;
    xyz[i * 3 + 1] = base[blk * 3 + 1] + lj;
    xyz[i * 3 + 2] = base[blk * 3 + 2] + lk;
    for (int s = 0; s < 27; ++s) {
      const int a = li + s / 9 - 1, b = lj + (s / 3) % 3 - 1, c = lk + s % 3 - 1;
      const int sx = a < 0 ? 0 : (a > 3 ? 2 : 1), sy = b < 0 ? 0 : (b > 3 ? 2 : 1), sz = c < 0 ? 0 : (c > 3 ? 2 : 1);
      const int sp = sx * 9 + sy * 3 + sz;
      const int64_t bb = sp == 13 ? blk : nbr[blk * 27 + sp];
      if (bb < 0) continue;
      adj[i * 27 + s] = id_of[bb * 64 + (((a & 3) << 4) | ((b & 3) << 2) | (c & 3))];
    }
  }
  // nested dissection: fronts in postorder, each front's pivots take the
  // next consecutive positions
  std::vector<int64_t> pos(n, -1);
  int64_t next = 0;
  std::vector<int32_t> work(n);
  std::iota(work.begin(), work.end(), 0);
  std::vector<int32_t> piv_nodes;  // all fronts' pivots, in position order
  // recursive helper (depth stays below ~64: each cut halves the box)
  std::function<int(int64_t, int64_t, int)> nd = [&](int64_t lo, int64_t hi, int depth) -> int {
    const int64_t cnt = hi - lo;
    std::vector<int> kids;
    int64_t slo = lo, shi = hi;  // the separator's range after the split
    if (cnt > kLeafNodes) {
      int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
      for (int64_t k = lo; k < hi; ++k)
        for (int d = 0; d < 3; ++d) {
          mn[d] = std::min(mn[d], xyz[work[k] * 3 + d]);
          mx[d] = std::max(mx[d], xyz[work[k] * 3 + d]);
        }
      int ax = 0;
      for (int d = 1; d < 3; ++d)
        if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
      if (mx[ax] > mn[ax]) {
        const int64_t mid = lo + cnt / 2;
        std::nth_element(work.begin() + lo, work.begin() + mid, work.begin() + hi,
                         [&](int32_t a, int32_t b) { return xyz[a * 3 + ax] < xyz[b * 3 + ax]; });
        const int32_t p = xyz[work[mid] * 3 + ax];
        // three-way partition: < p | > p | == p
        auto it1 = std::stable_partition(work.begin() + lo, work.begin() + hi,
                                         [&](int32_t a) { return xyz[a * 3 + ax] < p; });
        auto it2 = std::stable_partition(it1, work.begin() + hi, [&](int32_t a) { return xyz[a * 3 + ax] > p; });
        const int64_t a1 = it1 - work.begin(), a2 = it2 - work.begin();
        if (a1 > lo) kids.push_back(nd(lo, a1, depth + 1));
        if (a2 > a1) kids.push_back(nd(a1, a2, depth + 1));
        slo = a2;
        shi = hi;
      }
    }
    // this front: the separator (or the whole leaf box)
    std::sort(work.begin() + slo, work.begin() + shi, [&](int32_t a, int32_t b) {
      for (int d = 0; d < 3; ++d)
        if (xyz[a * 3 + d] != xyz[b * 3 + d]) return xyz[a * 3 + d] < xyz[b * 3 + d];
      return false;
    });
    Front F;
    F.pos0 = next;
    F.np = (int)(shi - slo);
    F.depth = depth;
    F.kids = kids;
    for (int64_t k = slo; k < shi; ++k) {
      pos[work[k]] = next++;
      piv_nodes.push_back(work[k]);
    }
    const int id = (int)fr_.size();
    for (int c : kids) fr_[c].parent = id;
    fr_.push_back(std::move(F));
    return id;
  };
  if (n > 0) nd(0, n, 0);
  // the nd() recursion hands out depths from the root; levels of the solve
  // are the depths
  maxdepth_ = 0;
  for (const Front& F : fr_) maxdepth_ = std::max(maxdepth_, F.depth);
  // boundary sets, postorder
  for (size_t t = 0; t < fr_.size(); ++t) {
    Front& F = fr_[t];
    const int64_t last = F.pos0 + F.np;
    std::vector<int64_t> B;
    for (int c : F.kids)
      for (int64_t q : fr_[c].bnd)
        if (q >= last) B.push_back(q);
    for (int64_t k = F.pos0; k < last; ++k) {
      const int32_t i = piv_nodes[k];
      for (int s = 0; s < 27; ++s) {
        const int32_t j = adj[i * 27 + s];
        if (j >= 0 && pos[j] >= last) B.push_back(pos[j]);
      }
    }
    std::sort(B.begin(), B.end());
    B.erase(std::unique(B.begin(), B.end()), B.end());
    F.bnd = std::move(B);
  }
  // var_of: slot -> first dof (3 * position)
  var_of_.assign(nb * 64, -1);
  for (int64_t i = 0; i < n; ++i) var_of_[slot[i]] = (int32_t)(3 * pos[i]);
  ndof_ = 3 * n;
  // local index of a position inside a front: pivots first, then boundary
  auto local = [&](const Front& F, int64_t q) -> int {
    if (q >= F.pos0 && q < F.pos0 + F.np) return (int)(q - F.pos0);
    const auto it = std::lower_bound(F.bnd.begin(), F.bnd.end(), q);
    if (it == F.bnd.end() || *it != q) return -1;
    return F.np + (int)(it - F.bnd.begin());
  };
  // device-side front table, work lists per depth, gather lists, maps
  std::vector<NdFrontDev> fd(fr_.size());
  int64_t mf = 0, wofs = 0, uofs = 0, bofs = 0, gofs = 0, ustore = 0;
  size_t maxm2 = 0, maxbp = 0;
  std::vector<int64_t> gptr, gidx, bidx;
  std::vector<int32_t> emap;
  for (size_t t = 0; t < fr_.size(); ++t) {
    Front& F = fr_[t];
    NdFrontDev& D = fd[t];
    D.p = 3 * F.np;
    D.b = 3 * (int)F.bnd.size();
    const int m = D.p + D.b;
    D.dof0 = 3 * F.pos0;
    D.mf_off = mf;
    D.mb_off = mf;  // MB has the same size and offsets in its own store
    mf += (int64_t)m * D.p;
    D.w_off = wofs;
    wofs += m;
    D.u_off = uofs;
    uofs += D.b;
    D.bidx_off = bofs;
    for (int64_t q : F.bnd)
      for (int c = 0; c < 3; ++c) bidx.push_back(3 * q + c);
    bofs += D.b;
    D.g_off = gofs;
    F.ustore_off = ustore;
    ustore += (int64_t)D.b * D.b;
    maxm2 = std::max(maxm2, (size_t)m * m);
    maxbp = std::max(maxbp, (size_t)D.b * D.p);
    // gather lists: children's update entries mapped into this front
    std::vector<std::vector<int64_t>> contrib(m);
    for (int c : F.kids) {
      const Front& C = fr_[c];
      const NdFrontDev& CD = fd[c];
      fr_[c].emap_off = (int64_t)emap.size();
      for (size_t k = 0; k < C.bnd.size(); ++k) {
        const int lq = local(F, C.bnd[k]);
        if (lq < 0) throw std::runtime_error("nested dissection: a child's boundary node is missing from its parent");
        for (int cc = 0; cc < 3; ++cc) {
          emap.push_back(3 * lq + cc);
          contrib[3 * lq + cc].push_back(CD.u_off + 3 * (int64_t)k + cc);
        }
      }
    }
    gptr.push_back((int64_t)gidx.size());
    for (int f = 0; f < m; ++f) {
      for (int64_t q : contrib[f]) gidx.push_back(q);
      gptr.push_back((int64_t)gidx.size());
    }
    gofs += m + 1;
    // stencil entries: pivot i, spoke s, neighbour j later in the front
    F.ent_off = (int64_t)ent_.size();
    for (int64_t k = F.pos0; k < F.pos0 + F.np; ++k) {
      const int32_t i = piv_nodes[k];
      const int li = (int)(k - F.pos0);
      for (int s = 0; s < 27; ++s) {
        const int32_t j = adj[i * 27 + s];
        if (j < 0) continue;
        const int lj = local(F, pos[j]);
        if (lj < 0 || lj < li) continue;  // eliminated earlier, or an earlier pivot's entry
        ent_.push_back(make_int4(slot[i], s, 3 * li, 3 * lj));
      }
    }
    F.nent = (int64_t)ent_.size() - F.ent_off;
  }
  mstore_ = mf;
  wlen_ = wofs;
  ulen_ = uofs;
  ustore_len_ = ustore;
  maxm2_ = maxm2;
  maxbp_ = maxbp;
  fdev_ = std::move(fd);
  gptr_ = std::move(gptr);
  gidx_ = std::move(gidx);
  bidx_ = std::move(bidx);
  emap_ = std::move(emap);
  // work items per depth: fgather/fgemv/bgather over (front, 0..m), bgemv over (front, 0..p)
  items_.clear();
  piv_items_.clear();
  lvl_off_.assign(maxdepth_ + 2, 0);
  piv_off_.assign(maxdepth_ + 2, 0);
  for (int d = 0; d <= maxdepth_; ++d) {
    lvl_off_[d] = (int64_t)items_.size();
    piv_off_[d] = (int64_t)piv_items_.size();
    for (size_t t = 0; t < fr_.size(); ++t) {
      if (fr_[t].depth != d) continue;
      const int m = fdev_[t].p + fdev_[t].b;
      for (int f = 0; f < m; ++f) items_.push_back(make_int2((int)t, f));
      for (int f = 0; f < fdev_[t].p; ++f) piv_items_.push_back(make_int2((int)t, f));
    }
  }
  lvl_off_[maxdepth_ + 1] = (int64_t)items_.size();
  piv_off_[maxdepth_ + 1] = (int64_t)piv_items_.size();
}

void NdBottom::upload(hipStream_t st) {
  upload_vec(d_fr_, fdev_, st);
  upload_vec(d_gptr_, gptr_, st);
  upload_vec(d_gidx_, gidx_, st);
  upload_vec(d_bidx_, bidx_, st);
  upload_vec(d_emap_, emap_, st);
  upload_vec(d_ent_, ent_, st);
  upload_vec(d_items_, items_, st);
  upload_vec(d_piv_items_, piv_items_, st);
  upload_vec(d_var_of_, var_of_, st);
  mfs_.alloc((size_t)mstore_ * 8);
  mbs_.alloc((size_t)mstore_ * 8);
  w_.alloc((size_t)std::max<int64_t>(wlen_, 1) * 8);
  u_.alloc((size_t)std::max<int64_t>(ulen_, 1) * 8);
  bx_.alloc((size_t)std::max<int64_t>(ndof_, 1) * 8);
  y_.alloc((size_t)std::max<int64_t>(ndof_, 1) * 8);
  x_.alloc((size_t)std::max<int64_t>(ndof_, 1) * 8);
  NDCHK(hipStreamSynchronize(st));
}

size_t NdBottom::device_bytes() const {
  size_t s = 0;
  for (const DevMem* d : {&d_fr_, &d_gptr_, &d_gidx_, &d_bidx_, &d_emap_, &d_ent_, &d_items_, &d_piv_items_,
                          &d_var_of_, &mfs_, &mbs_, &w_, &u_, &bx_, &y_, &x_})
    s += d->bytes;
  return s;
}

// ------------------------------------------------------------------ numeric
void NdBottom::factor(const double* ST, hipStream_t st, rocblas_handle rb) {
  DevMem F, T, U, info;
  F.alloc(std::max<size_t>(maxm2_, 1) * 8);
  T.alloc(std::max<size_t>(maxbp_, 1) * 8);
  U.alloc((size_t)std::max<int64_t>(ustore_len_, 1) * 8);
  info.alloc(2 * std::max<size_t>(fr_.size(), 1) * sizeof(rocblas_int));
  NDCHK(hipMemsetAsync(info.p, 0, info.bytes, st));
  double* Fp = static_cast<double*>(F.p);
  double* Tp = static_cast<double*>(T.p);
  double* Up = static_cast<double*>(U.p);
  rocblas_int* ip = static_cast<rocblas_int*>(info.p);
  const double one = 1.0, mone = -1.0;
  for (size_t t = 0; t < fr_.size(); ++t) {
    const Front& Fr = fr_[t];
    const NdFrontDev& D = fdev_[t];
    const int p = D.p, b = D.b, m = p + b;
    NDCHK(hipMemsetAsync(Fp, 0, (size_t)m * m * 8, st));
    if (Fr.nent > 0)
      k_nd_assemble<<<grid(Fr.nent * 9), 256, 0, st>>>((int)Fr.nent, static_cast<const int4*>(d_ent_.p) + Fr.ent_off,
                                                      ST, m, Fp);
    for (int c : Fr.kids) {
      const int bc = fdev_[c].b;
      if (bc == 0) continue;
      k_nd_extend_add<<<grid((int64_t)bc * bc), 256, 0, st>>>(bc, Up + fr_[c].ustore_off,
                                                              static_cast<const int*>(d_emap_.p) + fr_[c].emap_off, m,
                                                              Fp);
    }
    NDRB(rocsolver_dpotrf(rb, rocblas_fill_lower, p, Fp, m, ip + 2 * t));
    if (b > 0) {
      NDRB(rocblas_dtrsm(rb, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                         rocblas_diagonal_non_unit, b, p, &one, Fp, m, Fp + p, m));
      NDRB(rocblas_dsyrk(rb, rocblas_fill_lower, rocblas_operation_none, b, p, &mone, Fp + p, m, &one,
                         Fp + p + (int64_t)p * m, m));
      NDCHK(hipMemcpy2DAsync(Up + Fr.ustore_off, (size_t)b * 8, Fp + p + (int64_t)p * m, (size_t)m * 8,
                             (size_t)b * 8, b, hipMemcpyDeviceToDevice, st));
    }
    NDRB(rocsolver_dtrtri(rb, rocblas_fill_lower, rocblas_diagonal_non_unit, p, Fp, m, ip + 2 * t + 1));
    if (b > 0)
      NDRB(rocblas_dtrmm(rb, rocblas_side_right, rocblas_fill_lower, rocblas_operation_none,
                         rocblas_diagonal_non_unit, b, p, &one, Fp, m, Fp + p, m, Tp, b));
    k_nd_pack<<<grid((int64_t)m * p), 256, 0, st>>>(p, b, m, Fp, Tp, static_cast<double*>(mfs_.p) + D.mf_off,
                                                    static_cast<double*>(mbs_.p) + D.mb_off);
  }
  std::vector<rocblas_int> h(2 * fr_.size());
  if (!h.empty())
    NDCHK(hipMemcpyAsync(h.data(), info.p, h.size() * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  NDCHK(hipStreamSynchronize(st));
  for (size_t t = 0; t < fr_.size(); ++t)
    if (h[2 * t] != 0 || h[2 * t + 1] != 0)
      throw NdNotSpd("bottom Cholesky failed: front " + std::to_string(t) + " (" +
                     std::to_string(fdev_[t].p) + " pivots) is not positive definite at its pivot " +
                     std::to_string(h[2 * t] ? h[2 * t] : h[2 * t + 1]));
}

// ------------------------------------------------------------------ solve
void NdBottom::solve_pos(hipStream_t st) {
  const NdFrontDev* fr = static_cast<const NdFrontDev*>(d_fr_.p);
  const int2* items = static_cast<const int2*>(d_items_.p);
  const int2* pit = static_cast<const int2*>(d_piv_items_.p);
  double* w = static_cast<double*>(w_.p);
  double* u = static_cast<double*>(u_.p);
  double* y = static_cast<double*>(y_.p);
  double* x = static_cast<double*>(x_.p);
  const double* b = static_cast<const double*>(bx_.p);
  for (int d = maxdepth_; d >= 0; --d) {
    const int64_t o = lvl_off_[d], n = lvl_off_[d + 1] - o;
    if (n == 0) continue;
    k_nd_fgather<<<grid(n), 256, 0, st>>>((int)n, items + o, fr, static_cast<const int64_t*>(d_gptr_.p),
                                          static_cast<const int64_t*>(d_gidx_.p), b, u, w);
    k_nd_fgemv<<<grid(n, 4), 256, 0, st>>>((int)n, items + o, fr, static_cast<const double*>(mfs_.p), w, y, u);
  }
  for (int d = 0; d <= maxdepth_; ++d) {
    const int64_t o = lvl_off_[d], n = lvl_off_[d + 1] - o;
    const int64_t po = piv_off_[d], pn = piv_off_[d + 1] - po;
    if (n == 0) continue;
    k_nd_bgather<<<grid(n), 256, 0, st>>>((int)n, items + o, fr, static_cast<const int64_t*>(d_bidx_.p), y, x, w);
    k_nd_bgemv<<<grid(pn, 4), 256, 0, st>>>((int)pn, pit + po, fr, static_cast<const double*>(mbs_.p), w, x);
  }
}

template <class T>
void NdBottom::solve(int nb, const T* b, T* u, hipStream_t st) {
  const int* vo = static_cast<const int*>(d_var_of_.p);
  k_nd_in<T><<<grid((int64_t)nb * 64), 256, 0, st>>>(nb, vo, b, static_cast<double*>(bx_.p));
  solve_pos(st);
  k_nd_out<T><<<grid((int64_t)nb * 64), 256, 0, st>>>(nb, vo, static_cast<const double*>(x_.p), b, u);
}

template void NdBottom::solve<float>(int, const float*, float*, hipStream_t);
template void NdBottom::solve<double>(int, const double*, double*, hipStream_t);

NdStats NdBottom::stats() const {
  NdStats s;
  s.fronts = (int64_t)fr_.size();
  s.depth = maxdepth_ + 1;
  s.factor_entries = mstore_;
  for (const NdFrontDev& D : fdev_) s.max_front = std::max<int64_t>(s.max_front, D.p + D.b);
  return s;
}

}  // namespace spg
