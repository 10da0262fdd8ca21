// layout.hpp -- host-side construction of the narrow-band block hierarchy.
//
// Every level is a set of 4x4x4 node blocks (64 nodes, one wavefront) kept in
// Morton order; device arrays are structure-of-arrays per block.  The host
// part here reproduces the reference's page/flag semantics exactly:
//   * allocated nodes = nodes of SPGrid pages (4x4x8 nodes) that were set
//     (SPGrid_CPU_Solver.cpp:107-138, Multigrid_Hierarchy.h:158-195);
//   * ACTIVE_CELL, DIRICHLET_CELL propagate to the 8 nodes of the cell
//     (Flaging.h:93-142); BOUNDARY_NODE = 3 rings (Flaging.h:155-184);
//   * coarse index (x + B) >> 1 with B = (4,4,8) (SPGrid_Mask.h:183-187).
// Only blocks holding at least one ACTIVE node are shipped to the GPU: every
// other node carries zeros in all solver vectors.
#pragma once
#include <algorithm>
#include <cstdint>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

namespace spg {

// f(begin, end) over [0, n) on up to 16 threads (the ranges must be
// independent); the exception of the lowest range is rethrown after the join
template <class F>
inline void parallel_ranges(int64_t n, F&& f, int64_t min_chunk = 4096) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t nt = std::min<int64_t>(std::max<unsigned>(1u, std::min(hw, 16u)), (n + min_chunk - 1) / min_chunk);
  if (nt <= 1) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(nt);
  const int64_t chunk = (n + nt - 1) / nt;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t a = t * chunk, b = std::min(n, a + chunk);
    th.emplace_back([&, t, a, b] {
      try {
        if (a < b) f(a, b);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  }
  for (auto& x : th) x.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// reference flag bits (solver/Flags.h:10-24)
constexpr uint16_t F_ACTIVE_NODE = 0x0010;
constexpr uint16_t F_DIR_NODE_X = 0x0020;
constexpr uint16_t F_DIR_NODE = 0x00E0;
constexpr uint16_t F_ACTIVE_CELL = 0x0100;
constexpr uint16_t F_DIR_CELL_X = 0x0200;
constexpr uint16_t F_DIR_CELL = 0x0E00;
constexpr uint16_t F_BOUNDARY = 0x1000;

// compact device flag bits (one byte per node)
constexpr uint8_t D_ACTIVE = 0x01;
constexpr uint8_t D_DIR_X = 0x02;  // DIR_Y = 0x04, DIR_Z = 0x08
constexpr uint8_t D_DIR = 0x0E;
constexpr uint8_t D_BOUNDARY = 0x10;

constexpr int kPage[3] = {4, 4, 8};
constexpr int kB[3] = {4, 4, 8};

struct LayoutError : std::runtime_error {
  int code;
  LayoutError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

inline uint64_t bkey(int bx, int by, int bz) {
  return ((uint64_t)(uint32_t)(bx + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(by + (1 << 20)) << 21) |
         (uint64_t)(uint32_t)(bz + (1 << 20));
}

struct ProblemDesc {
  int32_t res[3];
  double dx, mu, lambda, penalty, min_stiffness, min_density;
  int32_t mg_levels, smooth_iters, boundary_iters;
  int64_t nblocks;
  const int32_t* block_base;
  const double* density;
  int64_t ndir;
  const int32_t* dir_coord;
  const int32_t* dir_axis;
  const double* dir_value;
};

// largest level vector (entries) the device kernels can index with int
int64_t vector_index_limit();

// distances from the level-0 boundary band are kept up to this value
constexpr int kGsDistCap = 9;

// SIMP map (topo_opt/fem_interface.h:343-353)
double stiffness_mapping(double penalty, double min_stiffness, double min_density, double d);

// Block coordinates -> id: a dense grid over a bounding box (plain array
// lookups), with a hash map for coordinates outside it or for boxes too
// large to hold densely.
struct BlockMap {
  static constexpr int64_t kMaxDense = int64_t(1) << 26;  // entries (256 MB)
  int lo[3] = {0, 0, 0}, dim[3] = {0, 0, 0};
  std::vector<int32_t> grid;
  std::unordered_map<uint64_t, int32_t> spill;
  void set_box(const int lo_[3], const int hi_[3]) {  // inclusive bounds
    int64_t vol = 1;
    for (int d = 0; d < 3; ++d) {
      lo[d] = lo_[d];
      dim[d] = hi_[d] >= lo_[d] ? hi_[d] - lo_[d] + 1 : 0;
      vol *= dim[d];
    }
    if (vol > 0 && vol <= kMaxDense) grid.assign((size_t)vol, -1);
  }
  int64_t index(int x, int y, int z) const {
    if (grid.empty()) return -1;
    const unsigned dx = (unsigned)(x - lo[0]), dy = (unsigned)(y - lo[1]), dz = (unsigned)(z - lo[2]);
    if (dx >= (unsigned)dim[0] || dy >= (unsigned)dim[1] || dz >= (unsigned)dim[2]) return -1;
    return ((int64_t)dx * dim[1] + dy) * dim[2] + dz;
  }
  int32_t find(int x, int y, int z) const {
    const int64_t i = index(x, y, z);
    if (i >= 0) return grid[(size_t)i];
    auto it = spill.find(bkey(x, y, z));
    return it == spill.end() ? -1 : it->second;
  }
  // stores id unless the coordinates already have one; returns the stored id
  int32_t insert(int x, int y, int z, int32_t id) {
    const int64_t i = index(x, y, z);
    if (i >= 0) {
      if (grid[(size_t)i] < 0) grid[(size_t)i] = id;
      return grid[(size_t)i];
    }
    return spill.emplace(bkey(x, y, z), id).first->second;
  }
};

// One level on the host while flags are being built.
struct HostLevel {
  BlockMap map;
  std::vector<int32_t> bc;       // [nb*3] block coords (block units)
  std::vector<uint8_t> alloc;    // block inside a set page
  std::vector<uint16_t> nflag;   // [nb*64] node flags
  std::vector<uint16_t> cflag;   // [nb*64] cell flags (cell at min corner)
  int32_t find(int bx, int by, int bz) const { return map.find(bx, by, bz); }
  int32_t get(int bx, int by, int bz);
  void mark_page(int x, int y, int z);  // node coords
  bool node_allocated(int x, int y, int z) const {
    int32_t b = find(x >> 2, y >> 2, z >> 2);
    return b >= 0 && alloc[b];
  }
  uint16_t node_flag(int x, int y, int z) const {
    int32_t b = find(x >> 2, y >> 2, z >> 2);
    return b < 0 ? 0 : nflag[(size_t)b * 64 + ((x & 3) << 4) + ((y & 3) << 2) + (z & 3)];
  }
};

// Final per-level GPU-facing description (host copies).
struct LevelDesc {
  int64_t nb = 0;
  std::vector<int32_t> base;      // [nb*3] node coords of the block origin
  std::vector<int32_t> nbr;       // [nb*27] neighbour block (-1)
  std::vector<uint8_t> flags;     // [nb*64] D_* bits
  std::vector<int32_t> restrict_src;  // [nb*27] fine blocks covering the restriction stencil (coarse levels)
  std::vector<int32_t> prolong_src;   // [nb*8] coarse blocks feeding this (fine) level's nodes
  std::vector<int32_t> bnd_list[8];   // boundary nodes per colour: blk*64+l
  std::vector<int32_t> dir_list;      // nodes with any Dirichlet bit: blk*64+l
  // boundary system: for boundary node i (list order, colours concatenated)
  // and spoke s = g + 8t, entry i*32 + g*4 + t holds the neighbour's index in
  // the boundary list (bnd_ci) or, for a neighbour outside it, its slot
  // blk*64+l (bnd_slot); -1 where absent or not applicable.
  std::vector<int32_t> bnd_ci, bnd_slot;
  // level-0 matrix-free boundary system: for corner I and boundary node i,
  // bnd_cslot[I*n + i] is the slot of cell (node - I) (-1 if absent) and
  // bnd_cnode[(I*n + i)*8 + J] the boundary index of that cell's node J
  // (-1 outside the boundary set)
  std::vector<int32_t> bnd_cslot;
  std::vector<int16_t> bnd_cnode;
  // level 0: distance of every slot from the boundary band (26-neighbour
  // steps, capped at kGsDistCap) and, per sweep direction (0 forward, 1
  // backward) and colour step s, the slots of that step's colour within
  // distance s + 1 of the band: the only ones whose update can depend on the
  // boundary phase that precedes the sweep (built only when SPG_OVERLAP
  // selects the overlapped first sweep)
  std::vector<uint8_t> gs_dist;
  std::vector<int32_t> gs_near[2][8];
  int64_t n_active = 0;
  bool is_explicit = false;
  // explicit levels: Galerkin data
  std::vector<int32_t> cell_row;      // [nb*64] row of the active cell at this min corner (-1)
  int64_t ncell = 0;                  // active coarse cells
  std::vector<int32_t> cell_fine_blocks;  // [ncell * (2^d/4)^3] level-0 blocks of each coarse cell (d>=2)
  std::vector<int32_t> cell_gpu_index;    // [ncell] blk*64+l of the cell's min corner
};

// An allocator that leaves elements uninitialised on resize: the per-node
// arrays of a level are hundreds of MB, and every element is written by a
// parallel loop, so a serial zero fill on allocation would only cost time.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
template <class T>
using NodeArray = std::vector<T, NoInitAlloc<T>>;

struct Layout {
  int nlev = 0;
  bool use_mg = true;
  int explicit_level = 2;
  std::vector<LevelDesc> lv;
  // level-0 cell data (per node slot = cell at min corner)
  NodeArray<double> E;            // stiffness mapping of active cells, 0 otherwise
  NodeArray<double> rho;          // input density (0 if none)
  // input block nodes -> level-0 slot (blk*64+l) or -1
  int64_t nin_blocks = 0;
  NodeArray<int32_t> in_node_slot;   // [nin_blocks*64]
  // Dirichlet values u_D at level-0 slots, and at input nodes outside the GPU set
  std::vector<int32_t> ud_slot;        // slot*3+axis, each at most once
  std::vector<double> ud_value;
  // input nodes outside the GPU blocks that carry a Dirichlet value:
  // (input node index)*3+axis and value, in input order (last write wins)
  std::vector<int64_t> ud_extra;
  std::vector<double> ud_extra_value;
  int64_t n_active_cells = 0;
  int64_t n_allocated_pages = 0;
};

// Builds the layout.  Throws LayoutError with the reference's failure modes.
Layout build_layout(const ProblemDesc& p);

// Position matrices P_p^T K0 P_p for a depth-d coarse cell: [8^d][576],
// q = (px*2^d + py)*2^d + pz, entry ((A*8+B)*3+v)*3+w.
std::vector<double> galerkin_position_matrices(int depth, const double* K0);

// Element tables K_mu, K_la [8][8][3][3] of the unit Q1 hex (node id x*4+y*2+z).
void element_tables(double* kmu, double* kla);

}  // namespace spg
