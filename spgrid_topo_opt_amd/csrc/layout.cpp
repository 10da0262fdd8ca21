// layout.cpp -- see layout.hpp.
#include "layout.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <array>
#include <cmath>
#include <exception>
#include <thread>
#include <cstdio>
#include <chrono>
#include <unordered_set>

namespace spg {

double stiffness_mapping(double penalty, double min_stiffness, double min_density, double d) {
  // topo_opt/fem_interface.h:343-353
  if (penalty != 0.0) return (1.0 - min_stiffness) * std::pow(std::max(min_density, d), penalty) + min_stiffness;
  return d;
}

// ------------------------------------------------------------- element tables
// Q1 hex on the unit cube, exact integrals as tensor products of the 1-D
// integrals of the linear shape functions phi_0 = 1-t, phi_1 = t:
//   mass m_ab = int phi_a phi_b, stiffness k_ab = int phi'_a phi'_b,
//   mixed g_ab = int phi'_a phi_b.
// K_mu[a][b][v][w] = delta_vw sum_d int dN_a/dx_d dN_b/dx_d + int dN_a/dx_w dN_b/dx_v
// K_la[a][b][v][w] = int dN_a/dx_v dN_b/dx_w
// (values match solver/CPU_Kernels/Stiffness_Matrices.h:8-688).
static double grad_product(int a, int b, int p, int q) {
  static const double m[2][2] = {{1.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 3.0}};
  static const double k[2][2] = {{1.0, -1.0}, {-1.0, 1.0}};
  static const double g[2][2] = {{-0.5, -0.5}, {0.5, 0.5}};  // g[a][b] = int phi'_a phi_b
  double r = 1.0;
  for (int d = 0; d < 3; ++d) {
    int ad = (a >> (2 - d)) & 1, bd = (b >> (2 - d)) & 1;
    if (d == p && d == q) r *= k[ad][bd];
    else if (d == p) r *= g[ad][bd];
    else if (d == q) r *= g[bd][ad];
    else r *= m[ad][bd];
  }
  return r;
}

void element_tables(double* kmu, double* kla) {
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w) {
          double lap = 0.0;
          if (v == w) lap = grad_product(a, b, 0, 0) + grad_product(a, b, 1, 1) + grad_product(a, b, 2, 2);
          int e = ((a * 8 + b) * 3 + v) * 3 + w;
          kmu[e] = lap + grad_product(a, b, w, v);
          kla[e] = grad_product(a, b, v, w);
        }
}

std::vector<double> galerkin_position_matrices(int depth, const double* K0) {
  const int n = 1 << depth;
  const int64_t npos = (int64_t)n * n * n;
  std::vector<double> out((size_t)npos * 576);
  for (int64_t q = 0; q < npos; ++q) {
    int p[3] = {(int)(q / (n * n)), (int)((q / n) % n), (int)(q % n)};
    double W[8][8];  // weight of coarse node A at fine node a
    for (int a = 0; a < 8; ++a)
      for (int A = 0; A < 8; ++A) {
        double w = 1.0;
        for (int d = 0; d < 3; ++d) {
          double t = (double)(p[d] + ((a >> (2 - d)) & 1)) / (double)n;
          w *= ((A >> (2 - d)) & 1) ? t : 1.0 - t;
        }
        W[a][A] = w;
      }
    // T[a][B][v][w] = sum_b K[a][b][v][w] W[b][B]
    double T[8][8][9];
    for (int a = 0; a < 8; ++a)
      for (int B = 0; B < 8; ++B)
        for (int e = 0; e < 9; ++e) {
          double s = 0.0;
          for (int b = 0; b < 8; ++b) s += K0[(a * 8 + b) * 9 + e] * W[b][B];
          T[a][B][e] = s;
        }
    double* M = out.data() + q * 576;
    for (int A = 0; A < 8; ++A)
      for (int B = 0; B < 8; ++B)
        for (int e = 0; e < 9; ++e) {
          double s = 0.0;
          for (int a = 0; a < 8; ++a) s += W[a][A] * T[a][B][e];
          M[(A * 8 + B) * 9 + e] = s;
        }
  }
  return out;
}

// ------------------------------------------------------------- host levels
int32_t HostLevel::get(int bx, int by, int bz) {
  const int32_t id = (int32_t)alloc.size();
  const int32_t got = map.insert(bx, by, bz, id);
  if (got != id) return got;
  bc.push_back(bx);
  bc.push_back(by);
  bc.push_back(bz);
  alloc.push_back(0);
  nflag.resize(nflag.size() + 64, 0);
  cflag.resize(cflag.size() + 64, 0);
  return id;
}

void HostLevel::mark_page(int x, int y, int z) {
  // a 4x4x8 page = two 4x4x4 blocks stacked in z
  int bx = x >> 2, by = y >> 2, bz0 = (z >> 3) * 2;
  int32_t b0 = get(bx, by, bz0);
  alloc[b0] = 1;
  int32_t b1 = get(bx, by, bz0 + 1);
  alloc[b1] = 1;
}

static inline int loc(int x, int y, int z) { return ((x & 3) << 4) | ((y & 3) << 2) | (z & 3); }

// Flag_Nodes (Flaging.h:127-142): every node takes the ACTIVE / DIRICHLET
// bits (>>4) of the cells of allocated blocks that have it as a corner.  The
// blocks that receive bits are created first (serially: a cell reaches the +
// neighbours on the axes where it sits on the block's upper face); the bits
// are then gathered node by node, so blocks are independent.
static void flag_nodes(HostLevel& H) {
  const size_t nb0 = H.alloc.size();
  // per block, the + neighbours (bit o) it reaches that do not exist yet;
  // they are then created in block order, as a serial pass would
  std::vector<uint8_t> missing(nb0, 0);
  parallel_ranges((int64_t)nb0, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      if (!H.alloc[b]) continue;
      unsigned upper = 0;  // axes (x 4, y 2, z 1) on whose upper face a flagged cell sits
      bool any = false;
      for (int l = 0; l < 64; ++l) {
        if (!(H.cflag[b * 64 + l] & (F_ACTIVE_CELL | F_DIR_CELL))) continue;
        any = true;
        upper |= (((l >> 4) == 3) << 2) | ((((l >> 2) & 3) == 3) << 1) | ((l & 3) == 3);
      }
      if (!any) continue;
      const int bx = H.bc[3 * b], by = H.bc[3 * b + 1], bz = H.bc[3 * b + 2];
      for (unsigned o = 1; o < 8; ++o)
        if ((o & upper) == o && H.find(bx + ((o >> 2) & 1), by + ((o >> 1) & 1), bz + (o & 1)) < 0)
          missing[b] |= (uint8_t)(1u << o);
    }
  });
  for (size_t b = 0; b < nb0; ++b) {
    if (!missing[b]) continue;
    const int bx = H.bc[3 * b], by = H.bc[3 * b + 1], bz = H.bc[3 * b + 2];
    for (unsigned o = 1; o < 8; ++o)
      if (missing[b] & (1u << o)) H.get(bx + ((o >> 2) & 1), by + ((o >> 1) & 1), bz + (o & 1));
  }
  const int64_t nb1 = (int64_t)H.alloc.size();
  parallel_ranges(nb1, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      const int bx = H.bc[3 * b], by = H.bc[3 * b + 1], bz = H.bc[3 * b + 2];
      int32_t src[8];  // this block and its - neighbours (bit set: shifted by -1 on that axis), if allocated
      bool any = false;
      for (int o = 0; o < 8; ++o) {
        const int32_t t = o == 0 ? (int32_t)b : H.find(bx - ((o >> 2) & 1), by - ((o >> 1) & 1), bz - (o & 1));
        src[o] = (t >= 0 && H.alloc[t]) ? t : -1;
        any |= src[o] >= 0;
      }
      if (!any) continue;
      for (int l = 0; l < 64; ++l) {
        const int i = l >> 4, j = (l >> 2) & 3, k = l & 3;
        uint16_t acc = 0;
        for (int c = 0; c < 8; ++c) {  // the cell whose corner c is this node
          const int ci = i - ((c >> 2) & 1), cj = j - ((c >> 1) & 1), ck = k - (c & 1);
          const int32_t t = src[((ci < 0) << 2) | ((cj < 0) << 1) | (ck < 0)];
          if (t >= 0) acc |= H.cflag[(size_t)t * 64 + loc(ci, cj, ck)];
        }
        H.nflag[(size_t)b * 64 + l] |= (uint16_t)((acc & (F_ACTIVE_CELL | F_DIR_CELL)) >> 4);
      }
    }
  });
}

// Flag_Boundary_Nodes with band 3 (Flaging.h:155-184), evaluated as a
// frontier expansion: in each pass an allocated ACTIVE node without Dirichlet
// bits becomes BOUNDARY if one of its 26 neighbours carried DIRICHLET or
// BOUNDARY bits at the start of the pass.
static void flag_boundary(HostLevel& H) {
  std::vector<std::array<int, 3>> seeds;
  for (size_t b = 0; b < H.alloc.size(); ++b)
    for (int l = 0; l < 64; ++l)
      if (H.nflag[b * 64 + l] & (F_DIR_NODE | F_BOUNDARY))
        seeds.push_back({H.bc[3 * b] * 4 + (l >> 4), H.bc[3 * b + 1] * 4 + ((l >> 2) & 3), H.bc[3 * b + 2] * 4 + (l & 3)});
  for (int pass = 0; pass < 3; ++pass) {
    std::vector<std::array<int, 3>> add;
    std::unordered_set<uint64_t> seen;
    for (auto& s : seeds)
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            int x = s[0] + dx, y = s[1] + dy, z = s[2] + dz;
            int32_t b = H.find(x >> 2, y >> 2, z >> 2);
            if (b < 0 || !H.alloc[b]) continue;
            uint16_t f = H.nflag[(size_t)b * 64 + loc(x, y, z)];
            if ((f & F_DIR_NODE) || !(f & F_ACTIVE_NODE) || (f & F_BOUNDARY)) continue;
            uint64_t k = bkey(x, y, z);
            if (!seen.insert(k).second) continue;
            add.push_back({x, y, z});
          }
    for (auto& a : add) {
      int32_t b = H.find(a[0] >> 2, a[1] >> 2, a[2] >> 2);
      H.nflag[(size_t)b * 64 + loc(a[0], a[1], a[2])] |= F_BOUNDARY;
    }
    seeds = std::move(add);
  }
}

static uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
  };
  return (spread(x) << 2) | (spread(y) << 1) | spread(z);
}

int64_t vector_index_limit() {
  // SPG_INDEX_LIMIT lowers the limit so that tests can reach the check
  // with small problems
  if (const char* e = std::getenv("SPG_INDEX_LIMIT")) {
    const long long v = std::atoll(e);
    if (v > 0 && v < INT32_MAX) return v;
  }
  return INT32_MAX;
}

// colour of step i of a level-0 sweep: 0..7 forward, 7..0 backward (the
// order of SPGrid_Gauss_Seidel.h:512-530; sweep_colour in kernels.hpp)
static int sweep_colour0(bool forward, int i) { return forward ? i : 7 - i; }

// Level-0 slots' distance from the boundary band and the near-band lists of
// every colour step.  A sweep's step-s update at a slot farther than s + 1
// from the band reads only values that the boundary phase before the sweep
// does not write, directly or through earlier steps of the same sweep
// (each earlier step brings the dependence at most one node closer), so the
// solver may run those updates while the boundary phase is still going.
static void band_distances(Layout& out) {
  if (out.lv.empty()) return;
  LevelDesc& D = out.lv[0];
  const int64_t ns = D.nb * 64;
  D.gs_dist.assign(ns, (uint8_t)kGsDistCap);
  std::vector<int32_t> front;
  for (int c = 0; c < 8; ++c)
    for (int32_t s : D.bnd_list[c]) {
      D.gs_dist[s] = 0;
      front.push_back(s);
    }
  for (int d = 1; d < kGsDistCap && !front.empty(); ++d) {
    std::vector<int32_t> next;
    for (int32_t s : front) {
      const int64_t g = s >> 6;
      const int q = s & 63, li = q >> 4, lj = (q >> 2) & 3, lk = q & 3;
      for (int sp = 0; sp < 27; ++sp) {
        if (sp == 13) continue;
        const int a = li + sp / 9 - 1, b = lj + (sp / 3) % 3 - 1, c = lk + sp % 3 - 1;
        const int rs = (a < 0 ? 0 : (a > 3 ? 2 : 1)) * 9 + (b < 0 ? 0 : (b > 3 ? 2 : 1)) * 3 + (c < 0 ? 0 : (c > 3 ? 2 : 1));
        const int32_t nb = rs == 13 ? (int32_t)g : D.nbr[(size_t)g * 27 + rs];
        if (nb < 0) continue;
        const int32_t t = nb * 64 + (((a & 3) << 4) | ((b & 3) << 2) | (c & 3));
        if (D.gs_dist[t] > d) {
          D.gs_dist[t] = (uint8_t)d;
          next.push_back(t);
        }
      }
    }
    front.swap(next);
  }
  for (int dir = 0; dir < 2; ++dir)
    for (int st = 0; st < 8; ++st) {
      const int colour = sweep_colour0(dir == 0, st);
      std::vector<int32_t>& L = D.gs_near[dir][st];
      L.clear();
      if (D.gs_dist.empty()) continue;
      for (int64_t g = 0; g < D.nb; ++g)
        for (int k = 0; k < 8; ++k) {
          const int q = ((((colour >> 2) & 1) + 2 * ((k >> 2) & 1)) << 4) |
                        ((((colour >> 1) & 1) + 2 * ((k >> 1) & 1)) << 2) | ((colour & 1) + 2 * (k & 1));
          const int32_t s = (int32_t)(g * 64 + q);
          if (D.gs_dist[s] <= st + 1) L.push_back(s);
        }
    }
}

// SPG_SETUP_TRACE=2: seconds per phase of build_layout on stderr
namespace {
struct PhaseTrace {
  bool on = false;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  PhaseTrace() {
    const char* e = std::getenv("SPG_SETUP_TRACE");
    on = e && std::atoi(e) >= 2;
  }
  void operator()(const char* what) {
    if (!on) return;
    const auto n = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[spg] layout %-28s %.3f s\n", what, std::chrono::duration<double>(n - t).count());
    t = n;
  }
};
}  // namespace

Layout build_layout(const ProblemDesc& p) {
  PhaseTrace trace;
  Layout out;
  if (p.nblocks <= 0) throw LayoutError("no density blocks", -1);
  if (!(p.dx > 0)) throw LayoutError("dx must be positive", -1);
  if (!(p.mu > 0) || p.lambda < 0) throw LayoutError("global_mu must be > 0 and global_lambda >= 0", -1);
  int L = p.mg_levels;
  out.use_mg = L != 0;
  out.nlev = L == 0 ? 1 : L;
  if (L < 0 || L > 12) throw LayoutError("mg_level out of range", -1);
  if (out.nlev == 2)
    throw LayoutError(
        "mg_level=2 is not supported: level 1 is a pass-through level without a matrix "
        "(Multigrid_Hierarchy.h:215-226); use 1 or >=3",
        -1);
  std::vector<HostLevel> H(out.nlev);

  // ---- level 0: pages of the input blocks (SPGrid_CPU_Solver.cpp:107-138)
  HostLevel& H0 = H[0];
  // block boxes: level 0 from the input blocks (pages add at most one block
  // in z), each coarse level from the map (x + B) >> 1 of the finer box; a
  // margin of 2 blocks, anything outside falls back to hashing
  std::vector<std::array<int, 6>> box(out.nlev);
  {
    std::array<int, 6> b = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
    for (int64_t i = 0; i < p.nblocks; ++i)
      for (int d = 0; d < 3; ++d) {
        b[d] = std::min(b[d], p.block_base[3 * i + d] >> 2);
        b[3 + d] = std::max(b[3 + d], p.block_base[3 * i + d] >> 2);
      }
    for (int d = 0; d < 3; ++d) {
      b[d] -= 2;
      b[3 + d] += 2;
    }
    box[0] = b;
    for (int l = 1; l < out.nlev; ++l)
      for (int d = 0; d < 3; ++d) {
        box[l][d] = (((box[l - 1][d] * 4 + kB[d]) >> 1) >> 2) - 2;
        box[l][3 + d] = (((box[l - 1][3 + d] * 4 + 3 + kB[d]) >> 1) >> 2) + 2;
      }
    for (int l = 0; l < out.nlev; ++l) H[l].map.set_box(box[l].data(), box[l].data() + 3);
  }
  // a page holds two blocks: room for twice the input blocks
  H0.bc.reserve((size_t)p.nblocks * 6);
  H0.alloc.reserve((size_t)p.nblocks * 2);
  H0.nflag.reserve((size_t)p.nblocks * 128);
  H0.cflag.reserve((size_t)p.nblocks * 128);
  for (int64_t b = 0; b < p.nblocks; ++b) {
    const int32_t* bb = p.block_base + 3 * b;
    H0.mark_page(bb[0], bb[1], bb[2]);
  }
  trace("level-0 pages");
  // every page marked so far holds two allocated blocks: its lower one (even
  // z) stands for it
  {
    int64_t npages = 0;
    bool bad = false;
    uint64_t first_bad = ~0ull;  // the violation reported: the smallest page key, as in key order
    for (size_t b = 0; b < H0.alloc.size(); ++b) {
      if (!H0.alloc[b] || (H0.bc[3 * b + 2] & 1)) continue;
      ++npages;
      // padding check (Multigrid_Hierarchy.h:145-157)
      const int o[3] = {H0.bc[3 * b] * 4, H0.bc[3 * b + 1] * 4, H0.bc[3 * b + 2] * 4};
      for (int d = 0; d < 3; ++d)
        if (o[d] < kB[d] || o[d] >= p.res[d] - 2 * kB[d]) {
          bad = true;
          first_bad = std::min(first_bad, bkey(o[0] >> 2, o[1] >> 2, o[2] >> 3));
        }
    }
    out.n_allocated_pages = npages;
    if (bad) {
      const uint64_t k = first_bad;
      const int pc[3] = {(int)((k >> 42) & 0x1fffff) - (1 << 20), (int)((k >> 21) & 0x1fffff) - (1 << 20),
                         (int)(k & 0x1fffff) - (1 << 20)};
      throw LayoutError("padding violated: page origin (" + std::to_string(pc[0] * 4) + "," +
                            std::to_string(pc[1] * 4) + "," + std::to_string(pc[2] * 8) + ") outside [B, res-2B)",
                        -2);
    }
  }
  trace("padding check");
  // cell moduli (SIMP) and ACTIVE_CELL (Flaging.h:93-106)
  // moduli per input cell (0 where the density is not positive or the cell
  // lies outside the set pages); the level-0 vectors take them at the end
  NodeArray<double> Ein((size_t)p.nblocks * 64);
  parallel_ranges(p.nblocks, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      const int32_t* bb = p.block_base + 3 * b;
      // a 4-aligned input block (the caller's layout) lies in one level-0 block
      const bool aligned = ((bb[0] | bb[1] | bb[2]) & 3) == 0;
      const int32_t hb0 = aligned ? H0.find(bb[0] >> 2, bb[1] >> 2, bb[2] >> 2) : -1;
      double* eb = Ein.data() + b * 64;
      std::fill(eb, eb + 64, 0.0);
      if (aligned && (hb0 < 0 || !H0.alloc[hb0])) continue;
      for (int l = 0; l < 64; ++l) {
        int x = bb[0] + (l >> 4), y = bb[1] + ((l >> 2) & 3), z = bb[2] + (l & 3);
        const int32_t hb = aligned ? hb0 : H0.find(x >> 2, y >> 2, z >> 2);
        if (hb < 0 || !H0.alloc[hb]) continue;  // cells outside set pages are dropped
        double d = p.density[b * 64 + l];
        if (std::isnan(d)) throw LayoutError("density contains NaN", -6);
        if (d > 0) {
          double e = stiffness_mapping(p.penalty, p.min_stiffness, p.min_density, d);
          eb[l] = e;
          if (e * p.mu != 0.0) H0.cflag[(size_t)hb * 64 + loc(x, y, z)] |= F_ACTIVE_CELL;
        }
      }
    }
  });
  trace("cell moduli");
  // block activation check (Validate_Blocks.h:46-69); a cell's nodes lie in
  // its own block or its + neighbours, whose allocation is looked up once
  parallel_ranges((int64_t)H0.alloc.size(), [&](int64_t b0, int64_t b1) {
  for (int64_t b = b0; b < b1; ++b) {
    int8_t al8[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    for (int l = 0; l < 64; ++l) {
      if (!(H0.cflag[b * 64 + l] & F_ACTIVE_CELL)) continue;
      int x = H0.bc[3 * b] * 4 + (l >> 4), y = H0.bc[3 * b + 1] * 4 + ((l >> 2) & 3), z = H0.bc[3 * b + 2] * 4 + (l & 3);
      for (int c = 1; c < 8; ++c) {
        const int X = x + ((c >> 2) & 1), Y = y + ((c >> 1) & 1), Z = z + (c & 1);
        const int o = (((X >> 2) - H0.bc[3 * b]) << 2) | (((Y >> 2) - H0.bc[3 * b + 1]) << 1) | ((Z >> 2) - H0.bc[3 * b + 2]);
        if (al8[o] < 0) al8[o] = H0.node_allocated(X, Y, Z) ? 1 : 0;
        if (!al8[o])
          throw LayoutError("block activation check failed at cell (" + std::to_string(x) + "," + std::to_string(y) +
                                "," + std::to_string(z) + "): its nodes are not all in allocated blocks",
                            -2);
      }
    }
  }
  });
  trace("activation check");
  // Dirichlet nodes -> Dirichlet cells (SPGrid_CPU_Solver.cpp:147-192)
  {
    std::unordered_set<uint64_t> dset[3];
    for (int64_t i = 0; i < p.ndir; ++i) {
      int v = p.dir_axis[i];
      if (v < 0 || v > 2) throw LayoutError("Dirichlet axis out of range", -1);
      dset[v].insert(bkey(p.dir_coord[3 * i], p.dir_coord[3 * i + 1], p.dir_coord[3 * i + 2]));
    }
    for (int64_t i = 0; i < p.ndir; ++i) {
      int v = p.dir_axis[i];
      int x = p.dir_coord[3 * i], y = p.dir_coord[3 * i + 1], z = p.dir_coord[3 * i + 2];
      bool all = true;
      for (int c = 1; c < 8 && all; ++c)
        all = dset[v].count(bkey(x + ((c >> 2) & 1), y + ((c >> 1) & 1), z + (c & 1))) != 0;
      if (!all) continue;
      int32_t hb = H0.find(x >> 2, y >> 2, z >> 2);
      if (hb < 0 || !H0.alloc[hb]) continue;
      H0.cflag[(size_t)hb * 64 + loc(x, y, z)] |= (uint16_t)(F_DIR_CELL_X << v);
    }
  }
  trace("Dirichlet cells");
  // ---- coarse levels (Multigrid_Hierarchy.h:158-195)
  for (int l = 1; l < out.nlev; ++l) {
    HostLevel& F = H[l - 1];
    HostLevel& C = H[l];
    // Fine cell x maps to coarse cell (x + B) >> 1: the cells of one fine
    // block land in a 2x2x2 group of cells of a single coarse block.  The
    // coarse blocks are created first, then every coarse cell gathers the
    // ACTIVE / DIRICHLET bits of its 8 fine cells x = 2X - B + {0, 1}.
    const size_t nbf = F.alloc.size();
    std::vector<uint8_t> flagged(nbf, 0);
    parallel_ranges((int64_t)nbf, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b) {
        if (!F.alloc[b]) continue;
        bool any = false;
        for (int q = 0; q < 64 && !any; ++q) any = (F.cflag[b * 64 + q] & (F_ACTIVE_CELL | F_DIR_CELL)) != 0;
        flagged[b] = any;
      }
    });
    for (size_t b = 0; b < nbf; ++b) {
      if (!flagged[b]) continue;
      int cb[3];
      for (int d = 0; d < 3; ++d) cb[d] = ((F.bc[3 * b + d] * 4 + kB[d]) >> 1) >> 2;
      C.get(cb[0], cb[1], cb[2]);
    }
    const int64_t nbc = (int64_t)C.alloc.size();
    parallel_ranges(nbc, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b) {
        int f0[3];  // first fine block under coarse block b: fine x from 2 * 4 cb - B
        for (int d = 0; d < 3; ++d) f0[d] = (2 * 4 * C.bc[3 * b + d] - kB[d]) >> 2;
        int32_t fb[8];
        for (int o = 0; o < 8; ++o) {
          const int32_t t = F.find(f0[0] + ((o >> 2) & 1), f0[1] + ((o >> 1) & 1), f0[2] + (o & 1));
          fb[o] = (t >= 0 && F.alloc[t]) ? t : -1;
        }
        for (int q = 0; q < 64; ++q) {
          const int X = C.bc[3 * b] * 4 + (q >> 4), Y = C.bc[3 * b + 1] * 4 + ((q >> 2) & 3), Z = C.bc[3 * b + 2] * 4 + (q & 3);
          uint16_t acc = 0;
          for (int c = 0; c < 8; ++c) {
            const int x = 2 * X - kB[0] + ((c >> 2) & 1), y = 2 * Y - kB[1] + ((c >> 1) & 1), z = 2 * Z - kB[2] + (c & 1);
            const int o = (((x >> 2) - f0[0]) << 2) | (((y >> 2) - f0[1]) << 1) | ((z >> 2) - f0[2]);
            if (fb[o] >= 0) acc |= F.cflag[(size_t)fb[o] * 64 + loc(x, y, z)];
          }
          C.cflag[(size_t)b * 64 + q] |= (uint16_t)(acc & (F_ACTIVE_CELL | F_DIR_CELL));
        }
      }
    });
    // pages of the coarse cells' nodes: all 8 for an ACTIVE cell, 1..7 for a
    // DIRICHLET one.  A cell's corner c lies in the block shifted by c & m,
    // m = the axes on whose upper face the cell sits; per block the reached
    // shifts are collected, then their pages are marked.
    std::vector<uint8_t> reached((size_t)nbc, 0);
    parallel_ranges(nbc, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b) {
        unsigned reach = 0;
        for (int q = 0; q < 64; ++q) {
          const uint16_t cf = C.cflag[(size_t)b * 64 + q];
          if (!(cf & (F_ACTIVE_CELL | F_DIR_CELL))) continue;
          const unsigned m = (((q >> 4) == 3) << 2) | ((((q >> 2) & 3) == 3) << 1) | ((q & 3) == 3);
          for (unsigned c = (cf & F_ACTIVE_CELL) ? 0 : 1; c < 8; ++c) reach |= 1u << (c & m);
        }
        reached[(size_t)b] = (uint8_t)reach;
      }
    });
    for (int64_t b = 0; b < nbc; ++b) {
      const unsigned reach = reached[(size_t)b];
      for (unsigned o = 0; o < 8; ++o)
        if (reach & (1u << o))
          C.mark_page((C.bc[3 * b] + ((o >> 2) & 1)) * 4, (C.bc[3 * b + 1] + ((o >> 1) & 1)) * 4,
                      (C.bc[3 * b + 2] + (o & 1)) * 4);
    }
  }
  trace("coarse levels");
  for (int l = 0; l < out.nlev; ++l) flag_nodes(H[l]);
  trace("node flags");
  for (int l = 0; l < out.nlev; ++l) flag_boundary(H[l]);

  trace("boundary flags");
  // ---- GPU block sets: allocated blocks with an ACTIVE node, Morton order
  std::vector<BlockMap> gmap(out.nlev);
  for (int l = 0; l < out.nlev; ++l) gmap[l].set_box(box[l].data(), box[l].data() + 3);
  out.lv.resize(out.nlev);
  for (int l = 0; l < out.nlev; ++l) {
    HostLevel& Hl = H[l];
    std::vector<std::pair<uint64_t, int32_t>> sel;
    for (size_t b = 0; b < Hl.alloc.size(); ++b) {
      if (!Hl.alloc[b]) continue;
      bool act = false;
      for (int q = 0; q < 64 && !act; ++q) act = (Hl.nflag[b * 64 + q] & F_ACTIVE_NODE) != 0;
      if (!act) continue;
      sel.push_back({morton3((uint32_t)(Hl.bc[3 * b] + 4096), (uint32_t)(Hl.bc[3 * b + 1] + 4096),
                             (uint32_t)(Hl.bc[3 * b + 2] + 4096)),
                     (int32_t)b});
    }
    std::sort(sel.begin(), sel.end());
    LevelDesc& D = out.lv[l];
    D.nb = (int64_t)sel.size();
    D.base.resize(D.nb * 3);
    D.flags.assign(D.nb * 64, 0);
    for (int64_t g = 0; g < D.nb; ++g) {
      int32_t b = sel[g].second;
      for (int d = 0; d < 3; ++d) D.base[3 * g + d] = Hl.bc[3 * b + d] * 4;
      gmap[l].insert(Hl.bc[3 * b], Hl.bc[3 * b + 1], Hl.bc[3 * b + 2], (int32_t)g);
    }
    // device flags, boundary and Dirichlet lists: per range of blocks, the
    // ranges' lists joined in block order
    {
      struct Part {
        std::vector<int32_t> bnd[8], dir;
        int64_t act = 0;
      };
      std::vector<Part> parts(64);
      const int64_t nper = (D.nb + 63) / 64;
      parallel_ranges(64, [&](int64_t r0, int64_t r1) {
        for (int64_t r = r0; r < r1; ++r) {
          Part& P = parts[(size_t)r];
          const int64_t g1 = std::min(D.nb, (r + 1) * nper);
          for (int64_t g = r * nper; g < g1; ++g) {
            const int32_t b = sel[(size_t)g].second;
            for (int q = 0; q < 64; ++q) {
              const uint16_t f = Hl.nflag[(size_t)b * 64 + q];
              uint8_t df = 0;
              if (f & F_ACTIVE_NODE) { df |= D_ACTIVE; P.act++; }
              if (f & 0x20) df |= 0x02;
              if (f & 0x40) df |= 0x04;
              if (f & 0x80) df |= 0x08;
              if (f & F_BOUNDARY) df |= D_BOUNDARY;
              D.flags[g * 64 + q] = df;
              if ((df & D_BOUNDARY)) {
                const int x = D.base[3 * g] + (q >> 4), y = D.base[3 * g + 1] + ((q >> 2) & 3), z = D.base[3 * g + 2] + (q & 3);
                P.bnd[((x & 1) << 2) | ((y & 1) << 1) | (z & 1)].push_back((int32_t)(g * 64 + q));
              }
              if (df & D_DIR) P.dir.push_back((int32_t)(g * 64 + q));
            }
          }
        }
      }, 1);
      for (const Part& P : parts) {
        D.n_active += P.act;
        for (int c = 0; c < 8; ++c) D.bnd_list[c].insert(D.bnd_list[c].end(), P.bnd[c].begin(), P.bnd[c].end());
        D.dir_list.insert(D.dir_list.end(), P.dir.begin(), P.dir.end());
      }
    }
    if (D.nb == 0) throw LayoutError("level " + std::to_string(l) + " has no active nodes", -1);
    // the device kernels index level vectors ([nb][3][64]) with 32-bit ints
    if (D.nb * 192 > vector_index_limit())
      throw LayoutError("level " + std::to_string(l) + " holds " + std::to_string(D.nb) +
                            " blocks: its vectors (192 entries per block) exceed the 32-bit index range of the "
                            "device kernels (at most " + std::to_string(vector_index_limit() / 192) + " blocks)",
                        -1);
    auto gfind = [&](int bx, int by, int bz) { return gmap[l].find(bx, by, bz); };
    D.nbr.resize(D.nb * 27);
    parallel_ranges(D.nb, [&](int64_t g0, int64_t g1) {
      for (int64_t g = g0; g < g1; ++g) {
        int bx = D.base[3 * g] >> 2, by = D.base[3 * g + 1] >> 2, bz = D.base[3 * g + 2] >> 2;
        for (int s = 0; s < 27; ++s) D.nbr[g * 27 + s] = gfind(bx + s / 9 - 1, by + (s / 3) % 3 - 1, bz + s % 3 - 1);
      }
    });
  }
  trace("block sets, neighbours");
  // boundary neighbour tables
  for (int l = 0; l < out.nlev; ++l) {
    LevelDesc& D = out.lv[l];
    std::vector<int32_t> all;
    for (int c = 0; c < 8; ++c) all.insert(all.end(), D.bnd_list[c].begin(), D.bnd_list[c].end());
    std::unordered_map<int32_t, int32_t> pos;
    pos.reserve(all.size() * 2);
    for (size_t i = 0; i < all.size(); ++i) pos.emplace(all[i], (int32_t)i);
    D.bnd_ci.assign(all.size() * 32, -1);
    D.bnd_slot.assign(all.size() * 32, -1);
    for (size_t i = 0; i < all.size(); ++i) {
      const int32_t g = all[i] >> 6, q = all[i] & 63;
      const int li = q >> 4, lj = (q >> 2) & 3, lk = q & 3;
      for (int sp = 0; sp < 27; ++sp) {
        const int a = li + sp / 9 - 1, b = lj + (sp / 3) % 3 - 1, c = lk + sp % 3 - 1;
        const int rs = (a < 0 ? 0 : (a > 3 ? 2 : 1)) * 9 + (b < 0 ? 0 : (b > 3 ? 2 : 1)) * 3 + (c < 0 ? 0 : (c > 3 ? 2 : 1));
        const int32_t nb = D.nbr[(size_t)g * 27 + rs];
        if (nb < 0) continue;
        const int32_t slot = nb * 64 + (((a & 3) << 4) | ((b & 3) << 2) | (c & 3));
        const size_t e = i * 32 + (size_t)(sp & 7) * 4 + (sp >> 3);
        auto it = pos.find(slot);
        if (it != pos.end()) D.bnd_ci[e] = it->second;
        else D.bnd_slot[e] = slot;
      }
    }
    if (l == 0 && !all.empty() && all.size() < 32768) {
      const size_t n = all.size();
      D.bnd_cslot.assign(8 * n, -1);
      D.bnd_cnode.assign(64 * n, (int16_t)-1);
      auto slot_of = [&](int32_t g, int a, int b, int c) -> int32_t {
        const int rs = (a < 0 ? 0 : (a > 3 ? 2 : 1)) * 9 + (b < 0 ? 0 : (b > 3 ? 2 : 1)) * 3 + (c < 0 ? 0 : (c > 3 ? 2 : 1));
        const int32_t nb = D.nbr[(size_t)g * 27 + rs];
        return nb < 0 ? -1 : nb * 64 + (((a & 3) << 4) | ((b & 3) << 2) | (c & 3));
      };
      for (size_t i = 0; i < n; ++i) {
        const int32_t g = all[i] >> 6, q = all[i] & 63;
        const int li = q >> 4, lj = (q >> 2) & 3, lk = q & 3;
        for (int I = 0; I < 8; ++I) {
          const int cx = li - ((I >> 2) & 1), cy = lj - ((I >> 1) & 1), cz = lk - (I & 1);
          D.bnd_cslot[(size_t)I * n + i] = slot_of(g, cx, cy, cz);
          for (int J = 0; J < 8; ++J) {
            const int32_t sl = slot_of(g, cx + ((J >> 2) & 1), cy + ((J >> 1) & 1), cz + (J & 1));
            if (sl < 0) continue;
            auto it = pos.find(sl);
            if (it != pos.end()) D.bnd_cnode[((size_t)I * n + i) * 8 + J] = (int16_t)it->second;
          }
        }
      }
    }
  }
  trace("boundary tables");
  // transfers
  for (int l = 0; l < out.nlev; ++l) {
    LevelDesc& D = out.lv[l];
    if (l >= 1) {
      D.restrict_src.resize(D.nb * 27);
      for (int64_t g = 0; g < D.nb; ++g) {
        int f0[3];
        for (int d = 0; d < 3; ++d) f0[d] = (2 * D.base[3 * g + d] - kB[d]) >> 2;
        for (int s = 0; s < 27; ++s) {
          D.restrict_src[g * 27 + s] = gmap[l - 1].find(f0[0] + s / 9 - 1, f0[1] + (s / 3) % 3 - 1, f0[2] + s % 3 - 1);
        }
      }
    }
    if (l + 1 < out.nlev) {
      D.prolong_src.resize(D.nb * 8);
      for (int64_t g = 0; g < D.nb; ++g) {
        int c0[3];
        for (int d = 0; d < 3; ++d) c0[d] = ((D.base[3 * g + d] + kB[d]) >> 1) >> 2;
        for (int s = 0; s < 8; ++s) {
          D.prolong_src[g * 8 + s] = gmap[l + 1].find(c0[0] + ((s >> 2) & 1), c0[1] + ((s >> 1) & 1), c0[2] + (s & 1));
        }
      }
    }
  }
  trace("transfer tables");
  // explicit levels (Multigrid_Hierarchy.h:215-226): cells and Galerkin tables
  int first_explicit = out.nlev == 1 ? 0 : out.explicit_level;
  if (out.use_mg)
    for (int l = first_explicit; l < out.nlev; ++l) {
      LevelDesc& D = out.lv[l];
      HostLevel& Hl = H[l];
      D.is_explicit = true;
      D.cell_row.assign(D.nb * 64, -1);
      int nper = l >= 2 ? (1 << (l - 2)) : 0;
      for (int64_t g = 0; g < D.nb; ++g) {
        int bx = D.base[3 * g] >> 2, by = D.base[3 * g + 1] >> 2, bz = D.base[3 * g + 2] >> 2;
        int32_t hb = Hl.find(bx, by, bz);
        for (int q = 0; q < 64; ++q) {
          if (!(Hl.cflag[(size_t)hb * 64 + q] & F_ACTIVE_CELL)) continue;
          D.cell_row[g * 64 + q] = (int32_t)D.ncell++;
          D.cell_gpu_index.push_back((int32_t)(g * 64 + q));
          if (l >= 2) {
            int C[3] = {D.base[3 * g] + (q >> 4), D.base[3 * g + 1] + ((q >> 2) & 3), D.base[3 * g + 2] + (q & 3)};
            int fb[3];
            for (int d = 0; d < 3; ++d) fb[d] = ((1 << l) * C[d] - ((1 << l) - 1) * kB[d]) >> 2;
            for (int i = 0; i < nper; ++i)
              for (int j = 0; j < nper; ++j)
                for (int k = 0; k < nper; ++k) {
                  D.cell_fine_blocks.push_back(gmap[0].find(fb[0] + i, fb[1] + j, fb[2] + k));
                }
          }
        }
      }
    }
  trace("explicit-level cells");
  // ---- level-0 cell data and input mapping
  {
    LevelDesc& D = out.lv[0];
    out.nin_blocks = p.nblocks;
    out.in_node_slot.resize(p.nblocks * 64);
    parallel_ranges(p.nblocks, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b) {
        const int32_t* bb = p.block_base + 3 * b;
        if (((bb[0] | bb[1] | bb[2]) & 3) == 0) {  // 4-aligned: one level-0 block
          const int32_t gb = gmap[0].find(bb[0] >> 2, bb[1] >> 2, bb[2] >> 2);
          for (int l = 0; l < 64; ++l) out.in_node_slot[b * 64 + l] = gb < 0 ? -1 : (int32_t)(gb * 64 + l);
          continue;
        }
        for (int l = 0; l < 64; ++l) {
          int x = bb[0] + (l >> 4), y = bb[1] + ((l >> 2) & 3), z = bb[2] + (l & 3);
          const int32_t gb = gmap[0].find(x >> 2, y >> 2, z >> 2);
          out.in_node_slot[b * 64 + l] = gb < 0 ? -1 : (int32_t)(gb * 64 + loc(x, y, z));
        }
      }
    });
    // level-0 moduli and densities through the input map: a slot holds the
    // input cell at its min corner (every ACTIVE cell is in a GPU block);
    // slots without an input cell stay 0.  Input nodes outside the GPU set
    // are collected per range for u_D below.
    out.E.resize(D.nb * 64);
    out.rho.resize(D.nb * 64);
    parallel_ranges(D.nb, [&](int64_t b0, int64_t b1) {
      std::fill(out.E.begin() + b0 * 64, out.E.begin() + b1 * 64, 0.0);
      std::fill(out.rho.begin() + b0 * 64, out.rho.begin() + b1 * 64, 0.0);
    });
    const int64_t nparts = std::min<int64_t>(64, std::max<int64_t>(1, p.nblocks / 1024));
    std::vector<int64_t> ncell(nparts, 0);
    std::vector<std::vector<int64_t>> off_nodes(nparts);
    parallel_ranges(nparts, [&](int64_t r0, int64_t r1) {
      for (int64_t r = r0; r < r1; ++r) {
        const int64_t b1 = std::min(p.nblocks, (r + 1) * ((p.nblocks + nparts - 1) / nparts));
        for (int64_t b = r * ((p.nblocks + nparts - 1) / nparts); b < b1; ++b)
          for (int l = 0; l < 64; ++l) {
            const int64_t i = b * 64 + l;
            const int32_t sl = out.in_node_slot[i];
            if (sl < 0) {
              off_nodes[r].push_back(i);
              continue;
            }
            const double d = p.density[i];
            if (d > 0 && Ein[i] * p.mu != 0.0) {
              out.E[sl] = Ein[i];
              ncell[r]++;
            }
            out.rho[sl] = d;
          }
      }
    }, 1);
    for (int64_t c : ncell) out.n_active_cells += c;
    // u_D (SPGrid_CPU_Solver.cpp:241-246,281-285): written at every listed node
    std::unordered_map<uint64_t, std::vector<int64_t>> in_nodes_off;  // for nodes outside the GPU set
    for (const auto& part : off_nodes)
      for (int64_t i : part) {
        const int64_t b = i / 64;
        const int l = (int)(i % 64);
        const int32_t* bb = p.block_base + 3 * b;
        in_nodes_off[bkey(bb[0] + (l >> 4), bb[1] + ((l >> 2) & 3), bb[2] + (l & 3))].push_back(i);
      }
    // one entry per (slot, axis): a repeated one keeps its first place and
    // the last value, as writing the list in order would leave it
    std::unordered_map<int32_t, size_t> ud_at;
    for (int64_t i = 0; i < p.ndir; ++i) {
      int x = p.dir_coord[3 * i], y = p.dir_coord[3 * i + 1], z = p.dir_coord[3 * i + 2];
      const int32_t gb = gmap[0].find(x >> 2, y >> 2, z >> 2);
      if (gb >= 0) {
        const int32_t e = (int32_t)((gb * 64 + loc(x, y, z)) * 3 + p.dir_axis[i]);
        auto [pos, fresh] = ud_at.emplace(e, out.ud_slot.size());
        if (!fresh) {
          out.ud_value[pos->second] = p.dir_value[i];
          continue;
        }
        out.ud_slot.push_back(e);
        out.ud_value.push_back(p.dir_value[i]);
      } else {
        auto jt = in_nodes_off.find(bkey(x, y, z));
        if (jt != in_nodes_off.end())
          for (int64_t n : jt->second) {
            out.ud_extra.push_back(n * 3 + p.dir_axis[i]);
            out.ud_extra_value.push_back(p.dir_value[i]);
          }
      }
    }
  }
  trace("level-0 data, input map");
  // the band distances serve only the overlapped first sweep (SPG_OVERLAP)
  if (const char* e = std::getenv("SPG_OVERLAP"); e && std::atoi(e) != 0) band_distances(out);
  trace("band distances");
  return out;
}

}  // namespace spg
