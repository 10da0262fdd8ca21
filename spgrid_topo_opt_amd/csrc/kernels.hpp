// kernels.hpp -- launch wrappers of the gfx950 kernels (kernels.hip).
//
// Level data layout (device): blocks of 4x4x4 nodes, local index
// l = i*16 + j*4 + k; a vector is stored per block and component:
// v[(blk*3 + c)*64 + l].  Explicit-level stencils hold 243 coefficients
// per node, e = ((s*3 + v)*3 + w), s = (di+1)*9 + (dj+1)*3 + (dk+1), grouped
// by (block, colour) as st_index() describes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace spg {

// constants of the strain-form cell operator per unit modulus: 9mu, 4.5la,
// 4.5mu, 1.5mu, 1.5la and 2mu+0.5la, all / 72 (k_apply_cell)
struct CellConsts {
  double d = 0, t = 0, s = 0, w = 0, q = 0, c3 = 0;
};
CellConsts cell_consts(double mu, double lambda);

struct LevelDev {
  int nb = 0;
  CellConsts cell;                 // level 0: the material's cell-operator constants
  const int* nbr = nullptr;        // [nb*27]
  const uint8_t* flags = nullptr;  // [nb*64]
  const int* rsrc = nullptr;       // [nb*27] restriction sources (level >= 1)
  const int* psrc = nullptr;       // [nb*8]  prolongation sources (level < L-1)
  // [nb*27] the 27 coarse nodes a fine block interpolates from, as
  // (coarse block << 6) | local index, -1 where absent (launch_prolong_table)
  const int* psrc27 = nullptr;
  const int* base = nullptr;       // [nb*3]  block origin (node coords)
  const int* bnd = nullptr;        // boundary nodes (blk*64+l), concatenated by colour
  int bnd_off[9] = {};             // colour c owns bnd[bnd_off[c] .. bnd_off[c+1])
  int bnd_max = 0;                 // longest colour list
  // boundary system (LDS-resident boundary phase): neighbour tables [n*32]
  // and coefficient rows [n*288] per precision, plus the b~ scratch [n*4]
  const int* bci = nullptr;
  const int* bslot = nullptr;
  const float* bA32 = nullptr;
  const double* bA64 = nullptr;
  float* bt32 = nullptr;
  double* bt64 = nullptr;
  // level-0 matrix-free boundary system: cell-node indices [8*n*8] (int16),
  // cell moduli [8*n] and inverse self blocks [n*8] per precision
  const int16_t* b0node = nullptr;
  const float* b0E32 = nullptr;
  const double* b0E64 = nullptr;
  const float* b0D32 = nullptr;
  const double* b0D64 = nullptr;
  // cooperative level-0 phase: boundary iterate [3*(n+1)] (8-byte entries,
  // either precision) and {barrier counter, error word}
  void* b0ub = nullptr;
  unsigned* coop = nullptr;
  int coop_off = 0;  // 1 after a barrier time-out: the list launches take over
  // explicit levels: dense maps [G H] of one boundary phase (forward,
  // backward colour order) for `bmap_sweeps` sweeps, and their vector
  // scratch ([6n] input then [3n] output)
  const float* bmap32[2] = {nullptr, nullptr};
  const double* bmap64[2] = {nullptr, nullptr};
  float* bmx32 = nullptr;
  double* bmx64 = nullptr;
  int bmap_sweeps = 0;
  // level 0: slots of the 27 neighbours of every boundary node ([n*27], -1 absent)
  const int* bnslot = nullptr;
  // level 0: the flags in colour-major order ([8][nb][8]: colour, block, k-th
  // node of the colour in the block), read by the colour sweep
  const uint8_t* flags_cm = nullptr;
};

// Colour of the i-th step of a Gauss-Seidel sweep: level 0 runs the colours
// 0..7 forward / 7..0 backward (SPGrid_Gauss_Seidel.h:765-875); the
// explicit levels use the reference's colour order
// (Multigrid_Hierarchy.h:577-617).
__host__ __device__ inline int sweep_colour(bool level0, bool forward, int i) {
  const unsigned ord = level0 ? (forward ? 0x76543210u : 0x01234567u) : (forward ? 0x02376451u : 0x15467320u);
  return (int)((ord >> (4 * i)) & 7u);
}

// Explicit-level stencil layout.  The 1944 coefficients of the 8 nodes of one
// colour in one block are contiguous; node k (0..7 inside the colour) and
// spoke s = g + 8t go to lane k*8+g of row t*9+m (t < 3) or, for the spokes
// 24..26, to the tail 1728 + m*24 + k*3 + g (m = v*3+w).
__host__ __device__ inline int64_t st_index(int64_t blk, int l, int e) {
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const int colour = ((li & 1) << 2) | ((lj & 1) << 1) | (lk & 1);
  const int k = ((li >> 1) << 2) | ((lj >> 1) << 1) | (lk >> 1);
  const int s = e / 9, m = e % 9, g = s & 7, t = s >> 3;
  const int64_t base = (blk * 8 + colour) * 1944;
  return t < 3 ? base + (t * 9 + m) * 64 + k * 8 + g : base + 1728 + m * 24 + k * 3 + g;
}

// device scalars of one PCG iteration
struct CgScalars {
  double rho, rho_old, sq, alpha, beta, nrm, nrm2;
  int singular;
  int pad;
};

// The element matrix K0 = mu K_mu + lambda K_la and the halo tables, in the
// device's constant memory (one set per device: Solver::activate uploads
// its material's before its kernels run).  Synchronous.
hipError_t upload_k0(const double* K0);

// level-0 operator: out = K in (mode 0), out = P(b - K in) (mode 1), out = P K in (mode 2)
template <class T>
void launch_apply(hipStream_t st, const LevelDev& L, const T* E, T dx, const T* in, const T* b, T* out, int mode);
// out = P K in (mode 2) and *dot = in . out (deterministic); `bpart` holds
// apply_dot_partials(nb) doubles, `part` reduce_partials(n) doubles
int64_t apply_dot_partials(int64_t nb);
template <class T>
void launch_apply_dot(hipStream_t st, const LevelDev& L, const T* E, T dx, const T* in, T* out, double* bpart,
                      double* part, double* dot, CgScalars* sc = nullptr, int stage = -1);
// level-0 GS: one colour over all blocks; all boundary sweeps of one phase.
// With `bcm` (b in colour-major order, launch_colour_major) and the level's
// flags_cm, the colour kernel reads b and the flags of its colour only.
template <class T>
void launch_gs0(hipStream_t st, const LevelDev& L, const T* E, T dx, T* u, const T* b, int colour,
                const uint8_t* dist = nullptr, int dmin = 0, const T* bcm = nullptr);
// v in colour-major order: vcm[((c * nb + blk) * 3 + comp) * 8 + k] = v at the
// k-th node of colour c in block blk (level-0 vectors)
template <class T>
void launch_colour_major(hipStream_t st, const LevelDev& L, const T* v, T* vcm);
// fills the level's psrc27 table (out, [nb*27]) from its psrc and base
void launch_prolong_table(hipStream_t st, const LevelDev& F, int* out);
void launch_colour_major_flags(hipStream_t st, int nb, const uint8_t* flags, uint8_t* fcm);
// one level-0 colour step over a list of slots (27 neighbour slots each in nslot)
template <class T>
void launch_gs0_list(hipStream_t st, const LevelDev& L, const int* list, const int* nslot, int n, const T* E, T dx,
                     T* u, const T* b);
// neighbour slots (27 per node) of a list of level slots
void launch_list_nslot(hipStream_t st, const LevelDev& L, int n, const int* list, int* out);
template <class T>
void launch_gs0_boundary(hipStream_t st, const LevelDev& L, const T* E, T dx, T* u, const T* b, int sweeps,
                         bool forward);
// explicit levels
template <class T>
void launch_st_residual(hipStream_t st, const LevelDev& L, const T* ST, const T* u, const T* b, T* r);
template <class T>
void launch_st_gs(hipStream_t st, const LevelDev& L, const T* ST, T* u, const T* b, int colour);
template <class T>
void launch_st_boundary(hipStream_t st, const LevelDev& L, const T* ST, T* u, const T* b, int sweeps, bool forward);
// transfers
// coarse = R fine; `coarse_zero` (optional) is cleared in the same launch
template <class T>
void launch_restrict(hipStream_t st, const LevelDev& C, const T* fine, T* coarse, T* coarse_zero = nullptr);
template <class T>
void launch_prolongate(hipStream_t st, const LevelDev& F, const T* coarse, T* fine);
// vectors
template <class T>
void launch_zero(hipStream_t st, T* v, int64_t n);
template <class T>
void launch_copy(hipStream_t st, const T* a, T* b, int64_t n);
void launch_cast_d2f(hipStream_t st, const double* a, float* b, int64_t n);
void launch_cast_f2d(hipStream_t st, const float* a, double* b, int64_t n);
template <class T>
void launch_axpy_d(hipStream_t st, double* u, const T* d, int64_t n);  // u += d (FP64 accumulate)
// reductions (two-stage, deterministic); partial buffer needs >= reduce_partials(n) doubles
int64_t reduce_partials(int64_t n);
template <class T>
void launch_dot(hipStream_t st, const T* a, const T* b, int64_t n, double* partial, double* out);
template <class T>
void launch_linf(hipStream_t st, const T* a, int64_t n, double* partial, double* out);
void launch_sumsq(hipStream_t st, const double* a, int64_t n, double* partial, double* out);
// PCG steps (scalars on device)
template <class T>
void launch_cg_direction(hipStream_t st, T* s, const T* z, const CgScalars* sc, int64_t n);  // s = beta s + z
template <class T>
void launch_cg_update(hipStream_t st, T* x, T* r, const T* s, const T* q, CgScalars* sc, int64_t n, int stage = -1,
                      T* zclear = nullptr);
void launch_cg_scalars(hipStream_t st, CgScalars* sc, int stage);  // 0: after rho, 1: after sq, 2: end
// (the reductions above take `sc, stage` to run a stage in their last kernel)
// projections
template <class T>
void launch_project(hipStream_t st, const LevelDev& L, T* v, bool nullspace);
// v = P P_null v, then *out = v . w (deterministic two-stage sum)
template <class T>
void launch_project_dot(hipStream_t st, const LevelDev& L, T* v, const T* w, double* partial, double* out,
                        CgScalars* sc = nullptr, int stage = -1);
// Galerkin assembly (FP64)
void launch_galerkin_gather(hipStream_t st, int64_t ncell, int depth, const int* cell_fine_blocks,
                            const int* cell_gpu_index, const double* E0, double* G);
void launch_galerkin_assemble(hipStream_t st, const LevelDev& L, const int* cell_row, const double* Cm, double dx,
                              double* ST);
void launch_stencil_to_float(hipStream_t st, const double* a, float* b, int64_t n);
// coefficient rows of the boundary nodes (FP64): level 0 from E and K0 (times
// dx), explicit levels from the stencils
void launch_bnd_coef0(hipStream_t st, const LevelDev& L, const double* E, double dx, double* A);
// level-0 matrix-free boundary system (FP64): cell moduli and inverse self blocks
void launch_bnd0_setup(hipStream_t st, const LevelDev& L, const int* cslot, const double* E, double dx, double* bE,
                       double* bD);
void launch_bnd_coef_st(hipStream_t st, const LevelDev& L, const double* ST, double* A);
// row stride of the dense boundary map of an n-node boundary set: 6n padded to 64
__host__ __device__ inline int64_t bmap_stride(int n) { return ((int64_t)6 * n + 63) / 64 * 64; }
// dense FP64 map [G H] (3n x 6n, row-major, stride bmap_stride(n)) of a boundary
// phase on an explicit level; the padding columns must hold zeros
void launch_bmap_build(hipStream_t st, const LevelDev& L, int sweeps, bool forward, double* M);
// bottom
void launch_dense_assemble(hipStream_t st, const LevelDev& L, const double* ST, const int* var_of, int64_t n,
                           double* A);
void launch_symmetrize_lower(hipStream_t st, double* A, int64_t n);
void launch_set_diagonal(hipStream_t st, double* A, int64_t n, double v);  // A[i + i n] = v
void launch_bnd_nslot(hipStream_t st, const LevelDev& L, int* out);  // fills LevelDev::bnslot
// density update on the device: err |= 1 NaN, 2 deactivation, 4 activation
void launch_density_check(hipStream_t st, int64_t n, const int* slot, const double* dens, const double* E64,
                          int* err);
void launch_density_update(hipStream_t st, int64_t n, const int* slot, const double* dens, double penalty,
                           double emin, double rmin, double* E64, float* E32, double* rho);
template <class T>
void launch_bottom_gather(hipStream_t st, const LevelDev& L, const int* var_of, const T* b, T* x);
template <class T>
void launch_bottom_scatter(hipStream_t st, const LevelDev& L, const int* var_of, const T* y, const T* b, T* u);
// symmetric matrix in packed lower-triangle 64x64 tiles: pack from a dense
// column-major FP64 matrix, and y = A x (x, y padded to nT*64; part holds
// symv_splits(nT) * nT * 64 partial sums)
template <class T>
void launch_pack_sym_tiles(hipStream_t st, const double* A, int64_t nv, int nT, T* tiles);
int symv_splits(int nT);
template <class T>
void launch_symv(hipStream_t st, int nT, const T* tiles, const T* x, T* part, T* y);
// sensitivity: per level-0 slot (cell): dc = -p rho^(p-1)(1-Emin) a_e^T K0 u_e ; obj partials
void launch_sensitivity(hipStream_t st, const LevelDev& L, const double* rho, const double* u, const double* a,
                        double penalty, double emin, double* dc, double* objcell);
// gather / scatter between input-block order and level-0 slots
void launch_scatter_input_vec(hipStream_t st, const int* in_slot, int64_t nin, const double* in3, double* v);
void launch_gather_input_vec(hipStream_t st, const int* in_slot, int64_t nin, const double* v, double* out3);
void launch_gather_input_scalar(hipStream_t st, const int* in_slot, int64_t nin, const double* v, double* out);
// per input block (nin = 64 x blocks): `hdr` int32 words of base
// coordinates (0 or 3, from base[3b..3b+2]) and 64 x {u_x, u_y, u_z,
// max_c |r_c|} doubles, as 32-bit words
void launch_pack_out4(hipStream_t st, const int* in_slot, int64_t nin, const double* u, const double* r,
                      const int32_t* base, int hdr, uint32_t* out);
// v[idx[i]] = val[i] (entry = slot * 3 + component) in parallel: the entries must be distinct
void launch_set_entries(hipStream_t st, double* v, const int* idx, const double* val, int64_t n);

}  // namespace spg
