/*
 * spgrid_topo_opt_amd.h -- C ABI of the MI355X narrow-band SPGrid elasticity
 * solver (libspgrid_topo_opt_amd.so).
 *
 * Two layers:
 *  1. The drop-in entry point `fem_solve(void*, void*)`, byte-compatible with
 *     the reference's serialized FEMInputs/FEMOutputs protocol
 *     (topo_opt/fem_interface.h:11 declaration, :252-276 input stream,
 *     :373-381 output stream, :414-443 ownership: *output receives a buffer
 *     allocated with new char[] that the caller releases with delete[]).
 *  2. A thin pointer API (`spg_*`) used by the Python driver (torch.autograd)
 *     and by the parity tests.  Plain pointers and sizes only.  Host-side
 *     arrays are marked `host`; arrays marked `device` are HIP device pointers
 *     on the solver's device.
 *
 * Every entry point returns 0 on success and a negative SPG_E_* code on
 * failure; no C++ exception and no abort crosses this boundary
 * (the reference can throw or abort: solver/CG/CONJUGATE_GRADIENT.cpp:49-52,
 * solver/CG/CG_SYSTEM.cpp:56-68).  `spg_last_error()` returns the message of
 * the last failure on the calling thread.
 */
#ifndef SPGRID_TOPO_OPT_AMD_H
#define SPGRID_TOPO_OPT_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define SPG_OK 0
#define SPG_E_INVALID (-1)      /* bad argument / unsupported configuration     */
#define SPG_E_VALIDATION (-2)   /* block activation or padding check failed     */
#define SPG_E_SINGULAR (-3)     /* CG detected s.q <= 0, or bottom not SPD      */
#define SPG_E_NOT_CONVERGED (-4)/* residual target not reached                  */
#define SPG_E_HIP (-5)          /* HIP runtime / device failure                  */
#define SPG_E_NAN (-6)          /* NaN in inputs or iterates                     */
#define SPG_E_VERSION (-7)      /* FEMInputs version mismatch                    */

/* ---- layer 1: the reference's C entry point ----------------------------
 * Replaces `extern "C" void fem_solve(void *input, void *output)`
 * (topo_opt/fem_interface.h:11; defined in solver/SPGrid_CPU_Solver.cpp:26).
 * `input` points at a serialized FEMInputs stream; `output` is really a
 * `void**` that receives a new[]-allocated serialized FEMOutputs.  Failures
 * are reported in FEMOutputs.success / error_code / info.  A solve that ran
 * to its end reports success = 1 even when the tolerance was not reached
 * (error_code SPG_E_NOT_CONVERGED), as the reference does
 * (SPGrid_CPU_Solver.cpp:368).                                               */
void fem_solve(void *input, void *output);

/* Fake solver used by the reference's interface test (u_p = rho + p):
 * replaces fem_solve_test (topo_opt/fem_interface.h:12,
 * topo_opt/interface_test.cpp:7-34).                                         */
void fem_solve_test(void *input, void *output);

/* ---- layer 2: pointer API --------------------------------------------- */
typedef struct spg_solver spg_solver;

/* Solver configuration: the FEMInputs fields and extras the solve honours. */
typedef struct {
  int32_t res[3];          /* FEMInputs::resolution                         */
  double dx;               /* FEMInputs::dx                                 */
  double mu, lambda;       /* FEMInputs::global_mu / global_lambda          */
  double penalty;          /* extra_doubles["penalty"] (SIMP p)             */
  double min_stiffness;    /* extra_doubles["minimum_stiffness"] (Emin)     */
  double min_density;      /* extra_doubles["minimum_density"]              */
  int32_t mg_levels;       /* FEMInputs::mg_level (0 = plain CG, 2 invalid) */
  int32_t smooth_iters;    /* FEMInputs::pre_and_post_smoothing_iter         */
  int32_t boundary_iters;  /* extra_ints["boundary_smoothing"]              */
  int32_t device;          /* HIP device ordinal                            */
  int32_t bottom_dense_max; /* largest bottom level (unknowns) solved with
                              the dense FP64 inverse; larger bottoms use the
                              nested-dissection Cholesky.  0: 16384, the
                              choice for many solves per factorization;
                              fem_solve, which factors once per call, uses
                              4096.  At most 60000.                         */
} spg_config;

/* Solve parameters (krylov + defect-correction knobs of FEMInputs). */
typedef struct {
  double tolerance;        /* target ||P(f-Ku)||_inf <= tolerance*||r0||_inf */
  double inner_tolerance;  /* per-pass inner CG reduction; <=0: each pass aims
                              at the target (an FP32 pass at most 1e-6 of its
                              starting residual)                             */
  int32_t max_iterations;  /* PCG iterations per pass                       */
  int32_t max_passes;      /* FP64 defect-correction passes (>=1)           */
  int32_t precision;       /* 0: FP32 inner / FP64 outer (mixed), 1: FP64   */
  int32_t print_residuals; /* krylov.print_residuals: print ||r||_inf of every
                              PCG iteration, one per line (CG/
                              CONJUGATE_GRADIENT.cpp:34-35); 1 stdout, as the
                              reference, 2 stderr, 0 off                     */
  int32_t warm_start;      /* the reference's warm start
                              (solver/SPGrid_CPU_Solver.cpp:241-285):
                              0 solve from u_D; 1 start from the solver's
                              last solution (or the guess given with
                              spg_set_initial_guess_device) if its residual
                              is below reference_linf, else from u_D;
                              2 always start from it (forced_reuse)        */
  int32_t reserved0;       /* zero                                          */
  double reference_linf;   /* warm starts: the stored zero-guess residual
                              of the first solve (state.initial_residual_linf,
                              SPGrid_CPU_Solver.cpp:318-326); the target is
                              tolerance * reference_linf.  <= 0: the
                              zero-guess residual of this problem           */
} spg_solve_params;

typedef struct {
  int32_t iterations, passes, converged, levels;
  double r0_linf, r_linf;  /* FP64 residual (L_inf) before / after          */
  double r_l2;             /* FP64 residual (L2) after                      */
  double setup_seconds, solve_seconds;
  int64_t active_nodes, active_cells, gpu_blocks;
  int64_t bottom_dofs;
  double hbm_bytes;        /* device memory held by the solver              */
  double r_start_linf;     /* FP64 residual (L_inf) of the starting guess   */
  int32_t warm_reused;     /* 1: the solve started from the previous u      */
  int32_t bottom_kind;     /* 0: dense FP64 inverse of the bottom matrix;
                              1: nested-dissection sparse FP64 Cholesky      */
  double bottom_bytes;     /* device bytes of the bottom solve's factors     */
} spg_info;

const char *spg_last_error(void);
const char *spg_version(void);
int spg_device_count(int32_t *count);

/* Build flags, multigrid hierarchy, Galerkin stencils (GPU FP64) and the
 * bottom FP64 Cholesky for a problem given as 4x4x4 density blocks
 * (FEMInputs::density, fem_interface.h:109-160; index i*16+j*4+k) and
 * Dirichlet nodes (FEMInputs::dirichlet_nodes).  All arrays host.           */
int spg_create(const spg_config *cfg, int64_t nblocks, const int32_t *block_base /*host [nblocks*3]*/,
               const double *density /*host [nblocks*64]*/, int64_t ndir, const int32_t *dir_coord /*host*/,
               const int32_t *dir_axis /*host*/, const double *dir_value /*host*/, spg_solver **out);
void spg_destroy(spg_solver *s);

/* Replace the densities (same block list) and rebuild the coarse operators;
 * flags and topology are kept (only rho changes between optimizer steps).   */
int spg_update_density(spg_solver *s, const double *density /*host [nblocks*64]*/);
/* The same update from a device array (input block order, [nblocks*64]):
 * SIMP map and checks run on the GPU, so nothing crosses the PCIe bus
 * (fem_interface.h:336-353, SPGrid_CPU_Solver.cpp:112-130).  The array must
 * be complete when called (the call runs on the solver's own stream); on a
 * NaN or a change of the active-cell set nothing is modified.               */
int spg_update_density_device(spg_solver *s, const double *density_dev);

/* Nodal forces (FEMInputs::forces, fem_interface.h:49-57).  Assignment
 * semantics: a later force on the same node replaces an earlier one.        */
int spg_set_forces(spg_solver *s, int64_t nforce, const int32_t *coord /*host*/, const double *value /*host*/);

/* Solve K(rho) u = f.  u_out: displacement at the input block nodes in the
 * reference's output order (nblocks*64*4: u_x,u_y,u_z,|r|_max), host.      */
int spg_solve(spg_solver *s, const spg_solve_params *prm, double *u_out4 /*host, may be NULL*/, spg_info *info);

/* Device-resident views of the last solution (FP64, [nblocks*64*3] in input
 * block order): copies into a device buffer owned by the caller.            */
int spg_get_displacement_device(spg_solver *s, double *u_dev /*device [nblocks*64*3]*/);

/* Starting guess for the next warm-started solve (spg_solve_params.
 * warm_start != 0), in input block order (device [nblocks*64*3]): the
 * previous displacement carried over to a rebuilt solver, as the
 * reference's kept SPGrid u channel is (SPGrid_CPU_Solver.cpp:58-75,
 * 241-285).  Dirichlet entries are reset to u_D by the solve.               */
int spg_set_initial_guess_device(spg_solver *s, const double *u_dev);

/* Adjoint solve K lambda = g for a right-hand side given per input block
 * node (device [nblocks*64*3]); result in the same layout.                  */
int spg_solve_adjoint_device(spg_solver *s, const spg_solve_params *prm, const double *g_dev, double *lam_dev,
                             spg_info *info);

/* Compliance sensitivity (objective/minimal_compliance.cpp:74-146):
 *   dc_e = -p rho_e^(p-1) (1-Emin) u_e^T K0 u_e ,  K0 = mu K_mu + la K_la
 * for every input cell (device [nblocks*64], 0 for rho<=0), and the objective
 * sum_e u_e^T K0 u_e (Emin + rho_e^p (1-Emin)) in *objective.  u_dev is a
 * displacement in input-block order (device [nblocks*64*3]); NULL uses the
 * last solution.  If lam_dev != NULL computes -p rho^(p-1)(1-Emin)
 * lam_e^T K0 u_e instead (adjoint form; no objective).                      */
int spg_sensitivity_device(spg_solver *s, const double *u_dev /*device or NULL*/,
                           const double *lam_dev /*device or NULL*/, double *dc_dev /*device*/,
                           double *objective /*host, may be NULL*/);

/* ---- kernel-level entry points (parity tests, benchmarks) --------------
 * Vectors are exchanged in "level node order": the solver's GPU block order,
 * 64 nodes per block (local i*16+j*4+k), components xyz interleaved, i.e.
 * value[(blk*64+l)*3+c].  spg_level_nodes returns the coordinates.          */
/* Sizes, level count, device bytes and setup time of a built solver (the
 * solve-independent fields of spg_info; the others are zero).  No
 * counterpart in the reference, whose solver reports these through logs.   */
int spg_get_info(spg_solver *s, spg_info *info);
int spg_num_levels(spg_solver *s, int32_t *levels);
int spg_level_size(spg_solver *s, int32_t level, int64_t *nblocks);
int spg_level_nodes(spg_solver *s, int32_t level, int32_t *xyz /*host [nb*64*3]*/, uint8_t *flags /*host [nb*64]*/);
/* op codes for spg_kernel */
#define SPG_OP_APPLY 0        /* out = K in (level 0, no projection)          */
#define SPG_OP_RESIDUAL 1     /* out = P(b - K in) (level 0)                  */
#define SPG_OP_GS_SWEEP 2     /* in (u) smoothed in place, b given; out = u    */
#define SPG_OP_GS_BOUNDARY 3  /* boundary-only sweep                          */
#define SPG_OP_RESTRICT 4     /* out(level) = R in(level-1)                   */
#define SPG_OP_PROLONGATE 5   /* out(level) = b(level) + P in(level+1)         */
#define SPG_OP_STENCIL_RES 6  /* out = P(b - A in) on an explicit level        */
#define SPG_OP_VCYCLE 7       /* out = V-cycle(b) (level 0)                    */
#define SPG_OP_BOTTOM 8       /* out = A^-1 b on the bottom level              */
#define SPG_OP_OPERATOR 9     /* out = P K in (level 0): the Krylov operator   */
/* Run one operator on host vectors (copied to/from the device);
 * precision 0 = FP32 kernels, 1 = FP64 kernels; bit 0 of `forward` selects
 * the GS colour order, bits 8+ the number of boundary sweeps (0 = one).    */
int spg_kernel(spg_solver *s, int32_t op, int32_t level, int32_t precision, int32_t forward, const double *in,
               const double *b, double *out);
/* FP64 Galerkin stencils of an explicit level: [nb*64*243] host, row-major
 * [node][i][j][k][v][w] (Elasticity_Matrix_Data::data, MG/Elasticity_Matrix.h:55-60). */
int spg_level_stencils(spg_solver *s, int32_t level, double *out);
/* Time `reps` launches of the level-0 stiffness apply on device-resident
 * data; returns the average kernel time in seconds (HIP events).            */
int spg_bench_apply(spg_solver *s, int32_t precision, int32_t reps, double *seconds);
/* Time `reps` runs of one operator (SPG_OP_*) of `level` on device-resident
 * vectors (HIP events on the solver's stream); returns seconds per run.
 * SPG_OP_GS_SWEEP times one full sweep (8 colour steps), SPG_OP_GS_BOUNDARY
 * one boundary phase of `arg` sweeps, SPG_OP_VCYCLE one V-cycle (level 0). */
int spg_bench_kernel(spg_solver *s, int32_t op, int32_t level, int32_t precision, int32_t arg, int32_t reps,
                     double *seconds);

/* ---- optimizer-side kernels on the same 4x4x4 cell blocks (no solver
 * handle; all arrays device; cell index block*64 + i*16 + j*4 + k; nbr is
 * [nblocks*27], spoke s = (dx+1)*9 + (dy+1)*3 + (dz+1), -1 where absent). */

/* Sensitivity filter smooth_dc (topo_opt/spgrid_topo_opt.cpp:527-574):
 *   out(x) = sum_o w(o) dc(x+o) rho(x+o) [x+o inside]
 *            / (sum_o w(o) [x+o inside] * (rho(x) + 1e-3)),
 *   w(o) = max(0, r - |o|), |o_i| <= floor(r) <= 3.  Every cell takes the
 *   formula, inside the container or not, as smooth_dc writes every cell of
 *   the fat blocks; a cell with no inside cell in its window gets 0 (the
 *   reference computes 0/0 there).                                           */
int spg_filter_sensitivity(int64_t nblocks, const int32_t *nbr, const uint8_t *inside, const double *dc,
                           const double *rho, double filter_size, double *out);

/* One trial of the optimality-criterion bisection
 * (topo_opt/optimizer/optimality_criterion.cpp:76-118) at multiplier mid.
 * mask bit 0: the cell is inside the container and is updated; bit 1: the
 * cell has a fixed density (flags.get_fixed_density()).  Updated cells take
 *   new = clamp(clamp(rho sqrt(max(0,-dc-dc_shift)/mid), rho-step, rho+step),
 *               min_density, 1)
 * or fixed_density for fixed cells; other cells keep rho.  dc_shift is
 * min(0, min -dc over the container) under smart_oc, else 0.  *volume = the
 * sum of new over updated cells (fixed ones left out when exclude_fixed),
 * a deterministic FP64 sum; with out != NULL the new densities are written
 * there and *change = max |new - rho| over updated cells.                   */
int spg_oc_trial(int64_t n, const uint8_t *mask, const double *rho, const double *dc, double mid, double step,
                 double min_density, double fixed_density, int32_t exclude_fixed, double dc_shift, double *out,
                 double *volume, double *change);

/* spg_oc_trial without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device or pinned memory (code under
 * HIP graph capture, tight optimizer loops, harnesses that forbid hipMalloc).
 * Same kernels and same result as spg_oc_trial, but
 *   - the per-workgroup partials live in a fixed device array of
 *     SPG_OC_NOALLOC_PARTS (volume, change) pairs (16 KiB), one copy per GPU,
 *     held until the process ends; the grid is capped at that many
 *     workgroups, which cover any n >= 0 by a grid-stride loop;
 *   - result (device, double[2]) receives {volume, change};
 *   - the kernels are enqueued on `stream` (a hipStream_t; NULL = the null
 *     stream) and the call returns without synchronizing; launch errors are
 *     reported, kernel faults surface at the caller's next synchronization.
 * Calls that overlap on one GPU (other streams, other host threads) share
 * that GPU's array and race: serialize them.                                 */
#define SPG_OC_NOALLOC_PARTS 1024
int spg_oc_trial_noalloc(int64_t n, const uint8_t *mask, const double *rho, const double *dc, double mid,
                         double step, double min_density, double fixed_density, int32_t exclude_fixed,
                         double dc_shift, double *out, double *result /*device [2]*/, void *stream);

#ifdef __cplusplus
}
#endif
#endif
