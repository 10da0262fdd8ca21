/*
 * spgrid_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room FP64 CPU restatement of the reference's narrow-band SPGrid
 * linear-elasticity solve (matrix-free Q1 apply, multigrid-preconditioned
 * projected CG, Galerkin coarse levels, exact bottom solve, one or more FP64
 * defect-correction passes) and of the compliance sensitivity.  It exists to
 * check the HIP implementation (tests, smoke check) and to time a CPU baseline
 * (bench.py cpu_baseline leg).  Nothing in the product links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning of each
 * part (which tests check it against what) is listed in tests/test_oracle_*.py.
 *
 * Parity unpinned: the exact rounding trajectory of ORC_MODE_REF (where the
 * reference rounds through float, and in which order its SIMD reductions
 * add).  The reference cannot be built here, so that mode is only checked to
 * converge to the pinned FP64 solution; every other function is pinned
 * (element tables, apply, transfers, Galerkin stencils, Gauss-Seidel, bottom
 * solve, PCG, sensitivity) against closed forms, scipy assemblies, dense
 * textbook algorithms or finite differences.
 */
#ifndef SPGRID_ORACLE_H
#define SPGRID_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node/cell flag bits: solver/Flags.h:10-24 */
#define ORC_ACTIVE_NODE 0x00010u
#define ORC_DIRICHLET_NODE_X 0x00020u
#define ORC_DIRICHLET_NODE_Y 0x00040u
#define ORC_DIRICHLET_NODE_Z 0x00080u
#define ORC_DIRICHLET_NODE 0x000E0u
#define ORC_ACTIVE_CELL 0x00100u
#define ORC_DIRICHLET_CELL_X 0x00200u
#define ORC_DIRICHLET_CELL 0x00E00u
#define ORC_BOUNDARY_NODE 0x01000u
#define ORC_BOUNDARY_NODE_TEMP 0x02000u

/* Rounding modes of the inner solve.
 * ORC_MODE_FP64: every operation in FP64 (the "plain FP64" solve).
 * ORC_MODE_REF : FP64 storage with the reference's float casts
 *                (CG/CG_VECTOR.cpp:25-113, MG/Multigrid_Hierarchy.h:439-451,577-618). */
#define ORC_MODE_FP64 0
#define ORC_MODE_REF 1

typedef struct {
  int32_t res[3];
  double dx;
  double mu0, lambda0;
  double penalty, min_stiffness, min_density;
  int32_t mg_levels;        /* FEMInputs::mg_level (0 = plain CG)           */
  int32_t smooth_iters;     /* pre_and_post_smoothing_iter                  */
  int32_t boundary_iters;   /* extra_ints["boundary_smoothing"]             */
  int64_t nblocks;          /* density blocks, 4x4x4 cells each             */
  const int32_t *block_base;/* [nblocks*3]                                  */
  const double *density;    /* [nblocks*64], index i*16+j*4+k               */
  int64_t ndir;
  const int32_t *dir_coord; /* [ndir*3] */
  const int32_t *dir_axis;  /* [ndir]   */
  const double *dir_value;  /* [ndir]   */
  int64_t nforce;
  const int32_t *force_coord; /* [nforce*3] */
  const double *force_value;  /* [nforce*3] */
} orc_problem;

typedef struct orc_system orc_system;

/* Element tables K_mu, K_la [8][8][3][3] derived from the Q1 weak form
 * (solver/CPU_Kernels/Stiffness_Matrices.h:8-688 hold the same numbers). */
void orc_element_tables(double *kmu /*576*/, double *kla /*576*/);

/* SIMP map: topo_opt/fem_interface.h:343-353 */
double orc_stiffness_mapping(double penalty, double min_stiffness, double min_density, double d);

/* Build flags, hierarchy, Galerkin stencils and the bottom factorization.
 * Returns NULL on failure with a message in err (reference: success=0 /
 * abort paths in SPGrid_CPU_Solver.cpp:131-145, CG_SYSTEM.cpp:56-75). */
orc_system *orc_build(const orc_problem *p, char *err, int errlen);
void orc_free(orc_system *s);

int orc_num_levels(const orc_system *s);
/* allocated nodes of a level (nodes of allocated 4x4x8 pages) */
int64_t orc_num_nodes(const orc_system *s, int level);
void orc_node_xyz(const orc_system *s, int level, int32_t *xyz);
void orc_node_flags(const orc_system *s, int level, uint32_t *flags);
/* level-0 cell moduli at allocated nodes (cell indexed by its min corner) */
void orc_cell_moduli(const orc_system *s, double *mu, double *la);

/* f = K u on level 0 (dx-scaled, all allocated nodes, no projection):
 * SPGrid_Linear_Elasticity.h:96-213 */
void orc_apply(const orc_system *s, const double *u, double *f);
/* One colour-ordered Gauss-Seidel sweep on `level` (matrix-free at 0,
 * stencil-based at explicit levels), optionally boundary-only:
 * SPGrid_Gauss_Seidel.h:877-1183, SPGrid_Stencil_Multiply.h:161-279,
 * Multigrid_Hierarchy.h:504-548 (no float cast). */
void orc_gs_sweep(const orc_system *s, int level, double *u, const double *b, int forward, int boundary);
/* coarse = R fine (+Dirichlet projection): Restriction.h:52-92, Multigrid_Hierarchy.h:560-568 */
void orc_restrict(const orc_system *s, int coarse_level, const double *fine, double *coarse);
/* fine += P coarse (+Dirichlet projection): Prolongation.h:57-99, Multigrid_Hierarchy.h:550-558 */
void orc_prolongate(const orc_system *s, int fine_level, double *fine, const double *coarse);
/* r = b - A u with stored stencils (+projection): Multigrid_Hierarchy.h:407-425 */
void orc_stencil_residual(const orc_system *s, int level, const double *u, const double *b, double *r);
/* 243 stencil coefficients [i][j][k][v][w] of every allocated node (0 where none) */
int orc_get_stencils(const orc_system *s, int level, double *out);
/* One V-cycle z = M^-1 b on level 0: Multigrid_Hierarchy.h:569-619 */
void orc_vcycle(orc_system *s, const double *b, double *z, int mode);

typedef struct {
  double tolerance;       /* target: ||r||_inf <= tolerance * ||r0||_inf (FP64 residual) */
  double inner_tolerance; /* per pass, relative to the pass's starting residual (<=0: =tolerance) */
  int32_t max_iterations; /* PCG iterations per pass */
  int32_t max_passes;     /* defect-correction passes */
  int32_t mode;           /* ORC_MODE_* */
} orc_solve_params;

typedef struct {
  int32_t iterations;     /* total PCG iterations */
  int32_t passes;
  int32_t converged;
  double r0_linf, r_linf; /* FP64 residual before / after */
  double setup_seconds, solve_seconds;
} orc_solve_info;

/* Full solve: SPGrid_CPU_Solver.cpp:241-341 (+ iterated defect correction).
 * u_out: [nnodes0*3] displacement at allocated level-0 nodes (u_D + delta).
 * r_out: [nnodes0*3] final FP64 residual P(f - K u) (may be NULL).
 * Returns 0 on success, negative on error (singular / not converged). */
int orc_solve(orc_system *s, const orc_solve_params *prm, double *u_out, double *r_out,
              orc_solve_info *info, char *err, int errlen);

/* Only the PCG iterations of one pass starting from the current state, for
 * bounded CPU timing: runs exactly `iters` iterations (no convergence stop). */
int orc_time_pcg_iterations(orc_system *s, int iters, int mode, double *seconds);

/* Compliance sensitivity: topo_opt/objective/minimal_compliance.cpp:74-146.
 * u: level-0 displacement at allocated nodes; for every active cell
 * (indexed by allocated node of its min corner) returns
 *   s_e  = u_e^T K0 u_e            (K0 = mu0 K_mu + lambda0 K_la, no dx)
 *   dc_e = -p rho^(p-1) (1-Emin) s_e
 * and the objective sum_e s_e (Emin + rho^p (1-Emin)).  rho is looked up from
 * the density blocks.  Cells without density get 0. */
double orc_sensitivity(const orc_system *s, const double *u, double *dc, double *se);
/* Generalised adjoint form: g_e = -p rho^(p-1)(1-Emin) lam_e^T K0 u_e */
void orc_adjoint_sensitivity(const orc_system *s, const double *u, const double *lam, double *g);

/* Reference output packing (SPGrid_CPU_Solver.cpp:343-364): for each input
 * block node (nblocks*64) returns (u_x,u_y,u_z,|r|_max). */
void orc_pack_displacements(const orc_system *s, const double *u, const double *r, double *out4);

#ifdef __cplusplus
}
#endif
#endif
