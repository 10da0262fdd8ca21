/*
 * spgrid_oracle.c -- TEST INFRASTRUCTURE ONLY (see spgrid_oracle.h).
 *
 * CPU FP64 restatement of the reference's narrow-band SPGrid elasticity solve.
 * Data layout is deliberately simple (a dense index box per level plus compact
 * arrays of allocated nodes) so that the restatement reads like the algorithm;
 * it shares no code with the HIP implementation it checks.
 *
 * Geometry conventions restated from the reference:
 *  - SPGrid pages hold 4x4x8 nodes (x,y,z) for the 128-byte FP64 node struct
 *    (SPGrid_CPU_Solver.cpp:44-49, SPGrid_Mask.h:31-35); "allocated" nodes are
 *    all nodes of pages that were set.
 *  - Coarse index of a fine node/cell: (x + B) >> 1 per axis, B = (4,4,8)
 *    (SPGrid_Mask.h:183-187 DownsampleOffset<block sizes>); the inverse map is
 *    2c - B (UpsampleOffset, SPGrid_Mask.h:204-208).
 *  - Cells are indexed by their minimum-corner node.
 */
#include "spgrid_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static const int PAGE[3] = {4, 4, 8};   /* page extent per axis */
static const int BPAD[3] = {4, 4, 8};   /* B in (x+B)>>1        */

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

/* ------------------------------------------------------------------------- */
/* Element tables from the Q1 weak form on the unit cube.                      */
/* a(u,v) = mu * int(grad u : grad v + grad u : grad v^T) + la * int(div u div v)
 * K[a][b][v][w] = a(N_b e_w, N_a e_v); node id = x*4+y*2+z.
 * Reference tables: solver/CPU_Kernels/Stiffness_Matrices.h:8-347 (K_mu),
 * :348-688 (K_la); derivation sketch Linear_Elasticity_System_Matrix.h:29-123. */
/* ------------------------------------------------------------------------- */
static double i1_pp(int a, int b) { return a == b ? 1.0 / 3.0 : 1.0 / 6.0; } /* int phi_a phi_b   */
static double i1_dd(int a, int b) { return a == b ? 1.0 : -1.0; }             /* int phi'_a phi'_b */
static double i1_dp(int a, int b) { (void)b; return a ? 0.5 : -0.5; }         /* int phi'_a phi_b  */

/* int dN_a/dx_p * dN_b/dx_q over the unit cube */
static double grad_int(int a, int b, int p, int q) {
  int ax[3] = {(a >> 2) & 1, (a >> 1) & 1, a & 1};
  int bx[3] = {(b >> 2) & 1, (b >> 1) & 1, b & 1};
  double r = 1.0;
  for (int d = 0; d < 3; ++d) {
    if (d == p && d == q) r *= i1_dd(ax[d], bx[d]);
    else if (d == p) r *= i1_dp(ax[d], bx[d]);
    else if (d == q) r *= i1_dp(bx[d], ax[d]);
    else r *= i1_pp(ax[d], bx[d]);
  }
  return r;
}

void orc_element_tables(double *kmu, double *kla) {
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w) {
          double lap = 0.0;
          if (v == w)
            for (int d = 0; d < 3; ++d) lap += grad_int(a, b, d, d);
          int id = ((a * 8 + b) * 3 + v) * 3 + w;
          kmu[id] = lap + grad_int(a, b, w, v);
          kla[id] = grad_int(a, b, v, w);
        }
}

double orc_stiffness_mapping(double penalty, double min_stiffness, double min_density, double d) {
  /* topo_opt/fem_interface.h:343-353 */
  if (penalty != 0.0) {
    double x = d > min_density ? d : min_density;
    return (1.0 - min_stiffness) * pow(x, penalty) + min_stiffness;
  }
  return d;
}

/* ------------------------------------------------------------------------- */
/* Level data                                                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
  int lo[3], n[3];        /* node box */
  int64_t nbox;
  int plo[3], pn[3];      /* page box */
  uint8_t *page;          /* page set */
  int32_t *idx;           /* box -> compact allocated node, -1 */
  uint32_t *bflags;       /* box-dense flags (cell + node bits, incl. non-allocated) */
  int64_t nn;
  int32_t *xyz;           /* [nn*3] */
  int32_t *nb;            /* [nn*27] neighbour compact index (-1 absent) */
  uint32_t *flags;        /* [nn] */
  int nb_stencil;         /* has stencils */
  double *st;             /* [nn*243] */
  uint32_t *spoke;        /* [nn] spoke-active bits (27) */
  int64_t ndl[3];
  int32_t *dl[3];         /* Dirichlet lists per axis */
  int64_t nbnd;
  double *u, *b, *r;      /* V-cycle work vectors [nn*3] */
  /* bottom solve (envelope Cholesky) */
  int64_t nvar;
  int32_t *var_of;        /* [nn] variable base index (3 per node) or -1 */
  int64_t *rowstart;      /* [nvar+1] offsets into L */
  int32_t *fc;            /* first column per row */
  double *L;
} lvl_t;

struct orc_system {
  orc_problem p;
  int nlev;
  int explicit_level;
  int use_mg;
  lvl_t lv[16];
  double kmu[576], kla[576];
  double *mu, *la;   /* level-0 cell moduli at allocated nodes */
  double *rho;       /* level-0 input density at allocated nodes (cell) */
  uint8_t *has_rho;
  /* input Dirichlet values (u_D) */
  int64_t nd;
  int32_t *d_node;   /* compact node or -1 */
  int32_t *d_axis;
  double *d_val;
  int32_t *d_xyz;
  /* forces */
  int32_t *f_node;
  /* PCG vectors at level 0 */
  double *x, *rr, *s, *q, *z, *fvec, *ubase;
};

static inline int64_t box_lin(const lvl_t *L, int x, int y, int z) {
  int a = x - L->lo[0], b = y - L->lo[1], c = z - L->lo[2];
  if (a < 0 || b < 0 || c < 0 || a >= L->n[0] || b >= L->n[1] || c >= L->n[2]) return -1;
  return ((int64_t)a * L->n[1] + b) * L->n[2] + c;
}
static inline int32_t node_at(const lvl_t *L, int x, int y, int z) {
  int64_t l = box_lin(L, x, y, z);
  return l < 0 ? -1 : L->idx[l];
}
static inline uint32_t bflag_at(const lvl_t *L, int x, int y, int z) {
  int64_t l = box_lin(L, x, y, z);
  return l < 0 ? 0u : L->bflags[l];
}
static inline int page_set(const lvl_t *L, int x, int y, int z) {
  int px = floordiv(x, PAGE[0]) - L->plo[0], py = floordiv(y, PAGE[1]) - L->plo[1],
      pz = floordiv(z, PAGE[2]) - L->plo[2];
  if (px < 0 || py < 0 || pz < 0 || px >= L->pn[0] || py >= L->pn[1] || pz >= L->pn[2]) return 0;
  return L->page[((int64_t)px * L->pn[1] + py) * L->pn[2] + pz];
}
static inline void mark_page(lvl_t *L, int x, int y, int z) {
  int px = floordiv(x, PAGE[0]) - L->plo[0], py = floordiv(y, PAGE[1]) - L->plo[1],
      pz = floordiv(z, PAGE[2]) - L->plo[2];
  if (px < 0 || py < 0 || pz < 0 || px >= L->pn[0] || py >= L->pn[1] || pz >= L->pn[2]) return;
  L->page[((int64_t)px * L->pn[1] + py) * L->pn[2] + pz] = 1;
}

/* set up a level box covering pages [pmin, pmax] with one margin page each side */
static void level_box(lvl_t *L, const int pmin[3], const int pmax[3]) {
  for (int d = 0; d < 3; ++d) {
    L->plo[d] = pmin[d] - 1;
    L->pn[d] = pmax[d] - pmin[d] + 3;
    L->lo[d] = L->plo[d] * PAGE[d];
    L->n[d] = L->pn[d] * PAGE[d] + 1;
  }
  L->nbox = (int64_t)L->n[0] * L->n[1] * L->n[2];
  L->page = (uint8_t *)calloc((size_t)L->pn[0] * L->pn[1] * L->pn[2], 1);
  L->bflags = (uint32_t *)calloc((size_t)L->nbox, sizeof(uint32_t));
  L->idx = (int32_t *)malloc((size_t)L->nbox * sizeof(int32_t));
}

/* enumerate allocated nodes (lexicographic x,y,z) */
static void level_enumerate(lvl_t *L) {
  int64_t nn = 0;
  for (int a = 0; a < L->n[0]; ++a)
    for (int b = 0; b < L->n[1]; ++b)
      for (int c = 0; c < L->n[2]; ++c) {
        int64_t l = ((int64_t)a * L->n[1] + b) * L->n[2] + c;
        int x = L->lo[0] + a, y = L->lo[1] + b, z = L->lo[2] + c;
        L->idx[l] = page_set(L, x, y, z) ? (int32_t)nn++ : -1;
      }
  L->nn = nn;
  L->xyz = (int32_t *)malloc((size_t)nn * 3 * sizeof(int32_t));
  L->flags = (uint32_t *)calloc((size_t)nn, sizeof(uint32_t));
  for (int a = 0; a < L->n[0]; ++a)
    for (int b = 0; b < L->n[1]; ++b)
      for (int c = 0; c < L->n[2]; ++c) {
        int64_t l = ((int64_t)a * L->n[1] + b) * L->n[2] + c;
        int32_t i = L->idx[l];
        if (i >= 0) {
          L->xyz[3 * i] = L->lo[0] + a;
          L->xyz[3 * i + 1] = L->lo[1] + b;
          L->xyz[3 * i + 2] = L->lo[2] + c;
        }
      }
  L->nb = (int32_t *)malloc((size_t)nn * 27 * sizeof(int32_t));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nn; ++i) {
    int x = L->xyz[3 * i], y = L->xyz[3 * i + 1], z = L->xyz[3 * i + 2];
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz)
          L->nb[i * 27 + (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)] = node_at(L, x + dx, y + dy, z + dz);
  }
  L->u = (double *)calloc((size_t)nn * 3, sizeof(double));
  L->b = (double *)calloc((size_t)nn * 3, sizeof(double));
  L->r = (double *)calloc((size_t)nn * 3, sizeof(double));
}

/* Flag_Nodes: solver/CPU_Kernels/Flaging.h:127-142 (cells of allocated nodes only) */
static void flag_nodes(lvl_t *L) {
  for (int64_t i = 0; i < L->nn; ++i) {
    int x = L->xyz[3 * i], y = L->xyz[3 * i + 1], z = L->xyz[3 * i + 2];
    int64_t l0 = box_lin(L, x, y, z);
    uint32_t nf = (L->bflags[l0] & (ORC_ACTIVE_CELL | ORC_DIRICHLET_CELL)) >> 4;
    if (!nf) continue;
    for (int c = 0; c < 8; ++c) {
      int64_t l = box_lin(L, x + ((c >> 2) & 1), y + ((c >> 1) & 1), z + (c & 1));
      if (l >= 0) L->bflags[l] |= nf;
    }
  }
}

/* Flag_Boundary_Nodes (band=3): Flaging.h:155-184 */
static void flag_boundary(lvl_t *L) {
  uint8_t *tmp = (uint8_t *)calloc((size_t)L->nn, 1);
  for (int pass = 0; pass < 3; ++pass) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L->nn; ++i) {
      int x = L->xyz[3 * i], y = L->xyz[3 * i + 1], z = L->xyz[3 * i + 2];
      uint32_t f = bflag_at(L, x, y, z);
      tmp[i] = 0;
      if (f & ORC_DIRICHLET_NODE) continue;
      if (!(f & ORC_ACTIVE_NODE)) continue;
      for (int dx = -1; dx <= 1 && !tmp[i]; ++dx)
        for (int dy = -1; dy <= 1 && !tmp[i]; ++dy)
          for (int dz = -1; dz <= 1; ++dz)
            if (bflag_at(L, x + dx, y + dy, z + dz) & (ORC_DIRICHLET_NODE | ORC_BOUNDARY_NODE)) {
              tmp[i] = 1;
              break;
            }
    }
    for (int64_t i = 0; i < L->nn; ++i)
      if (tmp[i]) L->bflags[box_lin(L, L->xyz[3 * i], L->xyz[3 * i + 1], L->xyz[3 * i + 2])] |= ORC_BOUNDARY_NODE;
  }
  free(tmp);
}

static void finalize_flags(lvl_t *L) {
  for (int v = 0; v < 3; ++v) L->ndl[v] = 0;
  for (int64_t i = 0; i < L->nn; ++i) {
    L->flags[i] = L->bflags[box_lin(L, L->xyz[3 * i], L->xyz[3 * i + 1], L->xyz[3 * i + 2])];
    for (int v = 0; v < 3; ++v)
      if (L->flags[i] & (ORC_DIRICHLET_NODE_X << v)) L->ndl[v]++;
  }
  for (int v = 0; v < 3; ++v) {
    L->dl[v] = (int32_t *)malloc((size_t)(L->ndl[v] + 1) * sizeof(int32_t));
    int64_t k = 0;
    for (int64_t i = 0; i < L->nn; ++i)
      if (L->flags[i] & (ORC_DIRICHLET_NODE_X << v)) L->dl[v][k++] = (int32_t)i;
  }
}

/* Dirichlet projection (CG_SYSTEM.cpp:178-191; Multigrid_Hierarchy.h Project calls) */
static void project_dirichlet(const lvl_t *L, double *v) {
  for (int a = 0; a < 3; ++a)
    for (int64_t k = 0; k < L->ndl[a]; ++k) v[3 * (int64_t)L->dl[a][k] + a] = 0.0;
}
/* Project_Null_Space with ACTIVE_NODE: CPU_Kernels/Project_Null_Space.h:28-37 */
static void project_null(const lvl_t *L, double *v) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < L->nn; ++i)
    if (!(L->flags[i] & ORC_ACTIVE_NODE)) v[3 * i] = v[3 * i + 1] = v[3 * i + 2] = 0.0;
}
static void cast_float(double *v, int64_t n, int mode) {
  if (mode != ORC_MODE_REF) return;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) v[i] = (double)(float)v[i];
}

/* ------------------------------------------------------------------------- */
/* Matrix-free level-0 operators                                               */
/* ------------------------------------------------------------------------- */
/* row of K at node n: f = dx * sum_I sum_J (mu_c Kmu[I][J] + la_c Kla[I][J]) u(n-I+J),
 * c = n - I  (SPGrid_Linear_Elasticity.h:22-94, 96-213); diag accumulates the
 * 3x3 self block (SPGrid_Gauss_Seidel.h:765-875). */
static void k_row(const orc_system *S, int64_t n, const double *u, double f[3], double diag[9]) {
  const lvl_t *L = &S->lv[0];
  const int32_t *nb = L->nb + n * 27;
  double acc[3] = {0, 0, 0};
  double dg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int I = 0; I < 8; ++I) {
    int ix = (I >> 2) & 1, iy = (I >> 1) & 1, iz = I & 1;
    int32_t c = nb[(1 - ix) * 9 + (1 - iy) * 3 + (1 - iz)];
    if (c < 0) continue;
    double mu = S->mu[c], la = S->la[c];
    if (mu == 0.0 && la == 0.0) continue;
    for (int J = 0; J < 8; ++J) {
      int jx = (J >> 2) & 1, jy = (J >> 1) & 1, jz = J & 1;
      int32_t m = nb[(jx - ix + 1) * 9 + (jy - iy + 1) * 3 + (jz - iz + 1)];
      const double *km = S->kmu + (I * 8 + J) * 9, *kl = S->kla + (I * 8 + J) * 9;
      if (m >= 0) {
        const double *um = u + 3 * (int64_t)m;
        for (int v = 0; v < 3; ++v)
          for (int w = 0; w < 3; ++w) acc[v] += (mu * km[v * 3 + w] + la * kl[v * 3 + w]) * um[w];
      }
      if (diag && J == I)
        for (int e = 0; e < 9; ++e) dg[e] += mu * km[e] + la * kl[e];
    }
  }
  double dx = S->p.dx;
  for (int v = 0; v < 3; ++v) f[v] = dx * acc[v];
  if (diag)
    for (int e = 0; e < 9; ++e) diag[e] = dx * dg[e];
}

void orc_apply(const orc_system *S, const double *u, double *f) {
  const lvl_t *L = &S->lv[0];
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < L->nn; ++n) k_row(S, n, u, f + 3 * n, NULL);
}

/* 3x3 symmetric inverse via minors (Multigrid_Hierarchy.h:64-101), with the
 * Dirichlet handling of SPGrid_Gauss_Seidel.h:1093-1144; d = {xx,yy,zz,xy,xz,yz} */
static void dirichlet_inverse(double d[6], uint32_t fl) {
  if (fl & ORC_DIRICHLET_NODE_X) { d[0] = 1.0; d[3] = 0; d[4] = 0; }
  if (fl & (ORC_DIRICHLET_NODE_X << 1)) { d[1] = 1.0; d[3] = 0; d[5] = 0; }
  if (fl & (ORC_DIRICHLET_NODE_X << 2)) { d[2] = 1.0; d[4] = 0; d[5] = 0; }
  double m11 = d[1] * d[2] - d[5] * d[5], m12 = d[5] * d[4] - d[2] * d[3], m13 = d[3] * d[5] - d[4] * d[1];
  double m22 = d[2] * d[0] - d[4] * d[4], m23 = d[4] * d[3] - d[0] * d[5], m33 = d[0] * d[1] - d[3] * d[3];
  double det = d[0] * m11 + d[3] * m12 + d[4] * m13;
  double inv = 1.0 / det;
  d[0] = m11 * inv; d[1] = m22 * inv; d[2] = m33 * inv; d[3] = m12 * inv; d[4] = m13 * inv; d[5] = m23 * inv;
  if (fl & ORC_DIRICHLET_NODE_X) { d[0] = 0; d[3] = 0; d[4] = 0; }
  if (fl & (ORC_DIRICHLET_NODE_X << 1)) { d[1] = 0; d[3] = 0; d[5] = 0; }
  if (fl & (ORC_DIRICHLET_NODE_X << 2)) { d[2] = 0; d[4] = 0; d[5] = 0; }
}

static inline int colour_of(const int32_t *xyz) {
  return ((xyz[0] & 1) << 2) | ((xyz[1] & 1) << 1) | (xyz[2] & 1);
}

/* one colour of the level-0 matrix-free GS: SPGrid_Gauss_Seidel.h:877-1183 */
static void gs0_colour(const orc_system *S, double *u, const double *b, int colour, int boundary) {
  const lvl_t *L = &S->lv[0];
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < L->nn; ++n) {
    if (colour_of(L->xyz + 3 * n) != colour) continue;
    uint32_t fl = L->flags[n];
    double f[3], dg[9];
    int update = (fl & ORC_ACTIVE_NODE) && (!boundary || (fl & ORC_BOUNDARY_NODE));
    if (update) {
      k_row(S, n, u, f, dg);
      double r[3] = {b[3 * n] - f[0], b[3 * n + 1] - f[1], b[3 * n + 2] - f[2]};
      double d[6] = {dg[0], dg[4], dg[8], dg[1], dg[2], dg[5]};
      dirichlet_inverse(d, fl);
      u[3 * n + 0] += d[0] * r[0] + d[3] * r[1] + d[4] * r[2];
      u[3 * n + 1] += d[3] * r[0] + d[1] * r[1] + d[5] * r[2];
      u[3 * n + 2] += d[4] * r[0] + d[5] * r[1] + d[2] * r[2];
    }
    for (int v = 0; v < 3; ++v)
      if (fl & (ORC_DIRICHLET_NODE_X << v)) u[3 * n + v] = 0.0;
  }
}

/* ------------------------------------------------------------------------- */
/* Stencil operators on explicit levels                                         */
/* ------------------------------------------------------------------------- */
static void st_row(const lvl_t *L, int64_t n, const double *u, double f[3]) {
  const double *st = L->st + n * 243;
  const int32_t *nb = L->nb + n * 27;
  double acc[3] = {0, 0, 0};
  for (int s = 0; s < 27; ++s) {
    int32_t m = nb[s];
    if (m < 0) continue;
    const double *um = u + 3 * (int64_t)m;
    for (int v = 0; v < 3; ++v)
      for (int w = 0; w < 3; ++w) acc[v] += um[w] * st[s * 9 + v * 3 + w];
  }
  f[0] = acc[0]; f[1] = acc[1]; f[2] = acc[2];
}

/* SPGrid_Stencil_Colored_Smooth: MG/SPGrid_Stencil_Multiply.h:161-279 */
static void gs_st_colour(const lvl_t *L, double *u, const double *b, int colour, int boundary) {
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < L->nn; ++n) {
    if (colour_of(L->xyz + 3 * n) != colour) continue;
    uint32_t fl = L->flags[n];
    if (!(fl & ORC_ACTIVE_NODE)) continue;
    if (boundary && !(fl & ORC_BOUNDARY_NODE)) continue;
    const double *st = L->st + n * 243 + 13 * 9;
    double d[6] = {st[0], st[4], st[8], st[1], st[2], st[5]};
    dirichlet_inverse(d, fl);
    double f[3];
    st_row(L, n, u, f);
    double r[3] = {b[3 * n] - f[0], b[3 * n + 1] - f[1], b[3 * n + 2] - f[2]};
    u[3 * n + 0] += d[0] * r[0] + d[3] * r[1] + d[4] * r[2];
    u[3 * n + 1] += d[3] * r[0] + d[1] * r[1] + d[5] * r[2];
    u[3 * n + 2] += d[4] * r[0] + d[5] * r[1] + d[2] * r[2];
  }
  project_dirichlet(L, u);
}

void orc_gs_sweep(const orc_system *S, int level, double *u, const double *b, int forward, int boundary) {
  /* colour orders: Multigrid_Hierarchy.h:512-530 (level 0), :533-536 (stencil levels) */
  static const int order_st[2][8] = {{1, 5, 4, 6, 7, 3, 2, 0}, {0, 2, 3, 7, 6, 4, 5, 1}};
  if (level == 0) {
    for (int c = 0; c < 8; ++c) gs0_colour(S, u, b, forward ? c : 7 - c, boundary);
  } else {
    const lvl_t *L = &S->lv[level];
    for (int c = 0; c < 8; ++c) gs_st_colour(L, u, b, order_st[forward ? 0 : 1][c], boundary);
  }
}

void orc_stencil_residual(const orc_system *S, int level, const double *u, const double *b, double *r) {
  const lvl_t *L = &S->lv[level];
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < L->nn; ++n) {
    double f[3];
    st_row(L, n, u, f);
    for (int v = 0; v < 3; ++v) r[3 * n + v] = b[3 * n + v] - f[v];
  }
  project_dirichlet(L, r);
}

/* ------------------------------------------------------------------------- */
/* Transfers                                                                   */
/* ------------------------------------------------------------------------- */
void orc_restrict(const orc_system *S, int cl, const double *fine, double *coarse) {
  const lvl_t *C = &S->lv[cl], *F = &S->lv[cl - 1];
  static const double w[3] = {0.5, 1.0, 0.5};
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < C->nn; ++n) {
    int fx = 2 * C->xyz[3 * n] - BPAD[0], fy = 2 * C->xyz[3 * n + 1] - BPAD[1], fz = 2 * C->xyz[3 * n + 2] - BPAD[2];
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          int32_t m = node_at(F, fx + i - 1, fy + j - 1, fz + k - 1);
          if (m < 0) continue;
          double wt = w[i] * w[j] * w[k];
          for (int v = 0; v < 3; ++v) acc[v] += wt * fine[3 * (int64_t)m + v];
        }
    for (int v = 0; v < 3; ++v) coarse[3 * n + v] = acc[v];
  }
  project_dirichlet(C, coarse);
}

void orc_prolongate(const orc_system *S, int fl, double *fine, const double *coarse) {
  const lvl_t *F = &S->lv[fl], *C = &S->lv[fl + 1];
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < F->nn; ++n) {
    int x[3] = {F->xyz[3 * n], F->xyz[3 * n + 1], F->xyz[3 * n + 2]};
    int c0[3];
    double wt[3][2];
    for (int d = 0; d < 3; ++d) {
      c0[d] = (x[d] + BPAD[d]) >> 1;
      if (x[d] & 1) { wt[d][0] = 0.5; wt[d][1] = 0.5; }
      else { wt[d][0] = 1.0; wt[d][1] = 0.0; }
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) {
          double w = wt[0][i] * wt[1][j] * wt[2][k];
          if (w <= 0) continue;
          int32_t m = node_at(C, c0[0] + i, c0[1] + j, c0[2] + k);
          if (m < 0) continue;
          for (int v = 0; v < 3; ++v) fine[3 * n + v] += w * coarse[3 * (int64_t)m + v];
        }
  }
  project_dirichlet(F, fine);
}

/* ------------------------------------------------------------------------- */
/* Galerkin stencils (MG/Galerkin_Coarsen.h:40-100 with the recursive cell     */
/* transfer of Linear_Elasticity_SIMD_Cell_Multiply.h:106-161, expressed as     */
/* per-position matrices P_p^T K P_p: trilinear restricted to a fine cell is    */
/* the composition of the reference's per-level cell prolongations).            */
/* ------------------------------------------------------------------------- */
static void position_matrix(const double *K, int depth, const int p[3], double *M) {
  int ncell = 1 << depth;
  double W[8][8]; /* W[a][A]: weight of coarse node A at fine node a */
  for (int a = 0; a < 8; ++a) {
    int ax[3] = {(a >> 2) & 1, (a >> 1) & 1, a & 1};
    for (int A = 0; A < 8; ++A) {
      int Ax[3] = {(A >> 2) & 1, (A >> 1) & 1, A & 1};
      double w = 1.0;
      for (int d = 0; d < 3; ++d) {
        double t = (double)(p[d] + ax[d]) / (double)ncell;
        w *= Ax[d] ? t : (1.0 - t);
      }
      W[a][A] = w;
    }
  }
  for (int A = 0; A < 8; ++A)
    for (int B = 0; B < 8; ++B)
      for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w) {
          double acc = 0.0;
          for (int a = 0; a < 8; ++a) {
            if (W[a][A] == 0.0) continue;
            for (int b = 0; b < 8; ++b) {
              if (W[b][B] == 0.0) continue;
              acc += W[a][A] * K[((a * 8 + b) * 3 + v) * 3 + w] * W[b][B];
            }
          }
          M[((A * 8 + B) * 3 + v) * 3 + w] = acc;
        }
}

static void build_stencils(orc_system *S, int level) {
  lvl_t *L = &S->lv[level];
  const lvl_t *L0 = &S->lv[0];
  int depth = level;
  int ncell_axis = 1 << depth;
  int64_t npos = (int64_t)ncell_axis * ncell_axis * ncell_axis;
  double *Mmu = (double *)malloc((size_t)npos * 576 * sizeof(double));
  double *Mla = (double *)malloc((size_t)npos * 576 * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < npos; ++q) {
    int p[3] = {(int)(q / (ncell_axis * ncell_axis)), (int)((q / ncell_axis) % ncell_axis), (int)(q % ncell_axis)};
    position_matrix(S->kmu, depth, p, Mmu + q * 576);
    position_matrix(S->kla, depth, p, Mla + q * 576);
  }
  L->st = (double *)calloc((size_t)L->nn * 243, sizeof(double));
  L->spoke = (uint32_t *)calloc((size_t)L->nn, sizeof(uint32_t));
  L->nb_stencil = 1;
  double dx = S->p.dx;
  /* each coarse cell's 24x24 matrix once (cell at its allocated min-corner node) */
  int32_t *crow = (int32_t *)malloc((size_t)L->nn * sizeof(int32_t));
  int64_t ncell = 0;
  for (int64_t n = 0; n < L->nn; ++n) {
    int64_t bl = box_lin(L, L->xyz[3 * n], L->xyz[3 * n + 1], L->xyz[3 * n + 2]);
    crow[n] = (L->bflags[bl] & ORC_ACTIVE_CELL) ? (int32_t)ncell++ : -1;
  }
  double *CA = (double *)calloc((size_t)(ncell > 0 ? ncell : 1) * 576, sizeof(double));
  uint8_t *cany = (uint8_t *)calloc((size_t)(ncell > 0 ? ncell : 1), 1);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t n = 0; n < L->nn; ++n) {
    if (crow[n] < 0) continue;
    double *A = CA + (int64_t)crow[n] * 576;
    int f0[3];
    for (int d = 0; d < 3; ++d) f0[d] = ncell_axis * L->xyz[3 * n + d] - (ncell_axis - 1) * BPAD[d];
    for (int64_t q = 0; q < npos; ++q) {
      int p[3] = {(int)(q / (ncell_axis * ncell_axis)), (int)((q / ncell_axis) % ncell_axis), (int)(q % ncell_axis)};
      int32_t c = node_at(L0, f0[0] + p[0], f0[1] + p[1], f0[2] + p[2]);
      if (c < 0) continue;
      double mu = S->mu[c], la = S->la[c];
      if (mu == 0.0 && la == 0.0) continue;
      cany[crow[n]] = 1;
      const double *mm = Mmu + q * 576, *ml = Mla + q * 576;
      for (int e = 0; e < 576; ++e) A[e] += mu * mm[e] + la * ml[e];
    }
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t n = 0; n < L->nn; ++n) {
    if (!(L->flags[n] & ORC_ACTIVE_NODE)) continue;
    double *st = L->st + n * 243;
    uint32_t spk = 0;
    for (int pc = 0; pc < 8; ++pc) {
      int px = (pc >> 2) & 1, py = (pc >> 1) & 1, pz = pc & 1;
      int32_t cn = node_at(L, L->xyz[3 * n] + px - 1, L->xyz[3 * n + 1] + py - 1, L->xyz[3 * n + 2] + pz - 1);
      if (cn < 0 || crow[cn] < 0 || !cany[crow[cn]]) continue;
      const double *A = CA + (int64_t)crow[cn] * 576;
      int I = ((1 - px) << 2) | ((1 - py) << 1) | (1 - pz); /* our node within the cell */
      for (int J = 0; J < 8; ++J) {
        int jx = (J >> 2) & 1, jy = (J >> 1) & 1, jz = J & 1;
        int s = (px + jx) * 9 + (py + jy) * 3 + (pz + jz);
        int nz = 0;
        for (int v = 0; v < 3; ++v)
          for (int w = 0; w < 3; ++w) {
            double a = A[((I * 8 + J) * 3 + v) * 3 + w];
            if (a != 0.0) nz = 1;
            st[s * 9 + v * 3 + w] += a;
          }
        if (nz) spk |= 1u << s;
      }
    }
    for (int e = 0; e < 243; ++e) st[e] *= dx;
    L->spoke[n] = spk;
  }
  free(CA);
  free(cany);
  free(crow);
  /* Delete_Dirichlet_Equations: MG/Delete_Dirichlet_Equations.h:24-51 */
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < L->nn; ++n) {
    if (!(L->flags[n] & ORC_ACTIVE_NODE)) continue;
    uint32_t fl = L->flags[n];
    double *st = L->st + n * 243;
    for (int s = 0; s < 27; ++s) {
      if (!(L->spoke[n] & (1u << s))) continue;
      int32_t m = L->nb[n * 27 + s];
      uint32_t mfl = (m >= 0 && (L->flags[m] & ORC_ACTIVE_NODE)) ? L->flags[m] : 0u;
      for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w) {
          if (fl & (ORC_DIRICHLET_NODE_X << v)) st[s * 9 + v * 3 + w] = (v == w && s == 13) ? 1.0 : 0.0;
          else if (mfl & (ORC_DIRICHLET_NODE_X << w)) st[s * 9 + v * 3 + w] = 0.0;
        }
    }
  }
  free(Mmu);
  free(Mla);
}

int orc_get_stencils(const orc_system *S, int level, double *out) {
  const lvl_t *L = &S->lv[level];
  if (!L->nb_stencil) return -1;
  memcpy(out, L->st, (size_t)L->nn * 243 * sizeof(double));
  return 0;
}

/* Bottom solve: exact SPD solve of the Dirichlet-deleted stencil matrix over
 * nodes with an active spoke (Elasticity_Matrix.cpp:15-164 builds the same
 * matrix; MKL PARDISO factorizes it there -- here an envelope Cholesky). */
static int factor_bottom(lvl_t *L, char *err, int errlen) {
  L->var_of = (int32_t *)malloc((size_t)L->nn * sizeof(int32_t));
  int64_t nv = 0;
  for (int64_t n = 0; n < L->nn; ++n) {
    if (L->spoke[n]) { L->var_of[n] = (int32_t)nv; nv += 3; }
    else L->var_of[n] = -1;
  }
  L->nvar = nv;
  L->fc = (int32_t *)malloc((size_t)(nv + 1) * sizeof(int32_t));
  for (int64_t n = 0; n < L->nn; ++n) {
    if (L->var_of[n] < 0) continue;
    int32_t mn = L->var_of[n];
    for (int s = 0; s < 27; ++s) {
      if (!(L->spoke[n] & (1u << s))) continue;
      int32_t m = L->nb[n * 27 + s];
      if (m >= 0 && L->var_of[m] >= 0 && L->var_of[m] < mn) mn = L->var_of[m];
    }
    for (int v = 0; v < 3; ++v) L->fc[L->var_of[n] + v] = mn;
  }
  L->rowstart = (int64_t *)malloc((size_t)(nv + 1) * sizeof(int64_t));
  L->rowstart[0] = 0;
  for (int64_t i = 0; i < nv; ++i) L->rowstart[i + 1] = L->rowstart[i] + (i - L->fc[i] + 1);
  L->L = (double *)calloc((size_t)L->rowstart[nv], sizeof(double));
  /* scatter lower triangle */
  for (int64_t n = 0; n < L->nn; ++n) {
    if (L->var_of[n] < 0) continue;
    const double *st = L->st + n * 243;
    for (int s = 0; s < 27; ++s) {
      if (!(L->spoke[n] & (1u << s))) continue;
      int32_t m = L->nb[n * 27 + s];
      if (m < 0 || L->var_of[m] < 0) continue;
      for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w) {
          int64_t i = L->var_of[n] + v, j = L->var_of[m] + w;
          if (j > i) continue;
          L->L[L->rowstart[i] + (j - L->fc[i])] = st[s * 9 + v * 3 + w];
        }
    }
  }
  /* envelope Cholesky, row by row */
  for (int64_t i = 0; i < nv; ++i) {
    double *Li = L->L + L->rowstart[i];
    int64_t fi = L->fc[i];
    for (int64_t j = fi; j < i; ++j) {
      double *Lj = L->L + L->rowstart[j];
      int64_t fj = L->fc[j];
      int64_t k0 = fi > fj ? fi : fj;
      double sum = Li[j - fi];
      for (int64_t k = k0; k < j; ++k) sum -= Li[k - fi] * Lj[k - fj];
      Li[j - fi] = sum / Lj[j - fj];
    }
    double sum = Li[i - fi];
    for (int64_t k = fi; k < i; ++k) sum -= Li[k - fi] * Li[k - fi];
    if (!(sum > 0.0)) {
      snprintf(err, errlen, "bottom Cholesky: matrix not positive definite at row %lld", (long long)i);
      return -1;
    }
    Li[i - fi] = sqrt(sum);
  }
  return 0;
}

static void bottom_solve(const lvl_t *L, const double *b, double *u) {
  int64_t nv = L->nvar;
  double *y = (double *)malloc((size_t)(nv > 0 ? nv : 1) * sizeof(double));
  for (int64_t n = 0; n < L->nn; ++n)
    if (L->var_of[n] >= 0)
      for (int v = 0; v < 3; ++v) y[L->var_of[n] + v] = b[3 * n + v];
  for (int64_t i = 0; i < nv; ++i) {
    const double *Li = L->L + L->rowstart[i];
    int64_t fi = L->fc[i];
    double s = y[i];
    for (int64_t k = fi; k < i; ++k) s -= Li[k - fi] * y[k];
    y[i] = s / Li[i - fi];
  }
  for (int64_t i = nv - 1; i >= 0; --i) {
    const double *Li = L->L + L->rowstart[i];
    int64_t fi = L->fc[i];
    y[i] /= Li[i - fi];
    for (int64_t k = fi; k < i; ++k) y[k] -= Li[k - fi] * y[i];
  }
  /* Exact_Solve copies b through the stencil array; nodes without a variable
   * keep b (Multigrid_Hierarchy.h:346-372). */
  for (int64_t n = 0; n < L->nn; ++n)
    for (int v = 0; v < 3; ++v) u[3 * n + v] = (L->var_of[n] >= 0) ? y[L->var_of[n] + v] : b[3 * n + v];
  free(y);
}

/* ------------------------------------------------------------------------- */
/* Build                                                                       */
/* ------------------------------------------------------------------------- */
static int cmp3(const void *a, const void *b) {
  const int32_t *x = (const int32_t *)a, *y = (const int32_t *)b;
  for (int d = 0; d < 4; ++d)
    if (x[d] != y[d]) return x[d] < y[d] ? -1 : 1;
  return 0;
}

orc_system *orc_build(const orc_problem *p, char *err, int errlen) {
  orc_system *S = (orc_system *)calloc(1, sizeof(orc_system));
  S->p = *p;
  orc_element_tables(S->kmu, S->kla);
  if (p->nblocks <= 0) { snprintf(err, errlen, "no density blocks"); free(S); return NULL; }
  int L_req = p->mg_levels;
  S->use_mg = L_req != 0;
  S->nlev = L_req == 0 ? 1 : L_req;
  if (S->nlev == 2) {
    /* The reference crashes here (no matrix for the pass-through level 1,
     * Multigrid_Hierarchy.h:215-226 + 342-361); the caller maps 2 -> 3
     * (spgrid_topo_opt.cpp:1033-1037). */
    snprintf(err, errlen, "mg_level=2 is not supported (level 1 is a pass-through level)");
    free(S);
    return NULL;
  }
  if (S->nlev > 15) { snprintf(err, errlen, "too many levels"); free(S); return NULL; }
  S->explicit_level = 2;

  /* ---- level 0 pages from block bases (SPGrid_CPU_Solver.cpp:107-130) ---- */
  lvl_t *L0 = &S->lv[0];
  int pmin[3] = {1 << 30, 1 << 30, 1 << 30}, pmax[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  for (int64_t bk = 0; bk < p->nblocks; ++bk)
    for (int d = 0; d < 3; ++d) {
      int pc = floordiv(p->block_base[3 * bk + d], PAGE[d]);
      if (pc < pmin[d]) pmin[d] = pc;
      if (pc > pmax[d]) pmax[d] = pc;
    }
  level_box(L0, pmin, pmax);
  for (int64_t bk = 0; bk < p->nblocks; ++bk)
    mark_page(L0, p->block_base[3 * bk], p->block_base[3 * bk + 1], p->block_base[3 * bk + 2]);
  /* padding check: Multigrid_Hierarchy.h:145-157 (page origin in [B, res-2B)) */
  for (int px = 0; px < L0->pn[0]; ++px)
    for (int py = 0; py < L0->pn[1]; ++py)
      for (int pz = 0; pz < L0->pn[2]; ++pz) {
        if (!L0->page[((int64_t)px * L0->pn[1] + py) * L0->pn[2] + pz]) continue;
        int o[3] = {(L0->plo[0] + px) * PAGE[0], (L0->plo[1] + py) * PAGE[1], (L0->plo[2] + pz) * PAGE[2]};
        for (int d = 0; d < 3; ++d)
          if (o[d] < BPAD[d] || o[d] >= p->res[d] - 2 * BPAD[d]) {
            snprintf(err, errlen, "padding violated: page origin (%d,%d,%d)", o[0], o[1], o[2]);
            orc_free(S);
            return NULL;
          }
      }
  level_enumerate(L0);
  S->mu = (double *)calloc((size_t)L0->nn, sizeof(double));
  S->la = (double *)calloc((size_t)L0->nn, sizeof(double));
  S->rho = (double *)calloc((size_t)L0->nn, sizeof(double));
  S->has_rho = (uint8_t *)calloc((size_t)L0->nn, 1);
  for (int64_t bk = 0; bk < p->nblocks; ++bk)
    for (int c = 0; c < 64; ++c) {
      int x = p->block_base[3 * bk] + (c >> 4), y = p->block_base[3 * bk + 1] + ((c >> 2) & 3),
          z = p->block_base[3 * bk + 2] + (c & 3);
      int32_t n = node_at(L0, x, y, z);
      if (n < 0) continue;
      double d = p->density[bk * 64 + c];
      S->rho[n] = d;
      S->has_rho[n] = 1;
      if (d > 0) {
        double e = orc_stiffness_mapping(p->penalty, p->min_stiffness, p->min_density, d);
        S->mu[n] = e * p->mu0;
        S->la[n] = e * p->lambda0;
      }
    }
  /* ACTIVE_CELL: Flaging.h:93-106 */
  for (int64_t n = 0; n < L0->nn; ++n)
    if (S->mu[n] != 0.0)
      L0->bflags[box_lin(L0, L0->xyz[3 * n], L0->xyz[3 * n + 1], L0->xyz[3 * n + 2])] |= ORC_ACTIVE_CELL;
  /* validation: Validate_Blocks.h:46-69 */
  for (int64_t n = 0; n < L0->nn; ++n) {
    if (!(S->mu[n] > 0)) continue;
    for (int c = 1; c < 8; ++c)
      if (!page_set(L0, L0->xyz[3 * n] + ((c >> 2) & 1), L0->xyz[3 * n + 1] + ((c >> 1) & 1), L0->xyz[3 * n + 2] + (c & 1))) {
        snprintf(err, errlen, "block activation check failed at (%d,%d,%d)", L0->xyz[3 * n], L0->xyz[3 * n + 1],
                 L0->xyz[3 * n + 2]);
        orc_free(S);
        return NULL;
      }
  }
  /* Dirichlet nodes -> cells: SPGrid_CPU_Solver.cpp:147-192 */
  S->nd = p->ndir;
  S->d_node = (int32_t *)malloc((size_t)(p->ndir + 1) * sizeof(int32_t));
  S->d_axis = (int32_t *)malloc((size_t)(p->ndir + 1) * sizeof(int32_t));
  S->d_val = (double *)malloc((size_t)(p->ndir + 1) * sizeof(double));
  S->d_xyz = (int32_t *)malloc((size_t)(p->ndir + 1) * 4 * sizeof(int32_t));
  {
    int32_t *keys = (int32_t *)malloc((size_t)(p->ndir + 1) * 4 * sizeof(int32_t));
    for (int64_t i = 0; i < p->ndir; ++i) {
      keys[4 * i] = p->dir_axis[i];
      keys[4 * i + 1] = p->dir_coord[3 * i];
      keys[4 * i + 2] = p->dir_coord[3 * i + 1];
      keys[4 * i + 3] = p->dir_coord[3 * i + 2];
      S->d_node[i] = node_at(L0, p->dir_coord[3 * i], p->dir_coord[3 * i + 1], p->dir_coord[3 * i + 2]);
      S->d_axis[i] = p->dir_axis[i];
      S->d_val[i] = p->dir_value[i];
      S->d_xyz[4 * i] = p->dir_axis[i];
      memcpy(S->d_xyz + 4 * i + 1, p->dir_coord + 3 * i, 3 * sizeof(int32_t));
    }
    qsort(keys, (size_t)p->ndir, 4 * sizeof(int32_t), cmp3);
    for (int64_t i = 0; i < p->ndir; ++i) {
      int v = p->dir_axis[i];
      if (v < 0 || v > 2) continue;
      int all = 1;
      for (int c = 1; c < 8 && all; ++c) {
        int32_t key[4] = {v, p->dir_coord[3 * i] + ((c >> 2) & 1), p->dir_coord[3 * i + 1] + ((c >> 1) & 1),
                          p->dir_coord[3 * i + 2] + (c & 1)};
        if (!bsearch(key, keys, (size_t)p->ndir, 4 * sizeof(int32_t), cmp3)) all = 0;
      }
      if (all) {
        int64_t l = box_lin(L0, p->dir_coord[3 * i], p->dir_coord[3 * i + 1], p->dir_coord[3 * i + 2]);
        if (l >= 0) L0->bflags[l] |= (ORC_DIRICHLET_CELL_X << v);
      }
    }
    free(keys);
  }
  /* ---- coarse levels: Multigrid_Hierarchy.h:122-144 (sizes), 158-195 ---- */
  for (int l = 1; l < S->nlev; ++l) {
    lvl_t *F = &S->lv[l - 1], *C = &S->lv[l];
    int cmin[3], cmax[3];
    for (int d = 0; d < 3; ++d) {
      int lo = F->lo[d], hi = F->lo[d] + F->n[d] - 1;
      cmin[d] = floordiv((lo + BPAD[d]) >> 1, PAGE[d]);
      cmax[d] = floordiv(((hi + BPAD[d]) >> 1) + 1, PAGE[d]);
    }
    level_box(C, cmin, cmax);
    for (int64_t n = 0; n < F->nn; ++n) {
      int x = F->xyz[3 * n], y = F->xyz[3 * n + 1], z = F->xyz[3 * n + 2];
      uint32_t cf = F->bflags[box_lin(F, x, y, z)];
      int X = (x + BPAD[0]) >> 1, Y = (y + BPAD[1]) >> 1, Z = (z + BPAD[2]) >> 1;
      int64_t cl = box_lin(C, X, Y, Z);
      if (cf & ORC_ACTIVE_CELL) {
        C->bflags[cl] |= ORC_ACTIVE_CELL;
        for (int c = 0; c < 8; ++c) mark_page(C, X + ((c >> 2) & 1), Y + ((c >> 1) & 1), Z + (c & 1));
      }
      for (int v = 0; v < 3; ++v)
        if (cf & (ORC_DIRICHLET_CELL_X << v)) {
          C->bflags[cl] |= (ORC_DIRICHLET_CELL_X << v);
          for (int c = 1; c < 8; ++c) mark_page(C, X + ((c >> 2) & 1), Y + ((c >> 1) & 1), Z + (c & 1));
        }
    }
    level_enumerate(C);
  }
  for (int l = 0; l < S->nlev; ++l) flag_nodes(&S->lv[l]);
  for (int l = 0; l < S->nlev; ++l) flag_boundary(&S->lv[l]);
  for (int l = 0; l < S->nlev; ++l) finalize_flags(&S->lv[l]);
  /* forces */
  S->f_node = (int32_t *)malloc((size_t)(p->nforce + 1) * sizeof(int32_t));
  for (int64_t i = 0; i < p->nforce; ++i)
    S->f_node[i] = node_at(L0, p->force_coord[3 * i], p->force_coord[3 * i + 1], p->force_coord[3 * i + 2]);
  /* stencils + bottom factorization: Multigrid_Hierarchy.h:215-240, 332-341 */
  if (S->use_mg) {
    int first = S->nlev == 1 ? 0 : S->explicit_level;
    for (int l = first; l < S->nlev; ++l) build_stencils(S, l);
    if (factor_bottom(&S->lv[S->nlev - 1], err, errlen) != 0) {
      orc_free(S);
      return NULL;
    }
  }
  int64_t n0 = L0->nn * 3;
  S->x = (double *)calloc((size_t)n0, sizeof(double));
  S->rr = (double *)calloc((size_t)n0, sizeof(double));
  S->s = (double *)calloc((size_t)n0, sizeof(double));
  S->q = (double *)calloc((size_t)n0, sizeof(double));
  S->z = (double *)calloc((size_t)n0, sizeof(double));
  S->fvec = (double *)calloc((size_t)n0, sizeof(double));
  S->ubase = (double *)calloc((size_t)n0, sizeof(double));
  return S;
}

void orc_free(orc_system *S) {
  if (!S) return;
  for (int l = 0; l < 16; ++l) {
    lvl_t *L = &S->lv[l];
    free(L->page); free(L->idx); free(L->bflags); free(L->xyz); free(L->nb); free(L->flags);
    free(L->st); free(L->spoke); free(L->u); free(L->b); free(L->r);
    for (int v = 0; v < 3; ++v) free(L->dl[v]);
    free(L->var_of); free(L->rowstart); free(L->fc); free(L->L);
  }
  free(S->mu); free(S->la); free(S->rho); free(S->has_rho);
  free(S->d_node); free(S->d_axis); free(S->d_val); free(S->d_xyz); free(S->f_node);
  free(S->x); free(S->rr); free(S->s); free(S->q); free(S->z); free(S->fvec); free(S->ubase);
  free(S);
}

int orc_num_levels(const orc_system *S) { return S->nlev; }
int64_t orc_num_nodes(const orc_system *S, int level) { return S->lv[level].nn; }
void orc_node_xyz(const orc_system *S, int level, int32_t *xyz) {
  memcpy(xyz, S->lv[level].xyz, (size_t)S->lv[level].nn * 3 * sizeof(int32_t));
}
void orc_node_flags(const orc_system *S, int level, uint32_t *flags) {
  memcpy(flags, S->lv[level].flags, (size_t)S->lv[level].nn * sizeof(uint32_t));
}
void orc_cell_moduli(const orc_system *S, double *mu, double *la) {
  memcpy(mu, S->mu, (size_t)S->lv[0].nn * sizeof(double));
  memcpy(la, S->la, (size_t)S->lv[0].nn * sizeof(double));
}

/* ------------------------------------------------------------------------- */
/* V-cycle: Multigrid_Hierarchy.h:569-619                                      */
/* ------------------------------------------------------------------------- */
static void residual0(const orc_system *S, const double *u, const double *b, double *r) {
  /* SPGrid_Linear_Elasticity::Residual then Project (CG_SYSTEM.cpp:118-130) */
  const lvl_t *L = &S->lv[0];
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < L->nn; ++n) {
    double f[3];
    k_row(S, n, u, f, NULL);
    for (int v = 0; v < 3; ++v) r[3 * n + v] = b[3 * n + v] - f[v];
  }
  project_dirichlet(L, r);
}

static void smooth_level(orc_system *S, int l, double *u, const double *b, int forward, int mode) {
  const lvl_t *L = &S->lv[l];
  int B = S->p.boundary_iters, I = S->p.smooth_iters;
  for (int it = 0; it < B; ++it) { orc_gs_sweep(S, l, u, b, forward, 1); cast_float(u, 3 * L->nn, mode); }
  for (int it = 0; it < I; ++it) { orc_gs_sweep(S, l, u, b, forward, 0); cast_float(u, 3 * L->nn, mode); }
  for (int it = 0; it < B; ++it) { orc_gs_sweep(S, l, u, b, forward, 1); cast_float(u, 3 * L->nn, mode); }
}

void orc_vcycle(orc_system *S, const double *b0, double *z, int mode) {
  int nl = S->nlev, ex = S->explicit_level;
  lvl_t *L0 = &S->lv[0];
  memcpy(L0->b, b0, (size_t)L0->nn * 3 * sizeof(double));
  memset(L0->u, 0, (size_t)L0->nn * 3 * sizeof(double));
  for (int l = 0; l < nl - 1; ++l) {
    lvl_t *L = &S->lv[l], *C = &S->lv[l + 1];
    int smoothed = (l == 0 || l >= ex);
    if (smoothed) smooth_level(S, l, L->u, L->b, 1, mode);
    memset(C->u, 0, (size_t)C->nn * 3 * sizeof(double));
    memset(C->b, 0, (size_t)C->nn * 3 * sizeof(double));
    if (smoothed) {
      if (l == 0) residual0(S, L->u, L->b, L->r);
      else orc_stencil_residual(S, l, L->u, L->b, L->r);
      cast_float(L->r, 3 * L->nn, mode);
      orc_restrict(S, l + 1, L->r, C->b);
    } else {
      orc_restrict(S, l + 1, L->b, C->b);
    }
  }
  lvl_t *Lb = &S->lv[nl - 1];
  bottom_solve(Lb, Lb->b, Lb->u);
  cast_float(Lb->u, 3 * Lb->nn, mode);
  for (int l = nl - 2; l >= 0; --l) {
    lvl_t *L = &S->lv[l];
    orc_prolongate(S, l, L->u, S->lv[l + 1].u);
    cast_float(L->u, 3 * L->nn, mode);
    project_null(L, L->u);
    if (l == 0 || l >= ex) smooth_level(S, l, L->u, L->b, 0, mode);
  }
  memcpy(z, L0->u, (size_t)L0->nn * 3 * sizeof(double));
}

/* ------------------------------------------------------------------------- */
/* PCG: CG/CONJUGATE_GRADIENT.cpp:21-61 with CG_SYSTEM/CG_VECTOR semantics      */
/* ------------------------------------------------------------------------- */
static double dot(const double *a, const double *b, int64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
static double linf(const double *a, int64_t n) {
  double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double v = fabs(a[i]);
    if (v > m) m = v;
  }
  return m;
}
/* d = c1*d2 + d3 with the CG_VECTOR float rounding (CG_VECTOR.cpp:96-114) */
static void axpy_copy(double *d1, double c1, const double *d2, const double *d3, int64_t n, int mode) {
  if (mode == ORC_MODE_REF) {
    float c = (float)c1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      float t = c * (float)d2[i];
      t = t + (float)d3[i];
      d1[i] = t;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) d1[i] = c1 * d2[i] + d3[i];
  }
}

static void precondition(orc_system *S, const double *r, double *z, int mode) {
  const lvl_t *L0 = &S->lv[0];
  if (S->use_mg) orc_vcycle(S, r, z, mode);
  else memcpy(z, r, (size_t)L0->nn * 3 * sizeof(double));
  /* KRYLOV_SYSTEM_BASE.cpp:24-31 (preconditioner does not commute) */
  project_dirichlet(L0, z);
  project_null(L0, z);
}

/* returns iterations (>=0) or -1 on singular; *conv set */
static int pcg(orc_system *S, double *x, double *r, double tol, int max_it, int mode, int *conv, int run_fixed) {
  const lvl_t *L0 = &S->lv[0];
  int64_t n = L0->nn * 3;
  double *s = S->s, *q = S->q, *z = S->z;
  memset(x, 0, (size_t)n * sizeof(double));
  memset(s, 0, (size_t)n * sizeof(double));
  memset(q, 0, (size_t)n * sizeof(double));
  project_dirichlet(L0, x);
  /* q = K x (x = 0) ; r -= q ; Project(r) */
  orc_apply(S, x, q);
  axpy_copy(r, -1.0, q, r, n, mode);
  project_dirichlet(L0, r);
  double rho_old = (double)FLT_MAX;
  *conv = 0;
  int it;
  for (it = 0;; ++it) {
    project_null(L0, r);
    double nrm = linf(r, n);
    if (!run_fixed && nrm <= tol) { *conv = 1; return it; }
    if (it == max_it) break;
    precondition(S, r, z, mode);
    double rho = dot(z, r, n);
    axpy_copy(s, rho / rho_old, s, z, n, mode);
    orc_apply(S, s, q);
    project_dirichlet(L0, q);
    double sq = dot(s, q, n);
    if (sq <= 0) return -1;
    double alpha = rho / sq;
    axpy_copy(x, alpha, s, x, n, mode);
    axpy_copy(r, -alpha, q, r, n, mode);
    rho_old = rho;
  }
  return it;
}

static void full_residual(orc_system *S, const double *u, double *r) {
  const lvl_t *L0 = &S->lv[0];
  residual0(S, u, S->fvec, r);
  (void)L0;
}

static void prepare_rhs(orc_system *S) {
  const lvl_t *L0 = &S->lv[0];
  int64_t n = L0->nn * 3;
  memset(S->fvec, 0, (size_t)n * sizeof(double));
  for (int64_t i = 0; i < S->p.nforce; ++i) {
    int32_t m = S->f_node[i];
    if (m < 0) continue;
    for (int v = 0; v < 3; ++v) S->fvec[3 * (int64_t)m + v] = S->p.force_value[3 * i + v];
  }
  project_dirichlet(L0, S->fvec);
  project_null(L0, S->fvec);
  memset(S->ubase, 0, (size_t)n * sizeof(double));
  for (int64_t i = 0; i < S->nd; ++i)
    if (S->d_node[i] >= 0 && S->d_axis[i] >= 0 && S->d_axis[i] < 3)
      S->ubase[3 * (int64_t)S->d_node[i] + S->d_axis[i]] = S->d_val[i];
}

int orc_solve(orc_system *S, const orc_solve_params *prm, double *u_out, double *r_out, orc_solve_info *info,
              char *err, int errlen) {
  const lvl_t *L0 = &S->lv[0];
  int64_t n = L0->nn * 3;
  double t0 = now_s();
  prepare_rhs(S);
  double *u = (double *)malloc((size_t)n * sizeof(double));
  memcpy(u, S->ubase, (size_t)n * sizeof(double));
  full_residual(S, u, S->rr);
  double r0 = linf(S->rr, n);
  double target = prm->tolerance * r0;
  int total = 0, passes = 0, conv = 0;
  double rn = r0;
  int maxp = prm->max_passes < 1 ? 1 : prm->max_passes;
  int rc = 0;
  if (r0 > 0) {
    for (int pass = 0; pass < maxp; ++pass) {
      double tolp = target;
      if (prm->inner_tolerance > 0) {
        double t = prm->inner_tolerance * rn;
        tolp = t > target ? t : target;
      }
      int c = 0;
      int it = pcg(S, S->x, S->rr, tolp, prm->max_iterations, prm->mode, &c, 0);
      if (it < 0) {
        snprintf(err, errlen, "CG detected singular (s.q <= 0) in pass %d", pass);
        rc = -1;
        break;
      }
      total += it;
      passes++;
      for (int64_t i = 0; i < n; ++i) u[i] += S->x[i];
      full_residual(S, u, S->rr);
      rn = linf(S->rr, n);
      if (rn <= target) { conv = 1; break; }
    }
  } else {
    conv = 1;
  }
  if (u_out) memcpy(u_out, u, (size_t)n * sizeof(double));
  if (r_out) memcpy(r_out, S->rr, (size_t)n * sizeof(double));
  if (info) {
    info->iterations = total;
    info->passes = passes;
    info->converged = conv;
    info->r0_linf = r0;
    info->r_linf = rn;
    info->setup_seconds = 0;
    info->solve_seconds = now_s() - t0;
  }
  free(u);
  if (rc == 0 && !conv) {
    snprintf(err, errlen, "not converged: residual %.3e of target %.3e", rn, target);
    return -2;
  }
  return rc;
}

int orc_time_pcg_iterations(orc_system *S, int iters, int mode, double *seconds) {
  const lvl_t *L0 = &S->lv[0];
  int64_t n = L0->nn * 3;
  prepare_rhs(S);
  full_residual(S, S->ubase, S->rr);
  (void)n;
  int conv = 0;
  double t0 = now_s();
  int it = pcg(S, S->x, S->rr, 0.0, iters, mode, &conv, 1);
  *seconds = now_s() - t0;
  return it;
}

/* ------------------------------------------------------------------------- */
/* Sensitivity: topo_opt/objective/minimal_compliance.cpp:74-146, hex_fem.cpp:52-83 */
/* ------------------------------------------------------------------------- */
static double cell_energy(const orc_system *S, int64_t c, const double *a, const double *b, const double *K0) {
  const lvl_t *L = &S->lv[0];
  const int32_t *nb = L->nb + c * 27;
  double ua[24], ub[24];
  for (int J = 0; J < 8; ++J) {
    int32_t m = nb[(((J >> 2) & 1) + 1) * 9 + (((J >> 1) & 1) + 1) * 3 + ((J & 1) + 1)];
    for (int w = 0; w < 3; ++w) {
      ua[3 * J + w] = m >= 0 ? a[3 * (int64_t)m + w] : 0.0;
      ub[3 * J + w] = m >= 0 ? b[3 * (int64_t)m + w] : 0.0;
    }
  }
  double sum = 0.0;
  for (int I = 0; I < 8; ++I)
    for (int v = 0; v < 3; ++v)
      for (int J = 0; J < 8; ++J)
        for (int w = 0; w < 3; ++w) sum += ua[3 * I + v] * K0[((I * 8 + J) * 3 + v) * 3 + w] * ub[3 * J + w];
  return sum;
}

double orc_sensitivity(const orc_system *S, const double *u, double *dc, double *se) {
  const lvl_t *L = &S->lv[0];
  double K0[576];
  for (int e = 0; e < 576; ++e) K0[e] = S->p.mu0 * S->kmu[e] + S->p.lambda0 * S->kla[e];
  double p = S->p.penalty, emin = S->p.min_stiffness, obj = 0.0;
#pragma omp parallel for reduction(+ : obj) schedule(static)
  for (int64_t c = 0; c < L->nn; ++c) {
    double d = S->rho[c];
    if (!S->has_rho[c] || !(d > 0)) {
      if (dc) dc[c] = 0.0;
      if (se) se[c] = 0.0;
      continue;
    }
    double s = cell_energy(S, c, u, u, K0);
    if (se) se[c] = s;
    if (dc) dc[c] = -(s * pow(d, p - 1.0) * p * (1.0 - emin));
    obj += s * (emin + pow(d, p) * (1.0 - emin));
  }
  return obj;
}

void orc_adjoint_sensitivity(const orc_system *S, const double *u, const double *lam, double *g) {
  const lvl_t *L = &S->lv[0];
  double K0[576];
  for (int e = 0; e < 576; ++e) K0[e] = S->p.mu0 * S->kmu[e] + S->p.lambda0 * S->kla[e];
  double p = S->p.penalty, emin = S->p.min_stiffness;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < L->nn; ++c) {
    double d = S->rho[c];
    if (!S->has_rho[c] || !(d > 0)) { g[c] = 0.0; continue; }
    g[c] = -(cell_energy(S, c, lam, u, K0) * pow(d, p - 1.0) * p * (1.0 - emin));
  }
}

void orc_pack_displacements(const orc_system *S, const double *u, const double *r, double *out4) {
  const lvl_t *L0 = &S->lv[0];
  const orc_problem *p = &S->p;
  for (int64_t bk = 0; bk < p->nblocks; ++bk)
    for (int c = 0; c < 64; ++c) {
      int x = p->block_base[3 * bk] + (c >> 4), y = p->block_base[3 * bk + 1] + ((c >> 2) & 3),
          z = p->block_base[3 * bk + 2] + (c & 3);
      int32_t n = node_at(L0, x, y, z);
      double *o = out4 + (bk * 64 + c) * 4;
      if (n >= 0) {
        double m = 0.0;
        for (int v = 0; v < 3; ++v) {
          o[v] = u[3 * (int64_t)n + v];
          double a = r ? fabs(r[3 * (int64_t)n + v]) : 0.0;
          if (a > m) m = a;
        }
        o[3] = m;
      } else {
        /* outside allocated pages: only the Dirichlet value written into u */
        o[0] = o[1] = o[2] = o[3] = 0.0;
        for (int64_t i = 0; i < S->nd; ++i)
          if (S->d_xyz[4 * i + 1] == x && S->d_xyz[4 * i + 2] == y && S->d_xyz[4 * i + 3] == z && S->d_axis[i] >= 0 &&
              S->d_axis[i] < 3)
            o[S->d_axis[i]] = S->d_val[i];
      }
    }
}
