"""GPU parity: every HIP kernel and the full solves against the C oracle.

Vectors are exchanged by node coordinates: the product reports its level-node
order (spg_level_nodes), the oracle its allocated nodes; values are mapped
through the coordinates.  FP64 kernels must agree to ~1e-12 relative, FP32
kernels to ~1e-5; solves are compared through displacement, residual and
compliance errors (tolerances stated per test).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from spgrid_topo_opt_amd import _lib
from spgrid_topo_opt_amd import fem_interface as FI
from spgrid_topo_opt_amd import workloads as W

pytestmark = pytest.mark.gpu


def _key(xyz):
    xyz = np.asarray(xyz, np.int64) + 4096
    return (xyz[:, 0] << 42) | (xyz[:, 1] << 21) | xyz[:, 2]


class Pair:
    """A problem built on both sides plus coordinate maps per level."""

    def __init__(self, prob):
        print(f"[pair] {prob.name} levels={prob.mg_levels}", flush=True)
        self.p = prob
        self.o = O.Oracle(prob)
        self.g = _lib.Solver(prob)
        self.map = []
        for lv in range(self.o.levels):
            gx, gf = self.g.level_nodes(lv)
            ox = self.o.xyz(lv)
            ok = _key(ox)
            order = np.argsort(ok)
            pos = np.searchsorted(ok[order], _key(gx))
            assert (pos < len(ok)).all()
            idx = order[pos]
            assert (ok[idx] == _key(gx)).all(), "product node missing from the oracle's allocated set"
            self.map.append((idx, gf))

    def to_o(self, lv, v):
        out = np.zeros((self.o.nnodes(lv), 3))
        out[self.map[lv][0]] = v
        return out

    def from_o(self, lv, v):
        return np.asarray(v)[self.map[lv][0]]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def blocky():
    return Pair(W.random_blocky(seed=11, n=(16, 12, 12), mg_levels=4))


@pytest.fixture(scope="module")
def box3():
    return Pair(W.box(16, 8, 8, seed=1, mg_levels=3))


def test_flags_match(blocky):
    P = blocky
    for lv in range(P.o.levels):
        idx, gf = P.map[lv]
        of = P.o.flags(lv)[idx]
        want = ((of & 0x10) != 0) * 1 | ((of >> 5) & 7) << 1 | ((of & 0x1000) != 0) * 16
        np.testing.assert_array_equal(gf, want.astype(np.uint8))
        # every active oracle node is present on the GPU
        assert ((P.o.flags(lv) & 0x10) != 0).sum() == ((gf & 1) != 0).sum()


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 2e-6)])
def test_apply_and_residual(blocky, prec, tol):
    P = blocky
    rng = np.random.default_rng(0)
    n = len(P.map[0][0])
    u = rng.standard_normal((n, 3))
    b = rng.standard_normal((n, 3))
    got = P.g.kernel(_lib.OP_APPLY, 0, u, precision=prec)
    want = P.from_o(0, P.o.apply(P.to_o(0, u)))
    assert rel(got, want) < tol
    got = P.g.kernel(_lib.OP_RESIDUAL, 0, u, b, precision=prec)
    want_full = P.to_o(0, b) - P.o.apply(P.to_o(0, u))
    dm = (P.o.flags(0)[:, None] >> (5 + np.arange(3))) & 1
    want_full[dm.astype(bool)] = 0
    assert rel(got, P.from_o(0, want_full)) < tol * 10


def _fp32_operator_forms(P, u, monkeypatch):
    """relative error of every FP32 form of the Krylov operator P K u against
    the FP64 operator on the same (FP32-rounded) input"""
    want = P.g.kernel(_lib.OP_OPERATOR, 0, u, precision=1)
    errs = {"cell (default)": rel(P.g.kernel(_lib.OP_OPERATOR, 0, u, precision=0), want)}
    monkeypatch.setenv("SPG_APPLY_CELL", "0")
    errs["rows, fp64 sums"] = rel(P.g.kernel(_lib.OP_OPERATOR, 0, u, precision=0), want)
    monkeypatch.setenv("SPG_DIFF_SUMS", "1")
    errs["rows, difference form"] = rel(P.g.kernel(_lib.OP_OPERATOR, 0, u, precision=0), want)
    monkeypatch.setenv("SPG_DIFF_SUMS", "0")
    monkeypatch.setenv("SPG_ACC64_MODES", "0")
    errs["rows, fp32 sums"] = rel(P.g.kernel(_lib.OP_OPERATOR, 0, u, precision=0), want)
    for k in ("SPG_APPLY_CELL", "SPG_DIFF_SUMS", "SPG_ACC64_MODES"):
        monkeypatch.delenv(k)
    return errs


def test_fp32_operator_difference_form_on_smooth_fields(blocky, monkeypatch):
    """the FP32 Krylov operator P K u on a smooth field, where direct FP32 row
    sums cancel large terms: the cures (the strain-form cell apply, the
    default; node rows with FP64 sums; the difference form K (u_m - u_n)
    with FP32 sums) stay far below the direct FP32 form's error against the
    FP64 apply"""
    P = blocky
    gx, gf = P.g.level_nodes(0)
    x = gx.astype(np.float64)
    x -= x.mean(0)
    # a translation (K annihilates it) plus a small long-wave part: the
    # direct sums cancel terms of the translation's size
    u = 1.0 + 1e-3 * np.sin(x / 7.0 + np.array([0.0, 1.0, 2.0]))
    u = u.astype(np.float32).astype(np.float64)  # the FP64 reference sees the same FP32 input
    errs = _fp32_operator_forms(P, u, monkeypatch)
    print("smooth field:", {k: f"{v:.2e}" for k, v in errs.items()})
    direct = errs["rows, fp32 sums"]
    for k in ("cell (default)", "rows, fp64 sums", "rows, difference form"):
        assert errs[k] < 1e-4 and errs[k] < 0.05 * direct, k
    assert errs["rows, fp64 sums"] < 1e-5


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 2e-6)])
def test_cell_apply_matches_oracle(blocky, prec, tol, monkeypatch):
    """the strain-form cell apply (SPG_APPLY_CELL=1): K u, P (b - K u) and the
    Krylov operator P K u against the oracle's apply"""
    P = blocky
    monkeypatch.setenv("SPG_APPLY_CELL", "1")
    rng = np.random.default_rng(1)
    n = len(P.map[0][0])
    u = rng.standard_normal((n, 3))
    b = rng.standard_normal((n, 3))
    Ku = P.o.apply(P.to_o(0, u))
    dm = ((P.o.flags(0)[:, None] >> (5 + np.arange(3))) & 1).astype(bool)
    got = P.g.kernel(_lib.OP_APPLY, 0, u, precision=prec)
    assert rel(got, P.from_o(0, Ku)) < tol
    want = P.to_o(0, b) - Ku
    want[dm] = 0
    got = P.g.kernel(_lib.OP_RESIDUAL, 0, u, b, precision=prec)
    assert rel(got, P.from_o(0, want)) < tol * 10
    want = Ku.copy()
    want[dm] = 0
    got = P.g.kernel(_lib.OP_OPERATOR, 0, u, precision=prec)
    assert rel(got, P.from_o(0, want)) < tol


def test_fp32_cell_apply_on_rotating_fields(blocky, monkeypatch):
    """a rigid rotation plus a translation and a small long-wave part, all
    rounded to FP32: the strain-form cell apply with FP32 sums (rotations and
    translations cancel within each cell) against the FP64 apply of the same
    input, beside the node-wise forms (FP64 row sums; difference form;
    direct FP32 sums).  The cell form's remaining error is the FP32 rounding
    of each cell's stresses, which cancel between neighbouring cells (about
    2e-5 here, measured also in a NumPy FP32 emulation of the kernel)."""
    P = blocky
    gx, gf = P.g.level_nodes(0)
    x = gx.astype(np.float64)
    x -= x.mean(0)
    om = np.array([0.3, -0.7, 0.5])
    u = np.cross(om[None, :], x) + np.array([5.0, -3.0, 2.0]) + 1e-3 * np.sin(x / 7.0 + np.array([0.0, 1.0, 2.0]))
    u = u.astype(np.float32).astype(np.float64)
    errs = _fp32_operator_forms(P, u, monkeypatch)
    print("rotating field:", {k: f"{v:.2e}" for k, v in errs.items()})
    cell = errs["cell (default)"]
    assert cell < 1e-4 and errs["rows, fp64 sums"] < 1e-6
    assert cell < 0.5 * errs["rows, difference form"] and cell < 0.05 * errs["rows, fp32 sums"]


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 5e-5)])
@pytest.mark.parametrize("forward,boundary", [(True, False), (False, False), (True, True), (False, True)])
def test_gs_sweep_level0(blocky, prec, tol, forward, boundary):
    P = blocky
    rng = np.random.default_rng(5)
    n = len(P.map[0][0])
    gf = P.map[0][1]
    u = rng.standard_normal((n, 3)) * 1e-6
    u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    b = rng.standard_normal((n, 3))
    op = _lib.OP_GS_BOUNDARY if boundary else _lib.OP_GS_SWEEP
    got = P.g.kernel(op, 0, u, b, precision=prec, forward=forward)
    want = P.from_o(0, P.o.gs_sweep(0, P.to_o(0, u), P.to_o(0, b), forward, boundary))
    assert rel(got, want) < tol


@pytest.mark.parametrize("kind", ["1", "2", "3", "3:4", "3:16"])
@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 5e-5)])
def test_gs_colour_kernels(blocky, prec, tol, kind, monkeypatch):
    """every level-0 colour kernel (SPG_GS_KERNEL=1: eight lanes per node;
    2: a cell corner per wavefront; 3: four lanes per node, with
    SPG_GS_QB blocks per workgroup)"""
    P = blocky
    kind, _, qb = kind.partition(":")
    monkeypatch.setenv("SPG_GS_KERNEL", kind)
    monkeypatch.setenv("SPG_GS_QB", qb or "8")
    rng = np.random.default_rng(6)
    n = len(P.map[0][0])
    gf = P.map[0][1]
    u = rng.standard_normal((n, 3)) * 1e-6
    u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    b = rng.standard_normal((n, 3))
    for fwd in (True, False):
        got = P.g.kernel(_lib.OP_GS_SWEEP, 0, u, b, precision=prec, forward=fwd)
        want = P.from_o(0, P.o.gs_sweep(0, P.to_o(0, u), P.to_o(0, b), fwd, False))
        assert rel(got, want) < tol


@pytest.mark.parametrize("prec,tol,vtol", [(1, 1e-11, 1e-10), (0, 5e-5, 1e-4)])
def test_gs_with_colour_major_b(prec, tol, vtol, monkeypatch):
    """SPG_GS_CMB=1: the level-0 colour sweeps read b and the flags from
    colour-major copies (by default only above the memory-side cache size).
    Sweeps in both directions and the V-cycle match the oracle, and solves
    give the node-major reads' iterations and bits."""
    monkeypatch.setenv("SPG_GS_CMB", "1")
    P = Pair(W.random_blocky(seed=11, n=(16, 12, 12), mg_levels=4))
    rng = np.random.default_rng(7)
    n = len(P.map[0][0])
    gf = P.map[0][1]
    u = rng.standard_normal((n, 3)) * 1e-6
    u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    b = rng.standard_normal((n, 3))
    for fwd in (True, False):
        got = P.g.kernel(_lib.OP_GS_SWEEP, 0, u, b, precision=prec, forward=fwd)
        want = P.from_o(0, P.o.gs_sweep(0, P.to_o(0, u), P.to_o(0, b), fwd, False))
        assert rel(got, want) < tol
    b[(gf & 1) == 0] = 0
    b[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    got = P.g.kernel(_lib.OP_VCYCLE, 0, None, b, precision=prec)
    want = P.from_o(0, P.o.vcycle(P.to_o(0, b), O.MODE_FP64))
    assert rel(got, want) < vtol
    monkeypatch.setenv("SPG_GS_CMB", "0")
    g0 = _lib.Solver(P.p)
    kw = dict(tolerance=1e-10, precision=1, max_iterations=500) if prec == 1 else dict(tolerance=1e-7)
    (o1, i1), (o0, i0) = P.g.solve(**kw) if prec == 1 else P.g.solve_mixed(**kw), \
        g0.solve(**kw) if prec == 1 else g0.solve_mixed(**kw)
    assert (i1["iterations"], i1["passes"]) == (i0["iterations"], i0["passes"])
    assert np.array_equal(o1, o0)
    g0.close()


@pytest.mark.parametrize("host", [True, False])
def test_gs_sweep_follows_a_density_update(host):
    """density updates (host and device) reach the level-0 colour sweeps and
    their boundary phase, whose inverse self blocks and boundary moduli are
    derived from the moduli: sweeps after the update match the oracle built
    on the new densities"""
    import dataclasses

    import torch

    p = W.random_blocky(seed=11, n=(16, 12, 12), mg_levels=4)
    P = Pair(p)
    rng = np.random.default_rng(8)
    d2 = np.where(p.density > 0, rng.uniform(0.05, 1.0, p.density.shape), 0.0)
    if host:
        P.g.update_density(d2)
    else:
        dd = torch.tensor(np.asarray(d2, np.float64).reshape(-1), device="cuda")
        P.g.update_density_device(dd.data_ptr())
        torch.cuda.synchronize()
    o2 = O.Oracle(dataclasses.replace(p, density=d2))
    n = len(P.map[0][0])
    gf = P.map[0][1]
    u = rng.standard_normal((n, 3)) * 1e-6
    u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    b = rng.standard_normal((n, 3))
    for prec, tol in ((1, 1e-11), (0, 5e-5)):
        for fwd in (True, False):
            for op, bnd in ((_lib.OP_GS_SWEEP, False), (_lib.OP_GS_BOUNDARY, True)):
                got = P.g.kernel(op, 0, u, b, precision=prec, forward=fwd)
                want = P.from_o(0, o2.gs_sweep(0, P.to_o(0, u), P.to_o(0, b), fwd, bnd))
                assert rel(got, want) < tol, (prec, fwd, bnd)
    P.g.close()


@pytest.mark.parametrize("path", ["1", "2", "3", "4", "5"])
def test_boundary_fallback_paths(blocky, path, monkeypatch):
    """every boundary smoother (1: one workgroup from global memory; 2:
    per-colour list launches on level 0; 3: coefficient rows in LDS on level
    0 too; 4: the matrix-free one-workgroup level-0 phase; 5: the matrix-free
    level-0 phase as one cooperative launch): same result as the oracle on
    every smoothed level"""
    P = blocky
    monkeypatch.setenv("SPG_BOUNDARY_PATH", path)
    rng = np.random.default_rng(11)
    for lv in [0] + list(range(2, P.o.levels)):
        n = len(P.map[lv][0])
        gf = P.map[lv][1]
        u = rng.standard_normal((n, 3)) * 1e-3
        u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
        b = rng.standard_normal((n, 3))
        for prec, tol in ((1, 1e-11), (0, 1e-4)):
            got = P.g.kernel(_lib.OP_GS_BOUNDARY, lv, u, b, precision=prec, forward=True)
            want = P.from_o(lv, P.o.gs_sweep(lv, P.to_o(lv, u), P.to_o(lv, b), True, True))
            assert rel(got, want) < tol, (lv, prec)


@pytest.mark.parametrize("path", ["0", "1", "5"])
def test_boundary_phase_multi_sweep(blocky, path, monkeypatch):
    """a whole boundary phase (as many sweeps as the V-cycle runs) on every
    smoothed level; on the explicit levels path 0 runs it as one dense map
    [G H] built from the phase on unit vectors, path 1 sweep by sweep"""
    P = blocky
    monkeypatch.setenv("SPG_BOUNDARY_PATH", path)
    k = P.p.boundary_iters
    rng = np.random.default_rng(12)
    for lv in [0] + list(range(2, P.o.levels)):
        n = len(P.map[lv][0])
        gf = P.map[lv][1]
        u = rng.standard_normal((n, 3)) * 1e-3
        u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
        b = rng.standard_normal((n, 3))
        for fwd in (True, False):
            want = P.to_o(lv, u)
            for _ in range(k):
                want = P.o.gs_sweep(lv, want, P.to_o(lv, b), fwd, True)
            want = P.from_o(lv, want)
            for prec, tol in ((1, 1e-11), (0, 1e-4)):
                got = P.g.kernel(_lib.OP_GS_BOUNDARY, lv, u, b, precision=prec, forward=fwd, sweeps=k)
                assert rel(got, want) < tol, (lv, fwd, prec)


@pytest.mark.parametrize("prec,tol", [(1, 1e-13), (0, 1e-6)])
def test_restriction_and_prolongation(blocky, prec, tol):
    P = blocky
    rng = np.random.default_rng(7)
    for lv in range(1, P.o.levels):
        nf, nc = len(P.map[lv - 1][0]), len(P.map[lv][0])
        r = rng.standard_normal((nf, 3))
        got = P.g.kernel(_lib.OP_RESTRICT, lv, r, precision=prec)
        want = P.from_o(lv, P.o.restrict(lv, P.to_o(lv - 1, r)))
        assert rel(got, want) < tol, lv
        uc = rng.standard_normal((nc, 3))
        uf = rng.standard_normal((nf, 3))
        got = P.g.kernel(_lib.OP_PROLONGATE, lv - 1, uc, uf, precision=prec)
        want = P.o.prolongate(lv - 1, P.to_o(lv - 1, uf), P.to_o(lv, uc))
        want[(P.o.flags(lv - 1) & 0x10) == 0] = 0  # null-space projection fused in the kernel
        assert rel(got, P.from_o(lv - 1, want)) < tol, lv


def test_galerkin_stencils_fp64(blocky):
    P = blocky
    for lv in range(2, P.o.levels):
        got = P.g.stencils(lv)
        want = P.o.stencils(lv)[P.map[lv][0]]
        act = (P.map[lv][1] & 1) != 0
        assert act.sum() > 0
        assert rel(got[act], want[act]) < 1e-12, lv
        assert np.abs(got[~act]).max() == 0


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 2e-5)])
def test_stencil_gs_and_residual(blocky, prec, tol):
    P = blocky
    rng = np.random.default_rng(9)
    for lv in range(2, P.o.levels):
        n = len(P.map[lv][0])
        gf = P.map[lv][1]
        u = rng.standard_normal((n, 3)) * 1e-3
        u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
        u[(gf & 1) == 0] = 0
        b = rng.standard_normal((n, 3))
        got = P.g.kernel(_lib.OP_STENCIL_RES, lv, u, b, precision=prec)
        want = P.from_o(lv, P.o.stencil_residual(lv, P.to_o(lv, u), P.to_o(lv, b)))
        assert rel(got, want) < tol
        for fwd in (True, False):
            for bnd in (False, True):
                op = _lib.OP_GS_BOUNDARY if bnd else _lib.OP_GS_SWEEP
                got = P.g.kernel(op, lv, u, b, precision=prec, forward=fwd)
                want = P.from_o(lv, P.o.gs_sweep(lv, P.to_o(lv, u), P.to_o(lv, b), fwd, bnd))
                assert rel(got, want) < tol * 10, (lv, fwd, bnd)


def test_bottom_solve_is_exact(blocky):
    P = blocky
    B = P.o.levels - 1
    rng = np.random.default_rng(3)
    n = len(P.map[B][0])
    gf = P.map[B][1]
    b = rng.standard_normal((n, 3))
    b[(gf & 1) == 0] = 0
    b[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    x = P.g.kernel(_lib.OP_BOTTOM, B, None, b, precision=1)
    res = P.o.stencil_residual(B, P.to_o(B, x), P.to_o(B, b))
    assert np.abs(res).max() < 1e-9 * np.abs(b).max()


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 2e-5)])
@pytest.mark.parametrize("geometry", ["blocky", "clamped_face"])
def test_overlapped_level0_smoothing_matches_sequential(geometry, prec, tol, monkeypatch):
    """the far part of the first level-0 sweep beside the boundary phase, the
    near-band part after it (SPG_OVERLAP=1) against the plain sequence
    (SPG_OVERLAP=0, the default): the same V-cycle up to the summation order of
    the near-band rows, and the same solve (measured: FP64 V-cycles agree to
    3e-17, FP32 to 6e-8)"""
    if geometry == "blocky":
        prob = W.random_blocky(seed=8, n=(24, 16, 16), mg_levels=3)
    else:
        prob = W.box(48, 24, 24, seed=2, mg_levels=3, load="edge")
    rng = np.random.default_rng(17)
    out = {}
    for ov in ("0", "1"):
        monkeypatch.setenv("SPG_OVERLAP", ov)
        g = _lib.Solver(prob)
        gx, gf = g.level_nodes(0)
        b = rng.standard_normal((len(gx), 3)) if ov == "0" else out["b"]
        b = b.copy()
        b[(gf & 1) == 0] = 0
        b[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
        z = g.kernel(_lib.OP_VCYCLE, 0, None, b, precision=prec)
        _, info = g.solve(tolerance=1e-8, max_iterations=500, precision=prec, max_passes=30)
        out["b"] = b
        out[ov] = (z, info["iterations"])
        g.close()
    print(f"[overlap] {geometry}: V-cycle difference {rel(out['1'][0], out['0'][0]):.2e}, "
          f"iterations {out['0'][1]} / {out['1'][1]}", flush=True)
    assert rel(out["1"][0], out["0"][0]) < tol
    # FP64: the same iterations; FP32: the near-band rows' other summation
    # order may move the rounding by one iteration
    assert abs(out["1"][1] - out["0"][1]) <= (0 if prec == 1 else 2)


@pytest.mark.parametrize("prec,tol", [(1, 1e-10), (0, 1e-4)])
def test_vcycle(blocky, prec, tol):
    P = blocky
    rng = np.random.default_rng(13)
    n = len(P.map[0][0])
    gf = P.map[0][1]
    b = rng.standard_normal((n, 3))
    b[(gf & 1) == 0] = 0
    b[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    got = P.g.kernel(_lib.OP_VCYCLE, 0, None, b, precision=prec)
    want = P.from_o(0, P.o.vcycle(P.to_o(0, b), O.MODE_FP64))
    assert rel(got, want) < tol


def _solution_error(P, out4, u_o):
    """relative L2 error of the displacement at the input block nodes"""
    want = P.o.pack(u_o)[..., :3]
    got = out4[..., :3]
    return np.linalg.norm(got - want) / np.linalg.norm(want)


@pytest.fixture(scope="module")
def wide_face():
    """a 64x32x32 box clamped on a whole face: about 4,400 level-0 boundary
    nodes, 9 groups of 64 nodes per colour"""
    return _lib.Solver(W.box(64, 32, 32, seed=3, mg_levels=3))


@pytest.mark.parametrize("launch", ["0", "1"])
@pytest.mark.parametrize("wg", ["", "1", "2", "4"])
def test_cooperative_boundary_phase_matches_list_path(wide_face, wg, launch, monkeypatch, capfd):
    """the single-launch level-0 phase (grid barriers between colour steps,
    SPG_COOP_WG workgroups, each taking several node groups when capped;
    ordinary launch or, SPG_COOP_LAUNCH=1, the cooperative launch API)
    against the per-colour list launches, whole phases in both colour
    orders; FP64 to 1e-12, FP32 to 1e-5 relative (only the summation order
    differs)"""
    s = wide_face
    monkeypatch.setenv("SPG_COOP_WG", wg)
    monkeypatch.setenv("SPG_COOP_LAUNCH", launch)
    gx, gf = s.level_nodes(0)
    n = len(gx)
    rng = np.random.default_rng(21)
    u = rng.standard_normal((n, 3)) * 1e-3
    u[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    b = rng.standard_normal((n, 3))
    for fwd in (True, False):
        for prec, tol in ((1, 1e-12), (0, 1e-5)):
            monkeypatch.setenv("SPG_BOUNDARY_PATH", "2")
            want = s.kernel(_lib.OP_GS_BOUNDARY, 0, u, b, precision=prec, forward=fwd, sweeps=3)
            monkeypatch.setenv("SPG_BOUNDARY_PATH", "5")
            got = s.kernel(_lib.OP_GS_BOUNDARY, 0, u, b, precision=prec, forward=fwd, sweeps=3)
            assert rel(got, want) < tol, (fwd, prec)
            assert np.abs(got - u).max() > 0
    assert "not launched" not in capfd.readouterr().err


def test_cooperative_barrier_time_out_is_reported(wide_face, monkeypatch):
    """a barrier wait cut to one wall-clock tick (set from the host) makes
    the workgroups leave the kernel early; the call reports the time-out as
    an error, once, and the next phase runs normally"""
    s = wide_face
    gx, gf = s.level_nodes(0)
    rng = np.random.default_rng(22)
    u = np.zeros((len(gx), 3))
    b = rng.standard_normal((len(gx), 3))
    monkeypatch.setenv("SPG_BOUNDARY_PATH", "5")
    monkeypatch.setenv("SPG_COOP_LIMIT_TICKS", "1")
    with pytest.raises(_lib.SpgError, match="timed out"):
        s.kernel(_lib.OP_GS_BOUNDARY, 0, u, b, precision=0, forward=True, sweeps=3)
    monkeypatch.delenv("SPG_COOP_LIMIT_TICKS")
    got = s.kernel(_lib.OP_GS_BOUNDARY, 0, u, b, precision=1, forward=True, sweeps=3)
    monkeypatch.setenv("SPG_BOUNDARY_PATH", "2")
    want = s.kernel(_lib.OP_GS_BOUNDARY, 0, u, b, precision=1, forward=True, sweeps=3)
    assert rel(got, want) < 1e-12


def test_solve_falls_back_to_list_launches_after_a_barrier_time_out(monkeypatch, capfd):
    """a barrier wait cut to one tick during a solve: the pass is repeated
    with the per-colour list launches (once per solver, with a message) and
    the solve still reaches the oracle's solution"""
    P = Pair(W.box(64, 32, 32, seed=3, mg_levels=3))
    u_o, _, _ = P.o.solve(tolerance=1e-12, max_iterations=500)
    monkeypatch.setenv("SPG_BOUNDARY_PATH", "5")
    monkeypatch.setenv("SPG_COOP_LIMIT_TICKS", "1")
    out, info = P.g.solve_mixed(tolerance=1e-8)
    assert info["converged"] == 1
    err = capfd.readouterr().err
    assert err.count("continuing with per-colour launches") == 1, err
    want = P.o.pack(u_o)[..., :3]
    assert np.linalg.norm(out[..., :3] - want) / np.linalg.norm(want) < 1e-6
    monkeypatch.delenv("SPG_COOP_LIMIT_TICKS")
    out2, info2 = P.g.solve_mixed(tolerance=1e-8)  # stays on the list launches
    assert info2["converged"] == 1 and "continuing" not in capfd.readouterr().err


@pytest.mark.parametrize("levels", [0, 1, 3, 4])
def test_fp64_solve_matches_oracle(levels):
    P = Pair(W.random_blocky(seed=21, n=(16, 12, 12), mg_levels=levels))
    tol = 1e-9 if levels == 0 else 1e-10
    u_o, r_o, io = P.o.solve(tolerance=tol, max_iterations=5000)
    assert io["rc"] == 0
    out, ig = P.g.solve(tolerance=tol, max_iterations=5000, precision=1)
    print("oracle", io["iterations"], "gpu", ig["iterations"], flush=True)
    assert ig["converged"] == 1
    assert abs(ig["iterations"] - io["iterations"]) <= max(2, io["iterations"] // 20)
    assert _solution_error(P, out, u_o) < 1e-7
    assert ig["r_linf"] <= tol * ig["r0_linf"]


@pytest.mark.parametrize("smooth,bnd", [(2, 5), (1, 0), (3, 1)])
def test_smoothing_counts_follow_the_oracle(smooth, bnd):
    """pre_and_post_smoothing_iter and boundary_smoothing other than the
    caller's 1 and 3 (the reference interface's default boundary count is
    5, fem_interface.h:298-300): V-cycles in both precisions and the FP64
    solve match the oracle with the same counts"""
    import dataclasses
    prob = dataclasses.replace(W.random_blocky(seed=23, n=(16, 12, 12), mg_levels=4), smooth_iters=smooth,
                               boundary_iters=bnd)
    P = Pair(prob)
    rng = np.random.default_rng(17)
    n = len(P.map[0][0])
    gf = P.map[0][1]
    b = rng.standard_normal((n, 3))
    b[(gf & 1) == 0] = 0
    b[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    want = P.from_o(0, P.o.vcycle(P.to_o(0, b), O.MODE_FP64))
    for prec, tol in ((1, 1e-10), (0, 1e-4)):
        got = P.g.kernel(_lib.OP_VCYCLE, 0, None, b, precision=prec)
        assert rel(got, want) < tol, (prec, rel(got, want))
    u_o, _, io = P.o.solve(tolerance=1e-10, max_iterations=500)
    out, ig = P.g.solve(tolerance=1e-10, max_iterations=500, precision=1)
    print("iterations: oracle", io["iterations"], "gpu", ig["iterations"], flush=True)
    assert ig["converged"] == 1
    assert abs(ig["iterations"] - io["iterations"]) <= max(2, io["iterations"] // 20)
    assert _solution_error(P, out, u_o) < 1e-7


def test_repeated_force_and_dirichlet_entries_keep_the_last():
    """a force node or a Dirichlet (node, axis) listed twice takes its last
    value, as the reference's in-order writes leave it
    (SPGrid_CPU_Solver.cpp:241-246, 289-296); the solve matches the oracle
    and the solve of the lists with the repeats removed, bit for bit"""
    import dataclasses
    p = W.box(16, 8, 8, seed=2, mg_levels=3)
    fc, fv = p.force_coord, p.force_value
    dc, da, dv = p.dir_coord, p.dir_axis, p.dir_value
    rep = dataclasses.replace(
        p, force_coord=np.concatenate([fc, fc[:2]]),
        force_value=np.concatenate([fv, fv[:2] * np.array([[2.0, -1.0, 0.5]])]),
        dir_coord=np.concatenate([dc, dc[:3]]), dir_axis=np.concatenate([da, da[:3]]),
        dir_value=np.concatenate([dv, np.array([1e-4, -2e-4, 3e-4])]))
    once = dataclasses.replace(
        p, force_value=np.concatenate([fv[:2] * np.array([[2.0, -1.0, 0.5]]), fv[2:]]),
        dir_value=np.concatenate([np.array([1e-4, -2e-4, 3e-4]), dv[3:]]))
    kw = dict(tolerance=1e-10, precision=1, max_iterations=500)
    P = Pair(rep)
    out, info = P.g.solve(**kw)
    out_once, info_once = _lib.Solver(once).solve(**kw)
    assert info["iterations"] == info_once["iterations"]
    assert np.array_equal(out, out_once)
    u_o, _, io = P.o.solve(tolerance=1e-12, max_iterations=500)
    assert io["rc"] == 0
    assert _solution_error(P, out, u_o) < 1e-8


def test_bottom_with_more_than_2_31_inverse_entries(monkeypatch):
    """a bottom level of more than 46341 unknowns (its dense inverse has more
    than 2^31 entries and takes the 64-bit factorization path) still solves
    exactly: FP64 MG-PCG iterations and solution agree with the oracle"""
    monkeypatch.setenv("SPG_BOTTOM", "dense")  # bottoms this large take the sparse solve by default
    P = Pair(W.ellipsoid_shell(256, 128, 128, thickness_cells=2, mg_levels=3))
    nv = P.g.info()["bottom_dofs"]
    print("bottom unknowns", nv, flush=True)
    assert nv * nv > 2 ** 31
    u_o, r_o, io = P.o.solve(tolerance=1e-8, max_iterations=500)
    assert io["rc"] == 0
    out, ig = P.g.solve(tolerance=1e-8, max_iterations=500, precision=1)
    print("oracle", io["iterations"], "gpu", ig["iterations"], flush=True)
    assert ig["converged"] == 1
    assert abs(ig["iterations"] - io["iterations"]) <= max(2, io["iterations"] // 20)
    assert _solution_error(P, out, u_o) < 1e-6


def test_mixed_solve_matches_fp64_oracle(box3):
    P = box3
    u_o, r_o, io = P.o.solve(tolerance=1e-12, max_iterations=500)
    out, ig = P.g.solve_mixed(tolerance=1e-6)
    assert ig["converged"] == 1 and ig["passes"] >= 1
    assert ig["r_linf"] <= 1e-6 * ig["r0_linf"]
    err = _solution_error(P, out, u_o)
    assert err < 1e-5, err
    # compliance f.u / dx against the oracle's (fem objective scale: f.u/dx = sum_e s_e E_e)
    c_o = P.o.sensitivity(u_o)[0]
    idx = {tuple(c): i for i, c in enumerate(P.o.xyz(0))}
    fu_o = sum(np.dot(v, u_o[idx[tuple(c)]]) for c, v in zip(P.p.force_coord, P.p.force_value))
    nodes = (P.p.block_base[:, None, :] + np.stack(np.meshgrid(range(4), range(4), range(4), indexing="ij"),
                                                   -1).reshape(64, 3)[None]).reshape(-1, 3)
    nidx = {tuple(c): i for i, c in enumerate(nodes)}
    fu_g = sum(np.dot(v, out.reshape(-1, 4)[nidx[tuple(c)], :3]) for c, v in zip(P.p.force_coord, P.p.force_value))
    assert abs(fu_g - fu_o) / abs(fu_o) < 1e-6
    assert abs(fu_o / P.p.dx - c_o) / c_o < 1e-9


@pytest.mark.parametrize("wide_max", ["0", "65536"])
def test_operator_dot_sum_forms_agree(box3, wide_max, monkeypatch):
    """the s.q partials of the PCG operator are summed by one workgroup up to
    SPG_SUM_WIDE_MAX blocks and in two stages beyond; both forms give the
    oracle's solution (FP64 solve to 1e-10, displacement error 1e-8)"""
    P = box3
    monkeypatch.setenv("SPG_SUM_WIDE_MAX", wide_max)
    u_o, _, _ = P.o.solve(tolerance=1e-12, max_iterations=500)
    out, ig = P.g.solve(tolerance=1e-10, precision=1, max_passes=1, max_iterations=500)
    assert ig["converged"] == 1
    assert _solution_error(P, out, u_o) < 1e-8


def test_cfg1_reference_pass_and_mixed():
    P = Pair(W.config("cfg1"))
    u_o, _, io = P.o.solve(tolerance=1e-10, max_iterations=500)
    # the reference's single pass at krylov tolerance 1e-4 (FP32 inner)
    out1, i1 = P.g.solve(tolerance=1e-4, max_iterations=500, precision=0)
    assert i1["converged"] == 1 and i1["passes"] == 1
    assert _solution_error(P, out1, u_o) < 1e-3
    out2, i2 = P.g.solve_mixed(tolerance=1e-6)
    assert i2["r_linf"] <= 1e-6 * i2["r0_linf"]
    assert _solution_error(P, out2, u_o) < 1e-5
    print("cfg1 iterations: oracle(1e-10)", io["iterations"], "ref-pass", i1["iterations"], "mixed", i2["iterations"],
          "passes", i2["passes"], flush=True)


def test_fem_solve_end_to_end(box3):
    P = box3
    u_o, r_o, io = P.o.solve(tolerance=1e-10, max_iterations=500)
    inp = FI.FEMInputs.from_problem(P.p, tolerance=1e-10, max_iterations=500,
                                    extra_ints={"defect_correction_iter": 20}, extra_doubles={"inner_tolerance": 1e-3})
    outp = FI.FEMInterface(inp).run()
    assert outp.success, outp.info
    assert (outp.displacement_base == P.p.block_base).all()
    want = P.o.pack(u_o)
    err = np.linalg.norm(outp.displacements[..., :3] - want[..., :3]) / np.linalg.norm(want[..., :3])
    assert err < 1e-8, err
    assert outp.residuals and outp.residuals[-1] <= 1e-10


def test_fem_solve_blocks_off_the_4_grid():
    """Input blocks whose x bases are 1 mod 4, plus a layer of empty blocks
    at the end: the nodes of that layer with local x = 3 fall outside every
    allocated page.  A Dirichlet value given at one of them appears only in
    its output entry (SPGrid_CPU_Solver.cpp:241-246).  The whole output
    matches the oracle's."""
    p = W.box(16, 8, 8, seed=3, mg_levels=3)
    sh = np.array([1, 0, 0], np.int32)
    extra = p.block_base[p.block_base[:, 0] == p.block_base[:, 0].max()] + np.array([4, 0, 0], np.int32)
    bb = (np.concatenate([p.block_base, extra]) + sh).astype(np.int32)
    node = extra[0] + sh + np.array([3, 1, 2], np.int32)
    q = W.Problem(**{**p.__dict__, "name": "box16x8x8 shifted", "block_base": bb,
                     "density": np.concatenate([p.density, np.zeros((len(extra), 64))]),
                     "dir_coord": np.concatenate([p.dir_coord + sh, node[None]]).astype(np.int32),
                     "dir_axis": np.concatenate([p.dir_axis, [1]]).astype(np.int32),
                     "dir_value": np.concatenate([p.dir_value, [0.25]]),
                     "force_coord": (p.force_coord + sh).astype(np.int32)})
    o = O.Oracle(q)
    u_o, _, _ = o.solve(tolerance=1e-10, max_iterations=500)
    want = o.pack(u_o)
    outp = FI.FEMInterface(FI.FEMInputs.from_problem(q, tolerance=1e-10, max_iterations=500,
                                                     extra_ints={"defect_correction_iter": 20},
                                                     extra_doubles={"inner_tolerance": 1e-3})).run()
    assert outp.success, outp.info
    assert (outp.displacement_base == bb).all()
    got = outp.displacements
    bi = int(np.flatnonzero((bb == node - np.array([3, 1, 2])).all(1))[0])
    l = 3 * 16 + 1 * 4 + 2
    assert got[bi, l].tolist() == [0.0, 0.25, 0.0, 0.0]
    assert want[bi, l].tolist() == [0.0, 0.25, 0.0, 0.0]
    err = np.linalg.norm(got[..., :3] - want[..., :3]) / np.linalg.norm(want[..., :3])
    assert err < 1e-8, err


def test_fem_solve_short_of_the_tolerance_still_succeeds(box3):
    """A solve stopped by its iteration cap before the tolerance reports
    success = 1, as the reference's does (SPGrid_CPU_Solver.cpp:326,368;
    its caller aborts on success = 0, objective/minimal_compliance.cpp:64-67),
    with error_code SPG_E_NOT_CONVERGED, the residual it reached, and the
    displacement of its last iterate.  krylov.max_iterations is ignored, as
    by the reference (500 iterations); the cap here is the
    pcg_max_iterations extra."""
    inp = FI.FEMInputs.from_problem(box3.p, tolerance=1e-12, max_iterations=2,
                                    extra_ints={"pcg_max_iterations": 2})
    outp = FI.FEMInterface(inp).run()
    assert outp.success, outp.info
    assert outp.error_code == _lib.SPG_E_NOT_CONVERGED
    assert "tolerance not reached" in outp.info
    assert outp.residuals and 1e-12 < outp.residuals[-1] < 1.0
    d = outp.displacements[..., :3]
    assert np.isfinite(d).all() and np.abs(d).max() > 0
    # krylov.max_iterations alone (the reference caller's default 50 here)
    # does not cap the solve
    outp = FI.FEMInterface(FI.FEMInputs.from_problem(box3.p, tolerance=1e-10, max_iterations=2,
                                                     extra_ints={"defect_correction_iter": 20},
                                                     extra_doubles={"inner_tolerance": 1e-3})).run()
    assert outp.success and outp.error_code == 0, outp.info
    assert outp.residuals[-1] <= 1e-10


def test_sensitivity_matches_oracle(box3):
    torch = pytest.importorskip("torch")
    P = box3
    out, ig = P.g.solve(tolerance=1e-10, max_iterations=500, precision=1)
    nin = P.p.block_base.shape[0] * 64
    dc = torch.zeros(nin, dtype=torch.float64, device="cuda")
    obj = P.g.sensitivity_device(dc.data_ptr())
    torch.cuda.synchronize()
    u_o, _, _ = P.o.solve(tolerance=1e-12, max_iterations=500)
    obj_o, dc_o, _ = P.o.sensitivity(u_o)
    assert abs(obj - obj_o) / obj_o < 1e-8
    # map oracle per-node cells to input-block order
    nodes = (P.p.block_base[:, None, :] + np.stack(np.meshgrid(range(4), range(4), range(4), indexing="ij"),
                                                   -1).reshape(64, 3)[None]).reshape(-1, 3)
    oi = {tuple(c): i for i, c in enumerate(P.o.xyz(0))}
    want = np.array([dc_o[oi[tuple(c)]] if tuple(c) in oi else 0.0 for c in nodes])
    got = dc.cpu().numpy()
    assert rel(got, want) < 1e-7


def test_errors_are_reported():
    # floating body (no Dirichlet cells): CG detects a singular system or does not converge
    p = W.box(8, 4, 4)
    q = W.Problem(**{**p.__dict__, "dir_coord": p.dir_coord[:0], "dir_axis": p.dir_axis[:0],
                     "dir_value": p.dir_value[:0], "mg_levels": 0})
    s = _lib.Solver(q)
    _, info = s.solve(tolerance=1e-12, max_iterations=50, precision=1, raise_on_error=False)
    assert info["rc"] in (_lib.SPG_E_SINGULAR, _lib.SPG_E_NOT_CONVERGED)
    with pytest.raises(_lib.SpgError):
        _lib.Solver(p.with_levels(2))


def test_print_residuals_and_static_info(box3, capfd):
    """krylov.print_residuals prints ||r||_inf once per PCG iteration check
    (CG/CONJUGATE_GRADIENT.cpp:34-35): iterations + 1 lines per pass;
    spg_get_info reports the same sizes as a solve"""
    g = box3.g
    st = g.info()
    _, info = g.solve(tolerance=1e-6, precision=1, max_iterations=300, want_output=False, print_residuals=2)
    err = capfd.readouterr().err
    lines = [float(x) for x in err.split() if x.strip()]
    assert len(lines) == info["iterations"] + info["passes"]
    assert lines[-1] <= lines[0]
    for k in ("levels", "active_nodes", "active_cells", "gpu_blocks", "bottom_dofs"):
        assert st[k] == info[k], k
    assert st["hbm_bytes"] > 0


def test_device_density_update_matches_host_update():
    """spg_update_density_device gives the solve of spg_update_density, and
    a NaN or a topology change raises without modifying the solver"""
    import torch
    p = W.box(16, 8, 8, seed=4, mg_levels=3)
    g = _lib.Solver(p)
    rng = np.random.default_rng(3)
    d = np.where(p.density > 0, rng.uniform(0.1, 1.0, p.density.shape), 0.0)
    g.update_density(d)
    out_h, ih = g.solve(tolerance=1e-10, precision=1, max_iterations=500)
    dd = torch.as_tensor(d.reshape(-1), device="cuda")
    g.update_density(p.density)  # something else in between
    torch.cuda.synchronize()
    g.update_density_device(dd.data_ptr())
    out_d, idv = g.solve(tolerance=1e-10, precision=1, max_iterations=500)
    # device and host pow may differ in the last bit of E: compare the
    # displacements (not the roundoff-level residual column) at 1e-10
    assert idv["iterations"] == ih["iterations"]
    assert rel(out_d[..., :3], out_h[..., :3]) < 1e-10
    for bad, code in ((np.nan, _lib.SpgError), (0.0, _lib.SpgError)):
        x = d.reshape(-1).copy()
        x[int(np.argmax(d.reshape(-1) > 0))] = bad
        xd = torch.as_tensor(x, device="cuda")
        torch.cuda.synchronize()
        with pytest.raises(code):
            g.update_density_device(xd.data_ptr())
    out_a, _ = g.solve(tolerance=1e-10, precision=1, max_iterations=500)
    assert rel(out_a[..., :3], out_d[..., :3]) == 0  # the failed updates changed nothing
    g.close()
