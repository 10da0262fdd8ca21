"""Pins the C oracle against checks that do not reuse its formulas.

Each oracle function is checked against one of: the reference's own numbers
(stiffness tables, when the reference sources are mounted), an independent
numpy/scipy model (tests/fem_scipy.py: Gauss-quadrature element matrix, global
sparse assembly, closed-form trilinear prolongation, dense multicolour GS),
exact direct solves, invariants (rigid modes, patch test, symmetry), and
central finite differences for the sensitivity.
"""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from spgrid_topo_opt_amd import workloads as W
from tests import fem_scipy as F

REF = "/root/reference/solver/CPU_Kernels/Stiffness_Matrices.h"


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def test_element_tables_match_gauss_quadrature():
    kmu, kla = O.element_tables()
    gmu, gla = F.element_tables_gauss()
    np.testing.assert_allclose(kmu, gmu, atol=1e-14)
    np.testing.assert_allclose(kla, gla, atol=1e-14)


def test_element_tables_properties():
    kmu, kla = O.element_tables()
    for K in (kmu, kla):
        M = K.transpose(0, 2, 1, 3).reshape(24, 24)
        np.testing.assert_allclose(M, M.T, atol=1e-15)
        # entries are integers / 72 (Stiffness_Matrices.h writes k./72.)
        np.testing.assert_allclose(K * 72, np.round(K * 72), atol=1e-12)
        # rigid-body null space
        np.testing.assert_allclose(M @ F.rigid_modes(), 0, atol=1e-14)
    mu, lam = W.lame()
    K0 = (mu * kmu + lam * kla).transpose(0, 2, 1, 3).reshape(24, 24)
    ev = np.linalg.eigvalsh(K0)
    assert (np.abs(ev[:6]) < 1e-9 * ev[-1]).all() and (ev[6:] > 1e-6 * ev[-1]).all()
    # closed-form spot values: int |grad N0|^2 + int (dN0/dx)^2 = 1/3 + 1/9
    assert abs(kmu[0, 0, 0, 0] - 4.0 / 9.0) < 1e-15
    assert abs(kla[0, 0, 0, 1] - 1.0 / 12.0) < 1e-15


@pytest.mark.skipif(not os.path.exists(REF), reason="reference sources not mounted")
def test_element_tables_equal_reference_header():
    txt = open(REF).read()
    nums = [float(x) / 72.0 for x in re.findall(r"(-?\d+)\./72\.", txt)]
    assert len(nums) == 2 * 576
    kmu, kla = O.element_tables()
    np.testing.assert_array_equal(np.array(nums[:576]), kmu.reshape(-1))
    np.testing.assert_array_equal(np.array(nums[576:]), kla.reshape(-1))


def _E(p, d):
    return (1 - p.min_stiffness) * np.maximum(p.min_density, d) ** p.penalty + p.min_stiffness


def _model(o, p):
    """scipy K on the oracle's allocated node set, moduli from the densities."""
    xyz = o.xyz(0)
    cc, dd = p.cell_coords()
    e = _E(p, dd)
    gmu, gla = F.element_tables_gauss()
    K = F.assemble_K(xyz, cc, e * p.mu, e * p.lam, p.dx, gmu, gla)
    return xyz, K


@pytest.fixture(scope="module")
def blocky():
    p = W.random_blocky(seed=3, n=(12, 8, 8), mg_levels=3)
    return p, O.Oracle(p)


def test_moduli_follow_simp(blocky):
    p, o = blocky
    mu, la = o.moduli()
    xyz = o.xyz(0)
    cc, dd = p.cell_coords()
    idx = F.node_index(xyz)
    e = _E(p, dd)
    for c, ee in zip(cc, e):
        i = idx[tuple(c)]
        assert mu[i] == pytest.approx(ee * p.mu, rel=1e-15)
        assert la[i] == pytest.approx(ee * p.lam, rel=1e-15)
    assert O.stiffness_mapping(3.0, 1e-9, 0.0, 0.5) == pytest.approx((1 - 1e-9) * 0.125 + 1e-9, rel=1e-15)
    assert O.stiffness_mapping(0.0, 1e-9, 0.0, 0.37) == 0.37  # p = 0: identity (fem_interface.h:349-351)


def test_apply_matches_scipy_assembly(blocky):
    p, o = blocky
    xyz, K = _model(o, p)
    rng = np.random.default_rng(0)
    u = rng.standard_normal((len(xyz), 3))
    f = o.apply(u)
    ref = (K @ u.reshape(-1)).reshape(-1, 3)
    np.testing.assert_allclose(f, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_apply_invariants():
    p = W.box(8, 4, 4, seed=0, lo=1.0, hi=1.0)
    o = O.Oracle(p)
    xyz = o.xyz(0).astype(float)
    # rigid motion -> zero force everywhere
    om = np.array([0.3, -0.2, 0.5])
    u = np.cross(om, xyz) + np.array([1.0, 2.0, -1.0])
    assert np.abs(o.apply(u)).max() < 1e-9 * p.mu * p.dx * np.abs(u).max()
    # patch test: linear field -> zero force at interior nodes of a homogeneous body
    A = np.array([[0.1, 0.2, -0.3], [0.05, -0.1, 0.2], [0.3, 0.1, 0.05]])
    f = o.apply(xyz @ A.T)
    lo = np.array([8, 8, 8])
    hi = lo + np.array([8, 4, 4])
    interior = ((xyz > lo) & (xyz < hi)).all(1)
    assert interior.sum() > 0
    assert np.abs(f[interior]).max() < 1e-9 * p.mu * p.dx
    # symmetry
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal((2, len(xyz), 3))
    assert abs(np.sum(a * o.apply(b)) - np.sum(b * o.apply(a))) < 1e-9 * abs(np.sum(a * o.apply(b)))


def _flags_model(p, xyz):
    """Independent node flags on level 0: ACTIVE, DIRICHLET_v, BOUNDARY (3 rings)."""
    idx = F.node_index(xyz)
    n = len(xyz)
    fl = np.zeros(n, np.uint32)
    cc, _ = p.cell_coords()
    for c in cc:
        for a in range(8):
            k = tuple(int(t) for t in c + F.NODES[a].astype(int))
            fl[idx[k]] |= 0x10
    dset = [set() for _ in range(3)]
    for c, a in zip(p.dir_coord, p.dir_axis):
        dset[a].add(tuple(int(t) for t in c))
    for v in range(3):
        for c in dset[v]:
            if all(tuple(int(t) for t in np.array(c) + F.NODES[a].astype(int)) in dset[v] for a in range(8)):
                for a in range(8):
                    k = tuple(int(t) for t in np.array(c) + F.NODES[a].astype(int))
                    if k in idx:
                        fl[idx[k]] |= 0x20 << v
    for _ in range(3):
        add = []
        for i in range(n):
            if fl[i] & 0xE0 or not fl[i] & 0x10:
                continue
            x = xyz[i]
            for d in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T:
                j = idx.get(tuple(int(t) for t in x + d), -1)
                if j >= 0 and fl[j] & (0xE0 | 0x1000):
                    add.append(i)
                    break
        for i in add:
            fl[i] |= 0x1000
    return fl


def test_level0_flags_match_model(blocky):
    p, o = blocky
    xyz = o.xyz(0)
    got = o.flags(0) & (0x10 | 0xE0 | 0x1000)
    want = _flags_model(p, xyz)
    np.testing.assert_array_equal(got, want)
    assert (got & 0x1000).sum() > 0 and (got & 0xE0).sum() > 0


def test_transfers_match_trilinear(blocky):
    p, o = blocky
    rng = np.random.default_rng(2)
    for lf in range(o.levels - 1):
        fx, cx = o.xyz(lf), o.xyz(lf + 1)
        P = F.prolongation(fx, cx)
        dc = F.dirichlet_mask(o.flags(lf + 1))
        df = F.dirichlet_mask(o.flags(lf))
        r = rng.standard_normal((len(fx), 3))
        got = o.restrict(lf + 1, r).reshape(-1)
        want = (P.T @ r.reshape(-1)) * (~dc)
        np.testing.assert_allclose(got, want, atol=1e-13)
        uc = rng.standard_normal((len(cx), 3))
        u0 = rng.standard_normal((len(fx), 3))
        got = o.prolongate(lf, u0, uc).reshape(-1)
        want = (u0.reshape(-1) + P @ uc.reshape(-1)) * (~df)
        np.testing.assert_allclose(got, want, atol=1e-13)


@pytest.mark.parametrize("levels", [3, 4])
def test_galerkin_stencils_match_PtKP(levels):
    p = W.random_blocky(seed=5, n=(16, 12, 12), mg_levels=levels)
    o = O.Oracle(p)
    xyz, K = _model(o, p)
    P = sp.identity(K.shape[0], format="csr")
    for lv in range(1, o.levels):
        P = P @ F.prolongation(o.xyz(lv - 1), o.xyz(lv))
        if lv < 2:
            continue
        A = (P.T @ K @ P).tocsr()
        flags = o.flags(lv)
        A = F.delete_dirichlet(A, F.dirichlet_mask(flags))
        st = o.stencils(lv)
        cx = o.xyz(lv)
        cidx = F.node_index(cx)
        active = np.nonzero(flags & 0x10)[0]
        assert len(active) > 0
        scale = np.abs(A).max()
        for i in active:
            for v in range(3):
                rv = np.zeros(3 * len(cx))
                for s in range(27):
                    d = np.array([s // 9, (s // 3) % 3, s % 3]) - 1
                    j = cidx.get(tuple(int(t) for t in cx[i] + d), -1)
                    if j >= 0:
                        rv[3 * j:3 * j + 3] = st[i].reshape(27, 3, 3)[s, v]
                want = A.getrow(3 * i + v).toarray().ravel()
                np.testing.assert_allclose(rv, want, atol=1e-12 * scale)


def _gs_dense(K, xyz, flags, u, b, forward, boundary):
    """Textbook multicolour point-block Gauss-Seidel (colours = coordinate parity)."""
    u = u.copy().reshape(-1)
    Kd = K.toarray()
    col = ((xyz[:, 0] & 1) << 2) | ((xyz[:, 1] & 1) << 1) | (xyz[:, 2] & 1)
    order = range(8) if forward else range(7, -1, -1)
    for c in order:
        for n in np.nonzero(col == c)[0]:
            fl = flags[n]
            sl = slice(3 * n, 3 * n + 3)
            upd = (fl & 0x10) and (not boundary or (fl & 0x1000))
            if upd:
                r = b.reshape(-1)[sl] - Kd[sl] @ u
                D = Kd[sl, sl].copy()
                dm = [(fl >> (5 + v)) & 1 for v in range(3)]
                for v in range(3):
                    if dm[v]:
                        D[v, :] = 0
                        D[:, v] = 0
                        D[v, v] = 1
                Di = np.linalg.inv(D)
                for v in range(3):
                    if dm[v]:
                        Di[v, :] = 0
                        Di[:, v] = 0
                u[sl] += Di @ r
            for v in range(3):
                if (fl >> (5 + v)) & 1:
                    u[3 * n + v] = 0
    return u.reshape(-1, 3)


@pytest.mark.parametrize("forward,boundary", [(True, False), (False, False), (True, True)])
def test_gs_sweep_matches_dense_multicolour(blocky, forward, boundary):
    p, o = blocky
    xyz, K = _model(o, p)
    flags = o.flags(0)
    rng = np.random.default_rng(4)
    u = rng.standard_normal((len(xyz), 3)) * 1e-3
    u[F.dirichlet_mask(flags).reshape(-1, 3)] = 0
    b = rng.standard_normal((len(xyz), 3))
    got = o.gs_sweep(0, u, b, forward, boundary)
    want = _gs_dense(K, xyz, flags, u, b, forward, boundary)
    np.testing.assert_allclose(got, want, atol=1e-10 * np.abs(want).max())


def _direct(o, p):
    xyz, K = _model(o, p)
    flags = o.flags(0)
    idx = F.node_index(xyz)
    f = np.zeros((len(xyz), 3))
    for c, v in zip(p.force_coord, p.force_value):
        if tuple(c) in idx:
            f[idx[tuple(c)]] = v
    free = ((flags & 0x10) != 0)[:, None] & ~F.dirichlet_mask(flags).reshape(-1, 3)
    fr = free.reshape(-1)
    u = np.zeros(3 * len(xyz))
    u[fr] = spla.spsolve(K[fr][:, fr].tocsc(), f.reshape(-1)[fr])
    return u.reshape(-1, 3), f


@pytest.mark.parametrize("levels", [0, 1, 3, 4])
def test_pcg_converges_to_direct_solve(levels):
    p = W.random_blocky(seed=7, n=(16, 12, 12), mg_levels=levels)
    o = O.Oracle(p)
    u_ref, _ = _direct(o, p)
    tol = 1e-10 if levels == 0 else 1e-12  # plain CG stagnates near 1e-11 in FP64
    u, r, info = o.solve(tolerance=tol, max_iterations=5000, max_passes=1)
    assert info["rc"] == 0, info
    err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    assert err < (1e-7 if levels == 0 else 1e-9), err
    if levels == 1:
        assert info["iterations"] == 1  # V-cycle is the exact bottom solve
    if levels >= 3:
        assert info["iterations"] < 120  # regression guard, not a pin


def test_reference_rounding_mode_and_defect_correction():
    p = W.random_blocky(seed=8, n=(16, 12, 12), mg_levels=3)
    o = O.Oracle(p)
    u_ref, _ = _direct(o, p)
    # one reference-style pass at 1e-4: float-rounded inner CG
    u1, _, i1 = o.solve(tolerance=1e-4, max_passes=1, mode=O.MODE_REF)
    assert i1["rc"] == 0
    assert i1["r_linf"] <= 1e-4 * i1["r0_linf"] * 1.0001
    # iterated defect correction reaches 1e-9 with float-rounded inner solves
    u2, _, i2 = o.solve(tolerance=1e-9, inner_tolerance=1e-3, max_passes=20, mode=O.MODE_REF)
    assert i2["rc"] == 0 and i2["passes"] > 1
    assert np.linalg.norm(u2 - u_ref) / np.linalg.norm(u_ref) < 1e-7


def test_compliance_identity_and_fd_sensitivity():
    p = W.box(8, 4, 4, seed=2, lo=0.2, hi=1.0, mg_levels=3, load="point")
    o = O.Oracle(p)
    u, _, info = o.solve(tolerance=1e-13, max_iterations=500)
    assert info["rc"] == 0
    obj, dc, se = o.sensitivity(u)
    # reference compliance = sum_e s_e E(rho_e) = f.u / dx
    xyz = o.xyz(0)
    idx = F.node_index(xyz)
    fu = sum(np.dot(v, u[idx[tuple(c)]]) for c, v in zip(p.force_coord, p.force_value))
    assert obj == pytest.approx(fu / p.dx, rel=1e-9)
    rng = np.random.default_rng(0)
    cc, dd = p.cell_coords()
    for k in rng.choice(len(cc), 6, replace=False):
        b, l = divmod(int(np.nonzero((p.density > 0).reshape(-1))[0][k]), 64)
        h = 1e-4 * p.density[b, l]
        vals = []
        for s in (+1, -1):
            q = W.Problem(**{**p.__dict__, "density": p.density.copy()})
            q.density[b, l] += s * h
            oq = O.Oracle(q)
            uq, _, iq = oq.solve(tolerance=1e-13, max_iterations=500)
            vals.append(oq.sensitivity(uq)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        i = idx[tuple(cc[k])]
        assert dc[i] == pytest.approx(fd, rel=2e-5)


def test_build_errors_follow_reference():
    p = W.box(8, 4, 4)
    # an active cell whose nodes are not all in allocated pages -> activation check
    q = W.Problem(**{**p.__dict__, "block_base": p.block_base.copy(), "density": p.density.copy()})
    keep = q.block_base[:, 0] < 16  # drop the node-only blocks of the x=max face
    q.block_base = q.block_base[keep]
    q.density = q.density[keep]
    with pytest.raises(RuntimeError, match="activation"):
        O.Oracle(q)
    # padding violation (block too close to the low boundary)
    r = W.Problem(**{**p.__dict__, "block_base": p.block_base - 8})
    with pytest.raises(RuntimeError, match="padding"):
        O.Oracle(r)
    with pytest.raises(RuntimeError, match="mg_level=2"):
        O.Oracle(p.with_levels(2))
