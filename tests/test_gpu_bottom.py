"""The nested-dissection Cholesky bottom solve (csrc/bottom_nd.hip) against
the oracle's exact bottom solve and the dense-inverse bottom.

The reference factorizes its bottom level with MKL PARDISO
(solver/MG/Elasticity_Matrix.cpp:15-269); any exact SPD solve agrees with it
to rounding, so the checks are: the bottom residual of the oracle's stencil
operator (1e-9 of the right-hand side), agreement with the dense bottom
(1e-10), and whole FP64 / mixed solves against the oracle.  SPG_BOTTOM=nd
selects the sparse bottom for bottoms the dense inverse also takes.
"""
import numpy as np
import pytest

from spgrid_topo_opt_amd import _lib
from spgrid_topo_opt_amd import workloads as W

from .test_gpu_parity import Pair, _solution_error, rel

pytestmark = pytest.mark.gpu


def _pair(prob, monkeypatch, kind):
    monkeypatch.setenv("SPG_BOTTOM", kind)
    P = Pair(prob)
    info = P.g.info()
    print(f"[bottom] {prob.name} levels={prob.mg_levels} {kind}: {info['bottom_dofs']} unknowns, kind "
          f"{info['bottom_kind']}, {info['bottom_bytes'] / 1e6:.1f} MB", flush=True)
    assert info["bottom_kind"] == (1 if kind == "nd" else 0)
    return P


@pytest.mark.parametrize("levels", [3, 4])
def test_nd_bottom_solve_is_exact_and_matches_dense(levels, monkeypatch):
    prob = W.random_blocky(seed=5, n=(24, 16, 16), mg_levels=levels)
    Pn = _pair(prob, monkeypatch, "nd")
    B = Pn.o.levels - 1
    rng = np.random.default_rng(11)
    n = len(Pn.map[B][0])
    gf = Pn.map[B][1]
    b = rng.standard_normal((n, 3))
    b[(gf & 1) == 0] = 0
    b[((gf[:, None] >> (1 + np.arange(3))) & 1).astype(bool)] = 0
    xn = Pn.g.kernel(_lib.OP_BOTTOM, B, None, b, precision=1)
    res = Pn.o.stencil_residual(B, Pn.to_o(B, xn), Pn.to_o(B, b))
    print(f"[bottom] nd residual {np.abs(res).max() / np.abs(b).max():.3e}", flush=True)
    assert np.abs(res).max() < 1e-9 * np.abs(b).max()
    # FP32 bottom: FP64 factors, rounded result
    x32 = Pn.g.kernel(_lib.OP_BOTTOM, B, None, b, precision=0)
    assert rel(x32, xn) < 1e-6
    Pn.g.close()
    Pd = _pair(prob, monkeypatch, "dense")
    xd = Pd.g.kernel(_lib.OP_BOTTOM, B, None, b, precision=1)
    assert rel(xn, xd) < 1e-10


def test_nd_bottom_solves_match_the_oracle(monkeypatch):
    P = _pair(W.random_blocky(seed=21, n=(16, 12, 12), mg_levels=3), monkeypatch, "nd")
    u_o, _, io = P.o.solve(tolerance=1e-10, max_iterations=500)
    out, ig = P.g.solve(tolerance=1e-10, max_iterations=500, precision=1)
    print("[bottom] oracle", io["iterations"], "gpu", ig["iterations"], flush=True)
    assert ig["converged"] == 1
    assert abs(ig["iterations"] - io["iterations"]) <= max(2, io["iterations"] // 20)
    assert _solution_error(P, out, u_o) < 1e-7
    out, ig = P.g.solve_mixed(tolerance=1e-6)
    assert ig["converged"] == 1 and ig["r_linf"] <= 1e-6 * ig["r0_linf"]
    assert _solution_error(P, out, u_o) < 1e-5


def test_nd_bottom_on_a_46k_unknown_bottom(monkeypatch):
    """the bottom of test_bottom_with_more_than_2_31_inverse_entries (more
    than 46,341 unknowns) through the sparse factorization: same FP64 MG-PCG
    iterations and solution as the oracle, and far fewer bytes than the
    dense inverse"""
    prob = W.ellipsoid_shell(256, 128, 128, thickness_cells=2, mg_levels=3)
    P = _pair(prob, monkeypatch, "nd")
    nd_bytes = P.g.info()["bottom_bytes"]
    nv = P.g.info()["bottom_dofs"]
    u_o, _, io = P.o.solve(tolerance=1e-8, max_iterations=500)
    out, ig = P.g.solve(tolerance=1e-8, max_iterations=500, precision=1)
    print("[bottom] oracle", io["iterations"], "gpu", ig["iterations"], flush=True)
    assert ig["converged"] == 1
    assert abs(ig["iterations"] - io["iterations"]) <= max(2, io["iterations"] // 20)
    assert _solution_error(P, out, u_o) < 1e-6
    # the dense inverse's FP64 and FP32 tiles hold about 1.5 n^2 * 8 bytes
    assert nd_bytes < 0.25 * 12.0 * nv * nv


def test_bottom_dense_max_picks_the_bottom(monkeypatch):
    """spg_config.bottom_dense_max: a bottom one unknown above the limit gets
    the nested-dissection Cholesky, one at the limit the dense inverse; both
    solve to the same FP64 solution"""
    monkeypatch.delenv("SPG_BOTTOM", raising=False)
    prob = W.random_blocky(seed=5, n=(24, 16, 16), mg_levels=3)
    s0 = _lib.Solver(prob)
    nv = s0.info()["bottom_dofs"]
    assert 0 < nv <= 16384 and s0.info()["bottom_kind"] == 0  # the default limit
    s0.close()
    outs = []
    for limit, kind in ((nv - 1, 1), (nv, 0)):
        s = _lib.Solver(prob, bottom_dense_max=limit)
        assert s.info()["bottom_kind"] == kind, (limit, nv)
        out, info = s.solve(tolerance=1e-10, max_iterations=500, precision=1)
        assert info["converged"] == 1
        outs.append(out)
        s.close()
    assert rel(outs[0], outs[1]) < 1e-8
