"""Pins tests/torch_reference.py (the checker of the filter and OC kernels)
against closed forms, on the CPU.

* Filter (topo_opt/spgrid_topo_opt.cpp:527-574): a constant field gives
  c rho / (rho(x) + 1e-3) at every cell whose window holds an inside cell;
  a unit impulse at x0 on a full box with rho = 1 gives
  max(0, r - |x0 - x|) / (W(x) * 1.001), W(x) the window weight sum.
* OC (topo_opt/optimizer/optimality_criterion.cpp:19-146): uniform data
  have the multiplier mid = g (rho0 / V)^2 and the new density V; fixed
  cells and smart_oc shift the target and the effective -dc in known ways.
"""
import itertools
import math

import numpy as np
import pytest
import torch

from spgrid_topo_opt_amd import workloads as W
from spgrid_topo_opt_amd.topo_opt import CellGrid
from tests import torch_reference as TR


def _grid(nx=12, ny=8, nz=8):
    p = W.box(nx, ny, nz, seed=0)
    return p, CellGrid(p, device="cpu")


def _cell_xyz(p, grid):
    loc = np.array([(i, j, k) for i in range(4) for j in range(4) for k in range(4)])
    return (p.block_base[:, None, :] + loc[None]).reshape(-1, 3)


def _window(r):
    h = int(math.floor(r))
    return {o: max(0.0, r - math.sqrt(sum(t * t for t in o))) for o in itertools.product(range(-h, h + 1), repeat=3)}


@pytest.mark.parametrize("r", [1.0, 1.5, 2.5])
def test_filter_constant_field(r):
    p, grid = _grid()
    ins = grid.inside
    rho = torch.where(ins, torch.full((grid.nb * 64,), 0.37, dtype=torch.float64), torch.zeros(grid.nb * 64,
                                                                                             dtype=torch.float64))
    dc = torch.full_like(rho, -2.5)
    out = TR.filter_sensitivity(grid, dc, rho, r).numpy()
    xyz = _cell_xyz(p, grid)
    inside = {tuple(x) for x, m in zip(xyz, ins.numpy()) if m}
    w = _window(r)
    for n, x in enumerate(xyz):
        has = any(wt > 0 and tuple(x + np.array(o)) in inside for o, wt in w.items())
        want = -2.5 * 0.37 / (rho[n].item() + 1e-3) if has else 0.0
        assert out[n] == pytest.approx(want, rel=1e-13, abs=1e-300)


@pytest.mark.parametrize("r", [1.5, 2.5])
def test_filter_unit_impulse(r):
    p, grid = _grid(16, 12, 12)
    ins = grid.inside
    rho = torch.where(ins, torch.ones(grid.nb * 64, dtype=torch.float64), torch.zeros(grid.nb * 64,
                                                                                     dtype=torch.float64))
    xyz = _cell_xyz(p, grid)
    idx = {tuple(x): n for n, x in enumerate(xyz)}
    x0 = np.array([W.OFFSET + 8, W.OFFSET + 6, W.OFFSET + 6])
    dc = torch.zeros_like(rho)
    dc[idx[tuple(x0)]] = 1.0
    out = TR.filter_sensitivity(grid, dc, rho, r).numpy()
    w = _window(r)
    wsum = sum(w.values())
    h = int(math.floor(r))
    for o in itertools.product(range(-h - 1, h + 2), repeat=3):
        x = x0 + np.array(o)
        want = max(0.0, r - math.sqrt(sum(t * t for t in o))) / (wsum * 1.001)
        assert out[idx[tuple(x)]] == pytest.approx(want, rel=1e-13, abs=1e-300)


def _oc_setup():
    p, grid = _grid()
    ins = grid.inside
    n = grid.nb * 64
    rho = torch.where(ins, torch.full((n,), 0.5, dtype=torch.float64), torch.zeros(n, dtype=torch.float64))
    return grid, ins, rho


def test_oc_uniform_multiplier():
    grid, ins, rho = _oc_setup()
    dc = torch.full_like(rho, -2.0)
    new, change, mid = TR.oc_update(rho, dc, ins, 0.4, 0.2)
    assert mid == pytest.approx(2.0 * (0.5 / 0.4) ** 2, rel=1e-6)
    np.testing.assert_allclose(new[ins].numpy(), 0.4, rtol=1e-6)
    assert (new[~ins] == 0).all()
    assert change == pytest.approx(0.1, rel=1e-5)


@pytest.mark.parametrize("exclude", [True, False])
def test_oc_fixed_cells(exclude):
    grid, ins, rho = _oc_setup()
    fixed = torch.zeros_like(ins)
    fixed[torch.nonzero(ins)[::7, 0]] = True
    N, Nf = ins.sum().item(), fixed.sum().item()
    V = 0.45
    dc = torch.full_like(rho, -2.0)
    new, _, mid = TR.oc_update(rho, dc, ins, V, 0.2, fixed=fixed, fixed_cell_density=1.0, exclude_fixed_cells=exclude)
    free = ins & ~fixed
    assert (new[fixed] == 1.0).all()
    # excluded: V of the free cells; included: the fixed cells' unit density counts
    want = V if exclude else (V * N - Nf) / (N - Nf)
    np.testing.assert_allclose(new[free].numpy(), want, rtol=1e-6)
    assert mid == pytest.approx(2.0 * (0.5 / want) ** 2, rel=1e-6)


@pytest.mark.parametrize("smart", [True, False])
def test_oc_smart_shift(smart):
    grid, ins, rho = _oc_setup()
    cells = torch.nonzero(ins)[:, 0]
    dc = torch.zeros_like(rho)
    a, b = cells[0::2], cells[1::2]
    dc[a] = 1.0   # -dc = -1: no stiffness gain
    dc[b] = -3.0  # -dc = 3
    V = 0.35
    new, _, mid = TR.oc_update(rho, dc, ins, V, 0.2, smart_oc=smart)
    # group a drops by the step limit; group b takes the rest of the volume
    np.testing.assert_allclose(new[a].numpy(), 0.3, rtol=1e-12)
    nb_ = (V * ins.sum().item() - 0.3 * len(a)) / len(b)
    np.testing.assert_allclose(new[b].numpy(), nb_, rtol=1e-6)
    g = 3.0 + 1.0 if smart else 3.0  # shift = min(0, min(-dc)) = -1 under smart_oc
    assert mid == pytest.approx(g * (0.5 / nb_) ** 2, rel=1e-6)
