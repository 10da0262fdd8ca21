"""Narrow-band maintenance (reference topo_opt/spgrid_topo_opt.cpp:92-241,
1404-1477, 1060-1160): the torch implementation against hand-derived sets and
against the loop restatement in oracle/narrow_band.py."""
import numpy as np
import pytest
import torch

from oracle.narrow_band import BandOracle
from spgrid_topo_opt_amd import narrow_band as NB
from spgrid_topo_opt_amd import workloads as W

O = W.OFFSET


def _container(nx=16, ny=8, nz=8):
    p = W.box(nx, ny, nz, seed=0)
    inside = (p.density > 0).reshape(-1)
    return p, inside


def _cell(p, x, y, z):
    """flat cell index of SPGrid cell (x, y, z) in the container"""
    bb = p.block_base // 4
    hit = np.nonzero((bb == np.array([x >> 2, y >> 2, z >> 2])).all(1))[0]
    assert len(hit) == 1
    return hit[0] * 64 + ((x & 3) << 4) + ((y & 3) << 2) + (z & 3)


def _blocks(p, mask):
    return {tuple(b) for b in (p.block_base[mask.cpu().numpy()] // 4).tolist()}


def _band(p, inside, device="cpu", force=None, **kw):
    hf = np.zeros_like(inside) if force is None else force
    return NB.NarrowBand(p.block_base, inside, hf, device=device, **kw)


def test_single_cell_inside_a_block():
    p, inside = _container()
    d = np.zeros(inside.shape)
    x, y, z = O + 5, O + 5, O + 5  # local (1,1,1) of block (O/4+1,)*3
    d[_cell(p, x, y, z)] = 0.5
    b = _band(p, inside)
    b.update(torch.as_tensor(d))
    B = (x >> 2, y >> 2, z >> 2)
    assert _blocks(p, b.thin) == {B}
    # (x +- 1) stays inside the block on every axis
    assert _blocks(p, b.fat) == {B}
    want = {(B[0] + i, B[1] + j, B[2] + k) for i in (0, 1) for j in (0, 1) for k in (0, 1)}
    assert _blocks(p, b.fatter) == want & {tuple(v) for v in (p.block_base // 4).tolist()}


def test_corner_cell_reaches_seven_neighbour_blocks():
    p, inside = _container()
    d = np.zeros(inside.shape)
    x, y, z = O + 4, O + 4, O + 4  # local (0,0,0)
    d[_cell(p, x, y, z)] = 0.5
    b = _band(p, inside)
    b.update(torch.as_tensor(d))
    B = (x >> 2, y >> 2, z >> 2)
    # offsets (+-1)^3 land in blocks B - {0,1}^3
    want = {(B[0] - i, B[1] - j, B[2] - k) for i in (0, 1) for j in (0, 1) for k in (0, 1)}
    cont = {tuple(v) for v in (p.block_base // 4).tolist()}
    assert _blocks(p, b.fat) == want & cont
    # margin 2 from local 1 reaches the lower block too
    d2 = np.zeros(inside.shape)
    d2[_cell(p, O + 5, O + 5, O + 5)] = 0.5
    b2 = _band(p, inside, fat_margin=2)
    b2.update(torch.as_tensor(d2))
    assert _blocks(p, b2.fat) == want & cont


def test_force_keeps_an_empty_block_and_threshold_uses_minimum_density():
    p, inside = _container()
    d = np.zeros(inside.shape)
    d[_cell(p, O + 5, O + 5, O + 5)] = 0.5
    hf = np.zeros_like(inside)
    far = _cell(p, O + 13, O + 1, O + 1)
    hf[far] = True
    b = _band(p, inside, force=hf)
    b.update_thin(torch.as_tensor(d))
    assert (O + 13) >> 2 in {t[0] for t in _blocks(p, b.thin)}
    # density 0.01 < minimum_density + threshold drops a block
    d[_cell(p, O + 9, O + 1, O + 1)] = 0.01
    b = _band(p, inside, minimum_density=0.05)
    b.update_thin(torch.as_tensor(d))
    assert _blocks(p, b.thin) == {((O + 5) >> 2,) * 3}


def test_connectivity_filter_keeps_the_large_component_and_raises_like_the_reference():
    p, inside = _container(32, 8, 8)
    d = np.zeros(inside.shape)
    # a row of 5 blocks along x and one separate block
    for bx in range(5):
        d[_cell(p, O + 4 * bx + 1, O + 1, O + 1)] = 0.5
    d[_cell(p, O + 4 * 7 + 1, O + 5, O + 5)] = 0.5
    b = _band(p, inside)
    b.update_thin(torch.as_tensor(d))
    b.filter_isolated()
    assert sorted(b.components) == [1, 5]
    assert len(_blocks(p, b.thin)) == 5
    # two components above 30%: the reference raises
    d = np.zeros(inside.shape)
    for bx in (0, 1):
        d[_cell(p, O + 4 * bx + 1, O + 1, O + 1)] = 0.5
    for bx in (5, 6):
        d[_cell(p, O + 4 * bx + 1, O + 1, O + 1)] = 0.5
    b = _band(p, inside)
    b.update_thin(torch.as_tensor(d))
    with pytest.raises(NB.BandError, match="more than two"):
        b.filter_isolated()
    # four single blocks: every component below 30%
    d = np.zeros(inside.shape)
    for bx in (0, 2, 4, 6):
        d[_cell(p, O + 4 * bx + 1, O + 1, O + 1)] = 0.5
    b = _band(p, inside)
    b.update_thin(torch.as_tensor(d))
    with pytest.raises(NB.BandError, match="fragmental"):
        b.filter_isolated()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matches_loop_restatement_over_several_updates(seed):
    p = W.random_blocky(seed=seed, n=(24, 16, 16))
    inside = (p.density > 0).reshape(-1)
    hf = NB.force_cells(p.block_base, inside, p.force_coord)
    rng = np.random.default_rng(seed)
    b = NB.NarrowBand(p.block_base, inside, hf, device="cpu", connectivity_filtering=False)
    o = BandOracle(p.block_base, inside, hf)
    d = np.where(inside, rng.uniform(0, 1, inside.shape), 0.0)
    for step in range(4):
        # zero out a growing region so blocks leave the band
        cut = p.block_base[:, 0].min() + 4 * (step + 1)
        blk = np.repeat(p.block_base[:, 0] < cut, 64)
        d = np.where(blk & (rng.uniform(size=d.shape) < 0.97), 0.0, d)
        b.update(torch.as_tensor(d))
        o.update(d, connectivity=False)
        for name in ("thin", "fat", "fatter"):
            np.testing.assert_array_equal(getattr(b, name).numpy(), o.mask(getattr(o, name)), err_msg=name)
        blocks, want = o.fem_density(d)
        np.testing.assert_array_equal(b.fem_blocks().numpy(), blocks)
        np.testing.assert_array_equal(b.fem_density(torch.as_tensor(d)).numpy(), want)


def test_connectivity_matches_restatement_on_random_islands():
    p, inside = _container(32, 16, 16)
    rng = np.random.default_rng(4)
    for trial in range(6):
        d = np.zeros(inside.shape)
        # a large slab plus a few random blocks
        slab = np.repeat((p.block_base[:, 1] == O) & (p.block_base[:, 2] == O), 64) & inside
        d[slab] = 0.5
        for c in rng.choice(np.nonzero(inside)[0], 3, replace=False):
            d[c] = 0.7
        b = _band(p, inside)
        o = BandOracle(p.block_base, inside, np.zeros_like(inside))
        try:
            o.update(d)
        except RuntimeError:
            with pytest.raises(NB.BandError):
                b.update(torch.as_tensor(d))
            continue
        b.update(torch.as_tensor(d))
        assert sorted(b.components) == sorted(o.components)
        for name in ("thin", "fat", "fatter"):
            np.testing.assert_array_equal(getattr(b, name).numpy(), o.mask(getattr(o, name)), err_msg=name)


def test_force_cells_are_the_eight_cells_around_each_force_node():
    p, inside = _container()
    hf = NB.force_cells(p.block_base, inside, p.force_coord)
    # edge load at x = nx face: the cells x = nx-1 (inside) around each node
    xs = set()
    for c in np.nonzero(hf)[0]:
        blk, l = divmod(int(c), 64)
        xs.add(int(p.block_base[blk, 0]) + (l >> 4))
    assert xs == {O + 15}
    # nodes (nx, 0, k), k = 0..nz: the inside cells (nx-1, 0, k-1..k) cover z = 0..nz-1 once each
    assert hf.sum() == 8
