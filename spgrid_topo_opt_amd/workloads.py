"""Synthetic narrow-band workloads in the reference's FEMInputs vocabulary.

A :class:`Problem` is what the caller of ``fem_solve`` packs into
``FEMInputs`` (topo_opt/spgrid_topo_opt.cpp:977-1200): 4x4x4 density blocks
(cell-only blocks plus density-0 "node-only" blocks so every node of an active
cell lies in an allocated block, spgrid_topo_opt.cpp:1084-1122), Dirichlet
nodes of whole clamped cells (spgrid_topo_opt.h:560-575,627-687), nodal forces
and solver knobs.

Densities are deterministic: rho = lo + (hi-lo)*U with U = SplitMix64(seed xor
linear cell index) / 2^64, the linear index being (i*ny + j)*nz + k over the
un-padded cell grid.  Domain cells (i,j,k) sit at SPGrid cell (i,j,k)+OFFSET.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

OFFSET = 8  # cell (0,0,0) of the domain lives at SPGrid cell (8,8,8)
E_DEFAULT = 1.0e6
NU_DEFAULT = 0.3


def lame(E=E_DEFAULT, nu=NU_DEFAULT):
    """Material(E, nu): topo_opt/hex_fem.h:19-22."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return mu, lam


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, idx):
    """U in [0,1) from SplitMix64(seed xor idx) (53 high bits)."""
    z = splitmix64(np.uint64(seed) ^ np.asarray(idx, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class Problem:
    name: str
    res: tuple
    dx: float
    mu: float
    lam: float
    block_base: np.ndarray      # (nb,3) int32, multiples of 4
    density: np.ndarray         # (nb,64) float64, index i*16+j*4+k
    dir_coord: np.ndarray       # (nd,3) int32
    dir_axis: np.ndarray        # (nd,) int32
    dir_value: np.ndarray       # (nd,) float64
    force_coord: np.ndarray     # (nf,3) int32
    force_value: np.ndarray     # (nf,3) float64
    penalty: float = 3.0
    min_stiffness: float = 1e-9
    min_density: float = 0.0
    mg_levels: int = 3
    smooth_iters: int = 1
    boundary_iters: int = 3
    domain: tuple = (0, 0, 0)
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def n_active_cells(self):
        return int((self.density > 0).sum())

    def cell_coords(self):
        """(N,3) SPGrid coords and (N,) density of every cell with density>0."""
        loc = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1).reshape(64, 3)
        nb = self.block_base.shape[0]
        coords = (self.block_base[:, None, :] + loc[None]).reshape(nb * 64, 3)
        d = self.density.reshape(-1)
        m = d > 0
        return coords[m].astype(np.int32), d[m]

    def active_nodes(self):
        """Sorted unique (M,3) coords of nodes touching an active cell."""
        cc, _ = self.cell_coords()
        offs = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], np.int32)
        return _unpack(np.unique(_pack((cc[:, None, :] + offs[None]).reshape(-1, 3)))).astype(np.int32)

    def with_levels(self, mg_levels):
        return dataclasses.replace(self, mg_levels=mg_levels)


_KB = np.int64(1 << 20)


def _pack(c):
    c = np.asarray(c, np.int64)
    return (c[..., 0] * _KB + c[..., 1]) * _KB + c[..., 2]


def _unpack(k):
    k = np.asarray(k, np.int64)
    return np.stack([k // (_KB * _KB), (k // _KB) % _KB, k % _KB], -1)


def _blocks_from_cells(cells, dens):
    """Group cells (SPGrid coords) into 4^3 blocks and add node-only blocks."""
    cells = np.asarray(cells, np.int64)
    bkey = cells >> 2
    ck = _pack(bkey)
    keys = [np.unique(ck)]
    top = (cells & 3) == 3  # a shift by +1 leaves the block only from local index 3
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                if a or b or c:
                    d = np.array([a, b, c])
                    sel = np.all(top | (d == 0), axis=1)
                    keys.append(np.unique(_pack((cells[sel] + d) >> 2)))
    keys = np.unique(np.concatenate(keys))  # sorted: lexicographic block order (x slowest)
    bid = np.searchsorted(keys, ck)
    loc = cells & 3
    li = loc[:, 0] * 16 + loc[:, 1] * 4 + loc[:, 2]
    density = np.zeros((keys.shape[0], 64), np.float64)
    density[bid, li] = dens
    return (_unpack(keys) * 4).astype(np.int32), density


def _clamp_cells(cells):
    """Dirichlet nodes (all 3 axes) of whole clamped cells, deduplicated."""
    cells = np.asarray(cells, np.int32).reshape(-1, 3)
    offs = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], np.int32)
    nodes = _unpack(np.unique(_pack((cells[:, None, :] + offs[None]).reshape(-1, 3)))) if len(cells) else \
        np.zeros((0, 3), np.int64)
    coord = np.repeat(nodes, 3, axis=0)
    axis = np.tile(np.arange(3, dtype=np.int32), nodes.shape[0])
    return coord.astype(np.int32), axis, np.zeros(coord.shape[0])


def _grid_density(nx, ny, nz, seed, lo, hi, ijk):
    lin = (ijk[:, 0].astype(np.uint64) * np.uint64(ny) + ijk[:, 1].astype(np.uint64)) * np.uint64(nz) + ijk[:, 2].astype(
        np.uint64)
    return lo + (hi - lo) * uniform(seed, lin)


def box(nx, ny, nz, seed=0, lo=0.05, hi=1.0, mg_levels=3, load="edge", name=None, E=E_DEFAULT, nu=NU_DEFAULT):
    """Dense cantilever box; clamp the x=0 cell layer; load on the x=max face.

    load="edge": total (0,-1,0) spread over the nodes (nx, 0, k), k=0..nz.
    load="point": (0,-1,0) at node (nx, 0, nz/2).
    """
    mu, lam = lame(E, nu)
    ijk = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    dens = _grid_density(nx, ny, nz, seed, lo, hi, ijk)
    cells = ijk + OFFSET
    bb, d = _blocks_from_cells(cells, dens)
    clamp = ijk[ijk[:, 0] == 0] + OFFSET
    dc, da, dv = _clamp_cells(clamp)
    if load == "edge":
        fc = np.array([[nx, 0, k] for k in range(nz + 1)], np.int32) + OFFSET
        fv = np.tile(np.array([0.0, -1.0 / (nz + 1), 0.0]), (nz + 1, 1))
    else:
        fc = np.array([[nx, 0, nz // 2]], np.int32) + OFFSET
        fv = np.array([[0.0, -1.0, 0.0]])
    res = (nx + 4 * OFFSET, ny + 4 * OFFSET, nz + 4 * OFFSET)
    return Problem(name or f"box{nx}x{ny}x{nz}", res, 1.0 / nx, mu, lam, bb, d, dc, da, dv, fc, fv,
                   mg_levels=mg_levels, domain=(nx, ny, nz), meta={"seed": seed, "lo": lo, "hi": hi})


def _box_filter(cells_ijk, dens, shape, passes):
    """Mean over the active cells of each 3x3x3 neighbourhood, `passes` times."""
    if passes <= 0:
        return dens
    nx, ny, nz = shape
    grid = np.zeros((nx + 2, ny + 2, nz + 2))
    mask = np.zeros((nx + 2, ny + 2, nz + 2))
    i, j, k = (cells_ijk[:, 0] + 1, cells_ijk[:, 1] + 1, cells_ijk[:, 2] + 1)
    mask[i, j, k] = 1.0
    cur = dens.copy()
    for _ in range(passes):
        grid[:] = 0
        grid[i, j, k] = cur
        s = np.zeros_like(grid)
        c = np.zeros_like(grid)
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for e in (-1, 0, 1):
                    s[1:-1, 1:-1, 1:-1] += grid[1 + a:nx + 1 + a, 1 + b:ny + 1 + b, 1 + e:nz + 1 + e]
                    c[1:-1, 1:-1, 1:-1] += mask[1 + a:nx + 1 + a, 1 + b:ny + 1 + b, 1 + e:nz + 1 + e]
        cur = s[i, j, k] / c[i, j, k]
    return cur


def sphere_shell(n=256, r_in=0.4, r_out=0.5, seed=0, lo=0.05, hi=1.0, filter_passes=2, mg_levels=5, name=None,
                 E=E_DEFAULT, nu=NU_DEFAULT, clamp_radius=0.03):
    """Cells of [-0.5,0.5]^3 (dx=1/n) whose centre satisfies r_in <= |x| <= r_out.

    Clamp: cells whose centre is within clamp_radius+dx of the south pole
    (0,-0.5,0); load (0,-1,0) at the active node nearest the north pole.
    """
    mu, lam = lame(E, nu)
    dx = 1.0 / n
    g = (np.arange(n) + 0.5) * dx - 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    m = (r >= r_in) & (r <= r_out)
    ijk = np.argwhere(m).astype(np.int64)
    del X, Y, Z, r, m
    dens = _grid_density(n, n, n, seed, lo, hi, ijk)
    dens = _box_filter(ijk, dens, (n, n, n), filter_passes)
    cells = ijk + OFFSET
    bb, d = _blocks_from_cells(cells, dens)
    cen = (ijk + 0.5) * dx - 0.5
    south = np.array([0.0, -0.5, 0.0])
    dist = np.linalg.norm(cen - south, axis=1)
    clamp = cells[dist <= clamp_radius + dx]
    dc, da, dv = _clamp_cells(clamp)
    # active node nearest the north pole: among nodes of the cells closest to it
    north = np.array([0.0, 0.5, 0.0])
    near = ijk[np.linalg.norm(cen - north, axis=1) <= 2 * dx + np.linalg.norm(cen - north, axis=1).min()]
    offs = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], np.int64)
    nodes = _unpack(np.unique(_pack((near[:, None, :] + offs[None]).reshape(-1, 3))))
    npos = nodes * dx - 0.5
    lnode = nodes[np.argmin(np.linalg.norm(npos - north, axis=1))] + OFFSET
    fc = lnode.reshape(1, 3).astype(np.int32)
    fv = np.array([[0.0, -1.0, 0.0]])
    res = (n + 4 * OFFSET,) * 3
    return Problem(name or f"shell{n}", res, dx, mu, lam, bb, d, dc, da, dv, fc, fv, mg_levels=mg_levels,
                   domain=(n, n, n), meta={"seed": seed, "load_node": fc[0].tolist()})


def band(nx=512, ny=256, nz=256, width=24, seed=0, lo=0.05, hi=1.0, mg_levels=6, name=None, E=E_DEFAULT,
         nu=NU_DEFAULT, clamp_radius=0.03):
    """Cells within `width` cells of the bounding-box surface (bridge-like band).

    Clamps: cell patches within clamp_radius of the two bottom corners
    (x=0 and x=max edges at y=0, z centre); uniform downward load spread over
    the top-face nodes (total (0,-1,0)).
    """
    mu, lam = lame(E, nu)
    dx = 1.0 / nx
    cells = []
    dens = []
    for i in range(nx):
        j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
        inb = (i < width) | (i >= nx - width) | (j < width) | (j >= ny - width) | (k < width) | (k >= nz - width)
        jj, kk = j[inb], k[inb]
        ijk = np.stack([np.full(jj.shape, i), jj, kk], -1).astype(np.int64)
        cells.append(ijk)
    ijk = np.concatenate(cells)
    dens = _grid_density(nx, ny, nz, seed, lo, hi, ijk)
    c = ijk + OFFSET
    bb, d = _blocks_from_cells(c, dens)
    cen = (ijk + 0.5) * dx
    corners = [np.array([0.0, 0.0, nz * dx / 2]), np.array([nx * dx, 0.0, nz * dx / 2])]
    sel = np.zeros(ijk.shape[0], bool)
    for cc in corners:
        sel |= np.linalg.norm(cen - cc, axis=1) <= clamp_radius
    dc, da, dv = _clamp_cells(c[sel])
    top = np.array([[i, ny, k] for i in range(nx + 1) for k in range(nz + 1)], np.int32) + OFFSET
    fv = np.tile(np.array([0.0, -1.0 / top.shape[0], 0.0]), (top.shape[0], 1))
    res = (nx + 4 * OFFSET, ny + 4 * OFFSET, nz + 4 * OFFSET)
    return Problem(name or f"band{nx}x{ny}x{nz}w{width}", res, dx, mu, lam, bb, d, dc, da, dv, top, fv,
                   mg_levels=mg_levels, domain=(nx, ny, nz), meta={"seed": seed})


def ellipsoid_shell(nx=2048, ny=1024, nz=1024, semi=(0.5, 0.25, 0.25), thickness_cells=16, seed=0, lo=0.05, hi=1.0,
                    mg_levels=6, name=None, E=E_DEFAULT, nu=NU_DEFAULT, progress=None):
    """Cells between the ellipsoid with semi-axes `semi` (domain units, centred,
    dx = 1/nx) and the one with every semi-axis reduced by thickness*dx.
    Clamp: cells within 0.03 of the -x pole; load (0,-1,0) at the +x pole node.
    `progress(msg)` is called every 256 x-slices and per build stage."""
    mu, lam = lame(E, nu)
    dx = 1.0 / nx
    a, b, c = semi
    t = thickness_cells * dx
    ai, bi, ci = a - t, b - t, c - t
    gy = (np.arange(ny) + 0.5) * dx - ny * dx / 2
    gz = (np.arange(nz) + 0.5) * dx - nz * dx / 2
    Y, Z = np.meshgrid(gy, gz, indexing="ij")
    parts = []
    for i in range(nx):
        x = (i + 0.5) * dx - nx * dx / 2
        outer = (x / a) ** 2 + (Y / b) ** 2 + (Z / c) ** 2 <= 1.0
        inner = (x / ai) ** 2 + (Y / bi) ** 2 + (Z / ci) ** 2 < 1.0 if ai > 0 else np.zeros_like(outer)
        m = outer & ~inner
        if m.any():
            jk = np.argwhere(m)
            parts.append(np.concatenate([np.full((jk.shape[0], 1), i), jk], 1))
        if progress and (i + 1) % 256 == 0:
            progress(f"ellipsoid shell: {i + 1}/{nx} x-slices")
    ijk = np.concatenate(parts).astype(np.int64)
    del parts
    dens = _grid_density(nx, ny, nz, seed, lo, hi, ijk)
    cc = ijk + OFFSET
    if progress:
        progress(f"ellipsoid shell: {ijk.shape[0]} cells, grouping into blocks")
    bb, d = _blocks_from_cells(cc, dens)
    if progress:
        progress(f"ellipsoid shell: {bb.shape[0]} blocks")
    cen = (ijk + 0.5) * dx - np.array([nx, ny, nz]) * dx / 2
    clamp = cc[np.linalg.norm(cen - np.array([-a, 0, 0]), axis=1) <= 0.03]
    dc, da, dv = _clamp_cells(clamp)
    fc = np.array([[nx // 2 + int(round(a / dx)), ny // 2, nz // 2]], np.int32)
    # nearest active node to the +x pole
    cand = cc[np.argmax(cen[:, 0])] + np.array([1, 0, 0])
    fc = cand.reshape(1, 3).astype(np.int32)
    fv = np.array([[0.0, -1.0, 0.0]])
    res = (nx + 4 * OFFSET, ny + 4 * OFFSET, nz + 4 * OFFSET)
    return Problem(name or f"ellipsoid{nx}", res, dx, mu, lam, bb, d, dc, da, dv, fc, fv, mg_levels=mg_levels,
                   domain=(nx, ny, nz), meta={"seed": seed})


def auto_mg_levels(container_max, bottom=64):
    """Caller rule: 1 + halvings until <= mg_bottom_size, 2 -> 3
    (topo_opt/spgrid_topo_opt.cpp:1022-1037)."""
    lv = 1
    n = container_max
    while n > bottom:
        n //= 2
        lv += 1
    return 3 if lv == 2 else lv


def random_blocky(seed=0, n=(16, 12, 12), fill=0.7, mg_levels=3, clamp_axis0=True):
    """Small irregular test geometry: random subset of 2x2x2 cell clusters."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    keep = rng.random((nx // 2, ny // 2, nz // 2)) < fill
    keep[0, :, :] = True
    m = np.repeat(np.repeat(np.repeat(keep, 2, 0), 2, 1), 2, 2)
    ijk = np.argwhere(m).astype(np.int64)
    dens = 0.05 + 0.95 * rng.random(ijk.shape[0])
    cells = ijk + OFFSET
    bb, d = _blocks_from_cells(cells, dens)
    dc, da, dv = _clamp_cells(cells[ijk[:, 0] == 0]) if clamp_axis0 else _clamp_cells(np.zeros((0, 3), np.int32))
    top = ijk[ijk[:, 1] == ijk[:, 1].max()]
    fc = (top[: max(1, top.shape[0] // 3)] + np.array([1, 1, 0]) + OFFSET).astype(np.int32)
    fc = np.unique(fc, axis=0)
    fv = np.tile(np.array([0.3, -1.0, 0.2]), (fc.shape[0], 1)) / fc.shape[0]
    mu, lam = lame()
    res = (nx + 4 * OFFSET, ny + 4 * OFFSET, nz + 4 * OFFSET)
    return Problem(f"blocky{seed}", res, 1.0 / nx, mu, lam, bb, d, dc, da, dv, fc, fv, mg_levels=mg_levels,
                   domain=(nx, ny, nz), meta={"seed": seed})


def counts(p: Problem):
    """(active cells, active nodes, DOF)."""
    nn = p.active_nodes().shape[0]
    return p.n_active_cells, nn, 3 * nn


def config(name: str, **kw) -> Problem:
    """The named benchmark / parity configurations."""
    if name == "cfg1":
        return box(64, 32, 32, seed=kw.get("seed", 0), lo=0.05, hi=1.0, mg_levels=3, load="edge", name="cfg1_box64x32x32")
    if name == "cfg2":
        return sphere_shell(kw.get("n", 256), seed=kw.get("seed", 0), mg_levels=kw.get("mg_levels", 5),
                            name="cfg2_shell%d" % kw.get("n", 256))
    if name == "cfg3":
        return box(128, 64, 64, seed=kw.get("seed", 0), lo=0.1, hi=1.0, mg_levels=kw.get("mg_levels", 4), load="point",
                   name="cfg3_box128x64x64")
    if name == "cfg4":
        return band(512, 256, 256, 24, seed=kw.get("seed", 0), mg_levels=kw.get("mg_levels", 6))
    if name == "cfg5":
        return ellipsoid_shell(seed=kw.get("seed", 0), progress=kw.get("progress"))
    raise KeyError(name)
