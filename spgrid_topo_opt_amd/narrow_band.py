"""Narrow-band maintenance on the GPU: the thin / fat / fatter block sets that
decide which blocks the FEM solve sees (reference:
topo_opt/spgrid_topo_opt.cpp:92-241 update_thin_page_map / update_fat_page_map,
:1404-1477 filter_isolated_blocks, :1289-1299 the per-iteration trigger,
:1060-1160 packing of the fatter blocks into FEMInputs).

All sets are boolean masks over the container blocks (4x4x4 cells; the
container holds every block touched by an inside cell or one of its nodes,
spgrid_topo_opt.cpp:361-367).  Cell arrays are (nblocks*64,) in block order,
cell index block*64 + i*16 + j*4 + k, as in the FEM inputs.

* thin   = fat blocks whose maximum density reaches minimum_density +
           active_threshold, or that hold a cell next to a force
           (spgrid_topo_opt.cpp:92-118);
* the connectivity filter keeps the one 6-connected component of thin blocks
  that holds more than 30% of them (:1404-1477);
* fat    = container blocks hit by a cell of a thin block whose density
           exceeds active_threshold, shifted by fat_margin*(+-1,+-1,+-1)
           (:155-187);
* fatter = container blocks at fat base + {0,1}^3 blocks (:189-200).

The FEM solve runs on the fatter blocks; cells of fatter blocks outside the
fat set carry density 0 (node-only blocks, :1087-1122).
"""
from __future__ import annotations

import numpy as np
import torch


class BandError(RuntimeError):
    """The reference's TC_ERROR cases of the connectivity filter."""


def _local_offsets(device):
    l = torch.arange(64, device=device)
    return torch.stack([l >> 4, (l >> 2) & 3, l & 3], 1)  # (64,3) i,j,k


class NarrowBand:
    """Container blocks plus the thin / fat / fatter masks of one optimizer.

    block_base: (nb,3) cell coordinates of the container blocks (multiples of 4);
    inside:     (nb*64,) cells inside the container;
    has_force:  (nb*64,) inside cells that share a node with a force
                (spgrid_topo_opt.h:580-590).
    Every mask starts as the whole container (spgrid_topo_opt.cpp:389-391).
    """

    def __init__(self, block_base, inside, has_force, minimum_density=0.0, active_threshold=1e-6, fat_margin=1,
                 fraction_to_keep=1.0, connectivity_filtering=True, device="cuda"):
        if fraction_to_keep != 1.0:
            # spgrid_topo_opt.cpp:121-150 indexes block_meta by the position in
            # the old thin list while filling it by fat-list position; only the
            # default (no selection) has a defined meaning
            raise ValueError("fraction_to_keep other than 1.0 is not supported")
        self.device = torch.device(device)
        bb = torch.as_tensor(np.asarray(block_base), dtype=torch.int64, device=self.device)
        if (bb % 4 != 0).any():
            raise ValueError("block bases must be multiples of 4")
        self.nb = bb.shape[0]
        self.base = bb
        self.bc = bb // 4
        self.inside = torch.as_tensor(np.asarray(inside), dtype=torch.bool, device=self.device).reshape(-1)
        self.has_force_block = torch.as_tensor(np.asarray(has_force), dtype=torch.bool,
                                               device=self.device).reshape(self.nb, 64).any(1)
        self.minimum_density = float(minimum_density)
        self.active_threshold = float(active_threshold)
        self.fat_margin = int(fat_margin)
        self.connectivity_filtering = bool(connectivity_filtering)
        # dense block-index volume with a margin so every shifted lookup is in range
        pad = (self.fat_margin + 3) // 4 + 1
        self.lo = self.bc.min(0).values - pad
        dims = self.bc.max(0).values + pad + 1 - self.lo
        self.dims = dims
        self.vol = torch.full((int(dims[0]), int(dims[1]), int(dims[2])), -1, dtype=torch.int64, device=self.device)
        rel = self.bc - self.lo
        self.vol[rel[:, 0], rel[:, 1], rel[:, 2]] = torch.arange(self.nb, device=self.device)
        # 6-neighbour table of the container blocks (for the connectivity filter)
        six = torch.tensor([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], device=self.device)
        self.nbr6 = self.lookup(self.bc[:, None, :] + six[None])
        self.loc = _local_offsets(self.device)
        full = torch.ones(self.nb, dtype=torch.bool, device=self.device)
        self.thin, self.fat, self.fatter = full.clone(), full.clone(), full.clone()
        self.components = []  # sizes of the last connectivity pass

    # ------------------------------------------------------------------
    def lookup(self, bc):
        """container block index of block coordinates (..., 3), -1 if absent"""
        rel = bc - self.lo
        ok = ((rel >= 0) & (rel < self.dims)).all(-1)
        rel = torch.where(ok[..., None], rel, torch.zeros_like(rel))
        idx = self.vol[rel[..., 0], rel[..., 1], rel[..., 2]]
        return torch.where(ok, idx, torch.full_like(idx, -1))

    def update_thin(self, density):
        """spgrid_topo_opt.cpp:92-118 (fraction_to_keep = 1)"""
        d = density.reshape(self.nb, 64)
        dmax = d.max(1).values
        keep = (dmax >= self.minimum_density + self.active_threshold) | self.has_force_block
        self.thin = self.fat & keep

    def filter_isolated(self):
        """spgrid_topo_opt.cpp:1404-1477: label the 6-connected components of
        the thin blocks; with more than one, keep the single component above
        30% of the thin blocks."""
        thin = self.thin
        n = int(thin.sum())
        if n == 0:
            raise BandError("no thin blocks left")
        big = self.nb
        lab = torch.where(thin, torch.arange(self.nb, device=self.device), torch.full_like(self.nbr6[:, 0], big))
        lab_ext = torch.cat([lab, lab.new_full((1,), big)])
        nbr = torch.where(self.nbr6 >= 0, self.nbr6, torch.full_like(self.nbr6, big))
        while True:
            m = torch.minimum(lab, lab_ext[nbr].min(1).values)
            m = torch.where(thin, m, lab)
            # pointer jumping: a label is itself a thin block's index
            m = torch.where(thin, m[m.clamp(max=big - 1)], m)
            m = torch.where(thin, torch.minimum(m, lab), m)
            if torch.equal(m, lab):
                break
            lab = m
            lab_ext = torch.cat([lab, lab.new_full((1,), big)])
        roots, sizes = torch.unique(lab[thin], return_counts=True)
        self.components = sizes.tolist()
        if len(self.components) == 1:
            return
        frac = sizes.to(torch.float64) / n
        big_ones = (frac > 0.3).nonzero().reshape(-1)
        if len(big_ones) > 1:
            raise BandError("There are more than two connected components with fraction > 30%. "
                            "Something may be wrong.")
        if len(big_ones) == 0:
            raise BandError("All compoments have fraction < 30%. How can this be so fragmental?")
        keep_root = roots[big_ones[0]]
        self.thin = thin & (lab == keep_root)

    def update_fat(self, density):
        """spgrid_topo_opt.cpp:155-200"""
        d = density.reshape(self.nb, 64)
        act = (d > self.active_threshold) & self.thin[:, None]
        bi, li = act.nonzero(as_tuple=True)
        coord = self.base[bi] + self.loc[li]
        fat = torch.zeros(self.nb, dtype=torch.bool, device=self.device)
        m = self.fat_margin
        for sx in (-1, 1):
            for sy in (-1, 1):
                for sz in (-1, 1):
                    o = torch.tensor([sx * m, sy * m, sz * m], device=self.device)
                    idx = self.lookup(torch.div(coord + o, 4, rounding_mode="floor"))
                    fat[idx[idx >= 0]] = True
        self.fat = fat
        fatter = torch.zeros_like(fat)
        fb = self.bc[fat]
        for c in range(8):
            o = torch.tensor([(c >> 2) & 1, (c >> 1) & 1, c & 1], device=self.device)
            idx = self.lookup(fb + o)
            fatter[idx[idx >= 0]] = True
        self.fatter = fatter

    def update(self, density):
        """the page-map update of one optimizer iteration (spgrid_topo_opt.cpp:1292-1299)"""
        density = density.to(self.device, torch.float64)
        self.update_thin(density)
        if self.connectivity_filtering:
            self.filter_isolated()
        self.update_fat(density)

    # ------------------------------------------------------------------
    def fem_blocks(self):
        """container indices of the blocks handed to the solve (fatter set)"""
        return self.fatter.nonzero().reshape(-1)

    def fem_density(self, density, blocks=None):
        """FEM input densities of the fatter blocks: max(rho, DBL_MIN) for
        inside cells of fat blocks, 0 elsewhere (spgrid_topo_opt.cpp:1087-1122)"""
        blocks = self.fem_blocks() if blocks is None else blocks
        d = density.reshape(self.nb, 64)[blocks]
        live = self.inside.reshape(self.nb, 64)[blocks] & self.fat[blocks, None]
        return torch.where(live, d.clamp(min=np.finfo(np.float64).tiny), torch.zeros_like(d))


def _block_keys(bc):
    bc = np.asarray(bc, np.int64)
    return (bc[..., 0] << 42) + (bc[..., 1] << 21) + bc[..., 2]


def block_lookup(block_base, bc):
    """index of each block coordinate (..., 3) in block_base, -1 if absent"""
    keys = _block_keys(np.asarray(block_base, np.int64) // 4 + (1 << 20))
    order = np.argsort(keys)
    sk = keys[order]
    q = _block_keys(np.asarray(bc, np.int64) + (1 << 20))
    pos = np.clip(np.searchsorted(sk, q), 0, len(sk) - 1)
    return np.where(sk[pos] == q, order[pos], -1)


def force_cells(block_base, inside, force_coord):
    """(nb*64,) inside cells that share a node with a force: the 8 cells
    around each force node (spgrid_topo_opt.h:580-590)."""
    nb = np.asarray(block_base).shape[0]
    out = np.zeros(nb * 64, bool)
    f = np.asarray(force_coord, np.int64).reshape(-1, 3)
    for c in range(8):
        cell = f - np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        b = block_lookup(block_base, cell >> 2)
        loc = ((cell[:, 0] & 3) << 4) + ((cell[:, 1] & 3) << 2) + (cell[:, 2] & 3)
        out[(b * 64 + loc)[b >= 0]] = True
    return out & np.asarray(inside).reshape(-1)
