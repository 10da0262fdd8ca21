"""TEST INFRASTRUCTURE ONLY: a plain-Python restatement of the reference's
narrow-band maintenance, used to check spgrid_topo_opt_amd.narrow_band.

Follows, loop for loop:
  update_thin_page_map    topo_opt/spgrid_topo_opt.cpp:92-118 (fraction_to_keep = 1)
  filter_isolated_blocks  topo_opt/spgrid_topo_opt.cpp:1404-1477 (BFS, 6-connectivity)
  update_fat_page_map     topo_opt/spgrid_topo_opt.cpp:155-200
  FEM input packing       topo_opt/spgrid_topo_opt.cpp:1060-1160
Blocks are keyed by their block coordinates (cell coordinates // 4); the
SPGrid linear offsets of the reference only order the loops, and no result
depends on that order (page-map sets, component sizes as a multiset).

Pinned by tests/test_narrow_band.py against hand-built cases whose sets
follow from the rules directly (a plus-shaped band, an isolated island, a
single active cell near a block corner).
"""
from collections import deque

import numpy as np


class BandOracle:
    def __init__(self, block_base, inside, has_force, minimum_density=0.0, active_threshold=1e-6, fat_margin=1):
        self.bc = [tuple(int(v) // 4 for v in b) for b in np.asarray(block_base).tolist()]
        self.index = {b: i for i, b in enumerate(self.bc)}
        self.inside = np.asarray(inside, bool).reshape(len(self.bc), 64)
        self.has_force = np.asarray(has_force, bool).reshape(len(self.bc), 64)
        self.minimum_density = minimum_density
        self.active_threshold = active_threshold
        self.fat_margin = fat_margin
        self.container = set(self.bc)
        self.thin = set(self.bc)
        self.fat = set(self.bc)
        self.fatter = set(self.bc)
        self.components = []

    def _d(self, density, b, l):
        return density[self.index[b] * 64 + l]

    def update_thin(self, density):
        thin = set()
        for b in self.fat:
            dmax = 0.0  # BlockMeta starts at density_max = 0
            force = False
            for l in range(64):
                dmax = max(dmax, self._d(density, b, l))
                force |= bool(self.has_force[self.index[b], l])
            if dmax >= self.minimum_density + self.active_threshold or force:
                thin.add(b)
        self.thin = thin

    def filter_isolated(self):
        visited = set()
        color = {}
        sizes = []
        for start in sorted(self.thin):
            if start in visited:
                continue
            q = deque()
            count = 0

            def visit(nb):
                nonlocal count
                if nb in self.thin and nb not in visited:
                    q.append(nb)
                    visited.add(nb)
                    color[nb] = len(sizes)
                    count += 1

            visit(start)
            while q:
                u = q.popleft()
                for o in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    visit((u[0] + o[0], u[1] + o[1], u[2] + o[2]))
            sizes.append(count)
        total = sum(sizes)
        assert total == len(self.thin) and sizes
        self.components = sizes
        if len(sizes) == 1:
            return
        big = [i for i, s in enumerate(sizes) if s / total > 0.3]
        if len(big) > 1:
            raise RuntimeError("There are more than two connected components with fraction > 30%.")
        if not big:
            raise RuntimeError("All compoments have fraction < 30%.")
        self.thin = {b for b in self.thin if color[b] == big[0]}

    def update_fat(self, density):
        m = self.fat_margin
        fat = set()
        for b in self.thin:
            for l in range(64):
                d = self._d(density, b, l)
                if d > self.active_threshold:
                    x, y, z = b[0] * 4 + (l >> 4), b[1] * 4 + ((l >> 2) & 3), b[2] * 4 + (l & 3)
                    for sx in (-1, 1):
                        for sy in (-1, 1):
                            for sz in (-1, 1):
                                nb = ((x + sx * m) // 4, (y + sy * m) // 4, (z + sz * m) // 4)
                                if nb in self.container:
                                    fat.add(nb)
        self.fat = fat
        fatter = set()
        for b in fat:
            for c in range(8):
                nb = (b[0] + ((c >> 2) & 1), b[1] + ((c >> 1) & 1), b[2] + (c & 1))
                if nb in self.container:
                    fatter.add(nb)
        self.fatter = fatter

    def update(self, density, connectivity=True):
        self.update_thin(density)
        if connectivity:
            self.filter_isolated()
        self.update_fat(density)

    def mask(self, s):
        out = np.zeros(len(self.bc), bool)
        for b in s:
            out[self.index[b]] = True
        return out

    def fem_density(self, density):
        """(blocks, (nblocks,64) densities) of the FEM inputs"""
        blocks = np.nonzero(self.mask(self.fatter))[0]
        out = np.zeros((len(blocks), 64))
        tiny = np.finfo(np.float64).tiny
        for r, bi in enumerate(blocks):
            only_nodes = self.bc[bi] not in self.fat
            for l in range(64):
                if self.inside[bi, l] and not only_nodes:
                    out[r, l] = max(density[bi * 64 + l], tiny)
        return blocks, out
