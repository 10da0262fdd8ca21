"""Independent numpy/scipy model used to pin the oracle (test infrastructure).

Nothing here shares code with the oracle: the element matrix comes from 2x2x2
Gauss quadrature of the Q1 shape-function gradients (not the oracle's closed
1-D integrals), the global operators are assembled as scipy sparse matrices,
and the prolongation is built from trilinear weights of node positions.
"""
import itertools

import numpy as np
import scipy.sparse as sp

NODES = np.array([[(a >> 2) & 1, (a >> 1) & 1, a & 1] for a in range(8)], dtype=float)
BPAD = np.array([4, 4, 8])


def shape_grads(x):
    """d N_a / d x_d at point x of the unit cube -> (8,3)."""
    g = np.zeros((8, 3))
    for a in range(8):
        n = NODES[a]
        f = [(x[d] if n[d] else 1 - x[d]) for d in range(3)]
        df = [(1.0 if n[d] else -1.0) for d in range(3)]
        g[a, 0] = df[0] * f[1] * f[2]
        g[a, 1] = f[0] * df[1] * f[2]
        g[a, 2] = f[0] * f[1] * df[2]
    return g


def element_tables_gauss():
    """K_mu, K_la [8][8][3][3] by 2x2x2 Gauss quadrature of
    mu (grad u : grad v + grad u : grad v^T) and la (div u div v)."""
    gp = [0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)]
    kmu = np.zeros((8, 8, 3, 3))
    kla = np.zeros((8, 8, 3, 3))
    for x in itertools.product(gp, gp, gp):
        G = shape_grads(np.array(x)) * 1.0  # weight 1/8 each, times 8 points -> volume 1
        w = 1.0 / 8.0
        for a in range(8):
            for b in range(8):
                lap = G[a] @ G[b]
                for v in range(3):
                    for wc in range(3):
                        kmu[a, b, v, wc] += w * ((lap if v == wc else 0.0) + G[a, wc] * G[b, v])
                        kla[a, b, v, wc] += w * G[a, v] * G[b, wc]
    return kmu, kla


def rigid_modes():
    """6 rigid-body modes on the 8 cell nodes -> (24,6)."""
    R = np.zeros((24, 6))
    for a in range(8):
        x = NODES[a]
        for v in range(3):
            R[3 * a + v, v] = 1.0
        # rotations: omega x x
        R[3 * a:3 * a + 3, 3] = np.cross([1, 0, 0], x)
        R[3 * a:3 * a + 3, 4] = np.cross([0, 1, 0], x)
        R[3 * a:3 * a + 3, 5] = np.cross([0, 0, 1], x)
    return R


def node_index(xyz):
    return {tuple(map(int, c)): i for i, c in enumerate(xyz)}


def assemble_K(xyz, cells_xyz, mu, la, dx, kmu, kla):
    """Global dx * sum_cells (mu K_mu + la K_la) over given node set."""
    idx = node_index(xyz)
    rows, cols, vals = [], [], []
    for c, m, l in zip(cells_xyz, mu, la):
        if m == 0 and l == 0:
            continue
        ids = []
        for a in range(8):
            key = (int(c[0] + NODES[a, 0]), int(c[1] + NODES[a, 1]), int(c[2] + NODES[a, 2]))
            ids.append(idx.get(key, -1))
        Ke = (m * kmu + l * kla) * dx
        for a in range(8):
            if ids[a] < 0:
                continue
            for b in range(8):
                if ids[b] < 0:
                    continue
                for v in range(3):
                    for w in range(3):
                        rows.append(3 * ids[a] + v)
                        cols.append(3 * ids[b] + w)
                        vals.append(Ke[a, b, v, w])
    n = 3 * len(xyz)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def prolongation(fine_xyz, coarse_xyz):
    """Raw trilinear prolongation (no projection): coarse node c sits at fine
    position 2c - B; weight = prod over axes of max(0, 1 - |x_f - (2c-B)|/2)."""
    cidx = node_index(coarse_xyz)
    rows, cols, vals = [], [], []
    for i, f in enumerate(fine_xyz):
        base = (np.asarray(f) + BPAD) // 2
        for o in itertools.product((0, 1), repeat=3):
            c = base + np.array(o)
            j = cidx.get(tuple(int(t) for t in c), -1)
            if j < 0:
                continue
            pos = 2 * c - BPAD
            w = np.prod(np.maximum(0.0, 1.0 - np.abs(np.asarray(f) - pos) / 2.0))
            if w == 0:
                continue
            for v in range(3):
                rows.append(3 * i + v)
                cols.append(3 * j + v)
                vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * len(fine_xyz), 3 * len(coarse_xyz)))


def dirichlet_mask(flags):
    m = np.zeros((len(flags), 3), bool)
    for v in range(3):
        m[:, v] = (flags & (0x20 << v)) != 0
    return m.reshape(-1)


def delete_dirichlet(A, dmask):
    """Rows of Dirichlet DOFs -> identity, columns -> 0 (other rows)."""
    A = A.tolil(copy=True)
    idx = np.nonzero(dmask)[0]
    A = A.tocsc()
    keep = sp.diags((~dmask).astype(float))
    A = (keep @ A @ keep).tolil()
    for i in idx:
        A[i, i] = 1.0
    return A.tocsr()
