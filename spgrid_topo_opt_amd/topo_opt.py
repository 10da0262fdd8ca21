"""Python driver seat (reference: scripts/topo_opt.py, topo_opt/spgrid_topo_opt.cpp
iterate() at :1289-1331): differentiable GPU FEM solves for torch.

* :class:`Compliance` -- ``torch.autograd.Function`` whose forward is the GPU
  MG-PCG solve and returns the reference-scaled compliance
  c = sum_e u_e^T K0 u_e E(rho_e)  (= f.u / dx), and whose backward is the
  hand-written sensitivity kernel dc/drho_e = -p rho_e^(p-1)(1-Emin) u_e^T K0 u_e
  (objective/minimal_compliance.cpp:74-146).  Only u is saved.
* :class:`Displacement` -- forward returns the displacement field (input block
  order); backward runs the adjoint MG-PCG solve K lambda = g on the same
  hierarchy and returns -dx E'(rho_e) lambda_e^T K0 u_e, so any objective of u
  (e.g. a load-node displacement) is differentiable.
* :func:`filter_sensitivity` / :func:`oc_update` -- the density-weighted cone
  filter (spgrid_topo_opt.cpp:527-574) and the bisection optimality-criterion
  update (optimizer/optimality_criterion.cpp:19-146), HIP kernels of the same
  library (spg_filter_sensitivity, spg_oc_trial).

Densities are tensors of shape (nblocks*64,) in the input block order of the
Problem (index block*64 + i*16 + j*4 + k), float64 on cuda.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import torch

from . import _lib
from . import narrow_band as _nb


class FEM:
    """A GPU solver bound to one block topology, with solve settings."""

    def __init__(self, prob, tolerance=1e-6, precision=0, inner_tolerance=0.0, max_passes=30, max_iterations=500,
                 device=0, warm_start=0):
        """warm_start: 0 every solve starts from u_D; 1 from the previous
        solution when its residual is below the first solve's zero-guess
        residual; 2 always from it (forced_reuse).  Under a warm start the
        tolerance is relative to that first residual, as the reference's
        kept state does (SPGrid_CPU_Solver.cpp:247-285, 318-326)."""
        self.prob = prob
        # a density update refactors the bottom for about one solve, so the
        # nested-dissection Cholesky takes bottoms above 4096 unknowns, as in
        # fem_solve (it factors faster than the dense inverse there)
        self.solver = _lib.Solver(prob, device=device, bottom_dense_max=4096)
        self.nin = prob.block_base.shape[0] * 64
        self.kw = dict(tolerance=tolerance, precision=precision, inner_tolerance=inner_tolerance if precision == 0
                       else 0.0, max_passes=max_passes if precision == 0 else 1, max_iterations=max_iterations)
        self.warm_start = warm_start
        self.reference_linf = 0.0  # zero-guess residual of the first solve
        self.has_guess = False
        self.last_info = None
        self._rho_host = None
        self._rho_dev = None
        self._inside_dev = None

    def set_density(self, rho: torch.Tensor):
        # cells of the design domain never reach zero density: the caller feeds
        # max(rho, DBL_MIN) so the solver keeps them (spgrid_topo_opt.cpp:1117-1122)
        if rho.is_cuda:
            # device path: SIMP map and checks on the GPU, no host round trip
            r = self._device_density(rho)
            if self._rho_dev is None or not torch.equal(r, self._rho_dev):
                torch.cuda.synchronize(rho.device)
                self.solver.update_density_device(r.data_ptr())
                self._rho_dev = r
            return
        r = rho.detach().to(torch.float64).cpu().numpy()
        inside = self.prob.density.reshape(-1) > 0
        r = np.where(inside, np.maximum(r, np.finfo(np.float64).tiny), r)
        if self._rho_host is None or not np.array_equal(r, self._rho_host):
            self.solver.update_density(r.reshape(-1, 64))
            self._rho_host = r.copy()

    def _device_density(self, rho):
        if self._inside_dev is None or self._inside_dev.device != rho.device:
            self._inside_dev = torch.as_tensor(self.prob.density.reshape(-1) > 0, device=rho.device)
        r = rho.detach().to(torch.float64).reshape(-1)
        return torch.where(self._inside_dev, r.clamp(min=np.finfo(np.float64).tiny), r).contiguous()

    def assume_density(self, rho: torch.Tensor):
        """Record `rho` (cuda) as the density the solver already holds, e.g.
        the one it was built with, so the next set_density(rho) is free."""
        self._rho_dev = self._device_density(rho)

    def close(self):
        self.solver.close()

    def solve(self) -> dict:
        warm = self.warm_start if (self.reference_linf > 0 or self.has_guess) else 0
        _, info = self.solver.solve(want_output=False, warm_start=warm, reference_linf=self.reference_linf,
                                    **self.kw)
        if self.warm_start and not self.reference_linf > 0:
            self.reference_linf = info["r0_linf"]
        self.has_guess = False
        self.last_info = info
        return info

    def carry_over(self, old: "FEM", old_blocks: torch.Tensor, new_blocks: torch.Tensor, nblocks_total: int):
        """Start the next warm solve from `old`'s solution: blocks are indices
        into a common block list; blocks new to this solver start at zero."""
        if not self.warm_start or old.last_info is None:
            return
        dev = new_blocks.device
        u_old = old.displacement(dev).reshape(-1, 192)
        pos = torch.full((nblocks_total,), -1, dtype=torch.int64, device=dev)
        pos[old_blocks] = torch.arange(old_blocks.numel(), device=dev)
        src = pos[new_blocks]
        u_new = torch.where((src >= 0)[:, None], u_old[src.clamp(min=0)], torch.zeros_like(u_old[:1]))
        u_new = u_new.reshape(-1).contiguous()
        torch.cuda.synchronize(dev)
        self.solver.set_initial_guess_device(u_new.data_ptr())
        self.reference_linf = old.reference_linf
        self.has_guess = True

    def displacement(self, device) -> torch.Tensor:
        u = torch.empty(self.nin * 3, dtype=torch.float64, device=device)
        torch.cuda.synchronize(device)
        self.solver.displacement_device(u.data_ptr())
        return u

    def sensitivity(self, u: torch.Tensor, lam: torch.Tensor = None):
        dc = torch.empty(self.nin, dtype=torch.float64, device=u.device)
        torch.cuda.synchronize(u.device)
        obj = self.solver.sensitivity_device(dc.data_ptr(), lam_ptr=None if lam is None else lam.data_ptr(),
                                             u_ptr=u.data_ptr())
        return obj, dc

    def adjoint(self, g: torch.Tensor) -> torch.Tensor:
        g = g.detach().to(torch.float64).contiguous()
        lam = torch.empty_like(g)
        torch.cuda.synchronize(g.device)
        kw = dict(self.kw)
        self.solver.solve_adjoint_device(g.data_ptr(), lam.data_ptr(), **kw)
        return lam


class Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, fem: FEM):
        fem.set_density(rho)
        fem.solve()
        u = fem.displacement(rho.device)
        obj, _ = fem.sensitivity(u)
        ctx.fem = fem
        ctx.save_for_backward(u)
        return torch.tensor(obj, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad):
        (u,) = ctx.saved_tensors
        _, dc = ctx.fem.sensitivity(u)
        return grad * dc, None


class Displacement(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, fem: FEM):
        fem.set_density(rho)
        fem.solve()
        u = fem.displacement(rho.device)
        ctx.fem = fem
        ctx.save_for_backward(u)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        (u,) = ctx.saved_tensors
        lam = ctx.fem.adjoint(grad_u)
        _, g = ctx.fem.sensitivity(u, lam)
        return g * ctx.fem.prob.dx, None


def compliance(rho, fem):
    return Compliance.apply(rho, fem)


def displacement(rho, fem):
    return Displacement.apply(rho, fem)


# --------------------------------------------------------------------------
# density filter and OC update (next to the solve on the same blocks)
# --------------------------------------------------------------------------
class CellGrid:
    """Input blocks as a cell grid with a 27-block neighbour table (torch)."""

    def __init__(self, prob, device="cuda"):
        bb = prob.block_base.astype(np.int64) // 4
        offs = np.array([(s // 9 - 1, (s // 3) % 3 - 1, s % 3 - 1) for s in range(27)], np.int64)
        nbr = _nb.block_lookup(prob.block_base, bb[:, None, :] + offs[None])
        self.nb = len(bb)
        self.nbr = torch.as_tensor(nbr, device=device)
        self.inside = torch.as_tensor(prob.density.reshape(-1) > 0, device=device)
        self.nbr32 = self.nbr.to(torch.int32).contiguous()
        self.inside_u8 = self.inside.to(torch.uint8).contiguous()

def filter_sensitivity(grid: CellGrid, dc, rho, filter_size=1.5):
    """dc1(x) = sum_o w(o) dc(x+o) rho(x+o) [x+o inside] / (sum_o w(o) [x+o inside] (rho(x)+1e-3)),
    w(o) = max(0, r - |o|) over |o_i| <= floor(r) (spgrid_topo_opt.cpp:527-574),
    by the HIP kernel spg_filter_sensitivity, for every cell (inside the
    container or not, as smooth_dc writes every cell of the fat blocks; 0
    where no window cell is inside)."""
    if not dc.is_cuda:
        raise RuntimeError("filter_sensitivity runs on the GPU (HIP kernel); got a CPU tensor")
    dc = dc.detach().to(torch.float64).contiguous()
    rho = rho.detach().to(torch.float64).contiguous()
    out = torch.empty_like(dc)
    torch.cuda.synchronize(dc.device)
    _lib.filter_sensitivity(grid.nb, grid.nbr32.data_ptr(), grid.inside_u8.data_ptr(), dc.data_ptr(), rho.data_ptr(),
                            float(filter_size), out.data_ptr())
    return out


def oc_update(rho, dc, inside, volume_fraction, step_limit=0.2, min_density=0.0, container_cells=None, fixed=None,
              fixed_cell_density=1.0, exclude_fixed_cells=True, smart_oc=False):
    """Bisection on the Lagrange multiplier (optimizer/optimality_criterion.cpp:19-146):
    new = clamp(old*sqrt(max(0,-dc-shift)/mid), old-step, old+step), clamped to
    [min_density, 1], on the `inside` cells; `fixed` cells (a bool mask) take
    fixed_cell_density instead and, with exclude_fixed_cells, neither count in
    the volume nor in the container size.  shift = min(0, min over the inside
    cells of -dc) with smart_oc, else 0.  The volume must equal volume_fraction
    times (container cells - fixed inside cells); container cells default to
    the updated cells.  Each trial is the HIP kernel spg_oc_trial; the
    bisection sequence is the reference's."""
    if not rho.is_cuda:
        raise RuntimeError("oc_update runs on the GPU (HIP kernel); got a CPU tensor")
    n = inside.sum().item() if container_cells is None else container_cells
    mask = inside.to(torch.uint8)
    n_fixed = 0
    if fixed is not None:
        fx = fixed & inside
        mask = mask | (fx.to(torch.uint8) << 1)
        if exclude_fixed_cells:
            n_fixed = int(fx.sum().item())
    target = volume_fraction * (n - n_fixed)
    rho = rho.detach().to(torch.float64).contiguous()
    dc = dc.detach().to(torch.float64).contiguous()
    mask = mask.contiguous()
    shift = min(0.0, (-dc[inside]).min().item()) if smart_oc and bool(inside.any()) else 0.0
    torch.cuda.synchronize(rho.device)
    kw = dict(fixed_density=fixed_cell_density, exclude_fixed=exclude_fixed_cells, dc_shift=shift)
    lower, upper = 1e-20, 1e20
    mid = 0.5 * (lower + upper)
    while lower * (1 + 1e-7) < upper:
        mid = 0.5 * (lower + upper)
        vol, _ = _lib.oc_trial(rho.numel(), mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), mid, step_limit,
                               min_density, **kw)
        if vol > target:
            lower = mid
        else:
            upper = mid
    new = torch.empty_like(rho)
    _, change = _lib.oc_trial(rho.numel(), mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), mid, step_limit,
                              min_density, new.data_ptr(), **kw)
    return new, change


class TopoOpt:
    """SIMP/OC loop on the GPU (reference iterate(), spgrid_topo_opt.cpp:1289-1331):
    narrow-band update, solve, sensitivity, filter, OC update.

    Without `grid_update_start` the block set stays the whole container.  With
    it, from that iteration on every iteration first updates the thin / fat /
    fatter block sets (narrow_band.NarrowBand); the solve then runs on the
    fatter blocks (rebuilt when the fat or fatter set changes), and the
    gradient clear, filter swap and OC update cover the fat blocks only, as
    clear_gradients (spgrid_topo_opt.h:869-878), smooth_dc
    (spgrid_topo_opt.cpp:527-574) and the OC optimizer
    (optimizer/optimality_criterion.cpp:29) do."""

    def __init__(self, prob, volume_fraction=0.1, step_limit=0.2, filter_size=1.5, grid_update_start=None,
                 active_threshold=1e-6, fat_margin=1, connectivity_filtering=True, fixed=None,
                 fixed_cell_density=1.0, exclude_fixed_cells=True, smart_oc=False, warm_starting=False,
                 forced_reuse=True, **fem_kw):
        """`fixed`: optional per-cell bool mask (input block order) of
        fixed-density cells (flags.set_fixed_density, e.g. the cells that
        add_force marks with fix_cells_near_force, spgrid_topo_opt.h:577-592);
        they start at and keep fixed_cell_density.  exclude_fixed_cells and
        smart_oc are the OC optimizer's options
        (optimizer/optimality_criterion.cpp:31-66).  warm_starting and
        forced_reuse are the caller's solver-state options
        (spgrid_topo_opt.cpp:1013-1015): each solve starts from the previous
        one's displacement, carried over when the band changes."""
        self.prob = prob
        if warm_starting:
            fem_kw = dict(fem_kw, warm_start=2 if forced_reuse else 1)
        self.fem_kw = fem_kw
        self.fem = FEM(prob, **fem_kw)
        self.grid = CellGrid(prob)
        self.volume_fraction = volume_fraction
        self.step_limit = step_limit
        self.filter_size = filter_size
        inside = self.grid.inside
        self.fixed = None if fixed is None else torch.as_tensor(fixed, device=inside.device).reshape(-1) & inside
        self.oc_kw = dict(fixed=self.fixed, fixed_cell_density=fixed_cell_density,
                          exclude_fixed_cells=exclude_fixed_cells, smart_oc=smart_oc)
        self.rho = torch.where(inside, torch.full_like(inside, volume_fraction, dtype=torch.float64),
                               torch.zeros_like(inside, dtype=torch.float64))
        if self.fixed is not None:
            self.rho = torch.where(self.fixed, torch.full_like(self.rho, fixed_cell_density), self.rho)
        self.history = []
        self.iter = 0
        self.grid_update_start = grid_update_start
        self.band = None
        if grid_update_start is not None:
            ins = (prob.density.reshape(-1) > 0)
            hf = _nb.force_cells(prob.block_base, ins, prob.force_coord)
            self.band = _nb.NarrowBand(prob.block_base, ins, hf, minimum_density=prob.min_density,
                                       active_threshold=active_threshold, fat_margin=fat_margin,
                                       connectivity_filtering=connectivity_filtering, device=inside.device)
            self.dc0 = torch.zeros_like(self.rho)
            self._fem_blocks = torch.arange(self.band.nb, device=inside.device)
            self._fem_fat = self.band.fat.clone()
            self._cells = torch.arange(64, device=inside.device)

    def close(self):
        self.fem.close()

    def iterate(self):
        if self.band is not None:
            return self._iterate_band()
        rho = self.rho.clone().requires_grad_(True)
        c = compliance(rho, self.fem)
        c.backward()
        dc = filter_sensitivity(self.grid, rho.grad.detach(), self.rho, self.filter_size)
        new, change = oc_update(self.rho, dc, self.grid.inside, self.volume_fraction, self.step_limit,
                                self.prob.min_density, **self.oc_kw)
        self.rho = new
        self.iter += 1
        rec = {"compliance": c.item(), "change": change, "iterations": self.fem.last_info["iterations"],
               "solve_s": self.fem.last_info["solve_seconds"]}
        self.history.append(rec)
        return rec

    def _iterate_band(self):
        band = self.band
        dev = self.rho.device
        t0 = time.perf_counter()
        if self.iter >= self.grid_update_start:
            band.update(self.rho)
        blocks = band.fem_blocks()
        torch.cuda.synchronize(dev)
        tb = time.perf_counter()
        rebuilt = False
        if not (torch.equal(blocks, self._fem_blocks) and torch.equal(band.fat, self._fem_fat)):
            # new topology: a solver for the fatter blocks (fem_solve rebuilds its hierarchy likewise)
            sub = dataclasses.replace(self.prob, block_base=self.prob.block_base[blocks.cpu().numpy()],
                                      density=band.fem_density(self.rho, blocks).cpu().numpy())
            old = self.fem
            self.fem = FEM(sub, **self.fem_kw)
            self.fem.carry_over(old, self._fem_blocks, blocks, self.band.nb)
            old.close()
            self._fem_blocks, self._fem_fat = blocks, band.fat.clone()
            rebuilt = True
        torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        rho_fem = band.fem_density(self.rho, blocks).reshape(-1).clone().requires_grad_(True)
        if rebuilt:
            self.fem.assume_density(rho_fem)  # built with exactly these densities
        c = compliance(rho_fem, self.fem)
        c.backward()
        torch.cuda.synchronize(dev)
        t2 = time.perf_counter()
        fat_cells = band.fat.repeat_interleave(64)
        idx = (blocks[:, None] * 64 + self._cells[None]).reshape(-1)
        live = (self.grid.inside & fat_cells)[idx]
        dc0 = torch.where(fat_cells, torch.zeros_like(self.dc0), self.dc0)
        dc0[idx[live]] = rho_fem.grad.detach()[live]
        dc1 = filter_sensitivity(self.grid, dc0, self.rho, self.filter_size)
        self.dc0 = torch.where(fat_cells, dc1, dc0)
        upd = self.grid.inside & fat_cells
        new, change = oc_update(self.rho, self.dc0, upd, self.volume_fraction, self.step_limit,
                                self.prob.min_density, container_cells=self.grid.inside.sum().item(), **self.oc_kw)
        self.rho = new
        torch.cuda.synchronize(dev)
        t3 = time.perf_counter()
        self.iter += 1
        rec = {"compliance": c.item(), "change": change, "iterations": self.fem.last_info["iterations"],
               "solve_s": self.fem.last_info["solve_seconds"], "fem_blocks": int(blocks.numel()),
               "fat_blocks": int(band.fat.sum()), "thin_blocks": int(band.thin.sum()), "rebuilt": rebuilt,
               "band_s": t1 - t0, "band_update_s": tb - t0, "rebuild_s": t1 - tb, "fem_s": t2 - t1,
               "filter_oc_s": t3 - t2}
        self.history.append(rec)
        return rec
