"""Plain-torch references of the optimizer-side HIP kernels (test use only):
the sensitivity filter (topo_opt/spgrid_topo_opt.cpp:527-574) and the
optimality-criterion bisection (topo_opt/optimizer/optimality_criterion.cpp:
19-146), written with tensor ops over the same blocks.  Pinned by closed-form
cases in tests/test_torch_reference_pinning.py."""
import numpy as np
import torch


def halo(grid, v, width=1):
    """(nb*64,) cell values -> (nb, 4+2w, 4+2w, 4+2w) with neighbour halos (0 outside)."""
    w = width
    v = v.reshape(grid.nb, 4, 4, 4)
    padded = torch.cat([v.new_zeros(1, 4, 4, 4), v], 0)  # index 0 = absent block
    n = 4 + 2 * w
    out = v.new_zeros(grid.nb, n, n, n)
    for s in range(27):
        o = (s // 9 - 1, (s // 3) % 3 - 1, s % 3 - 1)
        src = padded[grid.nbr[:, s] + 1]
        sl_dst, sl_src = [], []
        for d in range(3):
            if o[d] < 0:
                sl_dst.append(slice(0, w))
                sl_src.append(slice(4 - w, 4))
            elif o[d] > 0:
                sl_dst.append(slice(4 + w, 4 + 2 * w))
                sl_src.append(slice(0, w))
            else:
                sl_dst.append(slice(w, 4 + w))
                sl_src.append(slice(0, 4))
        out[:, sl_dst[0], sl_dst[1], sl_dst[2]] = src[:, sl_src[0], sl_src[1], sl_src[2]]
    return out


def filter_sensitivity(grid, dc, rho, filter_size=1.5):
    h = int(np.floor(filter_size))
    inside = grid.inside.to(torch.float64)
    hv = halo(grid, dc * rho * inside, h)
    hm = halo(grid, inside, h)
    num = torch.zeros(grid.nb, 4, 4, 4, dtype=torch.float64, device=dc.device)
    den = torch.zeros_like(num)
    for a in range(-h, h + 1):
        for b in range(-h, h + 1):
            for c in range(-h, h + 1):
                w = max(0.0, filter_size - float(np.sqrt(a * a + b * b + c * c)))
                if w == 0.0:
                    continue
                sl = (slice(None), slice(h + a, h + a + 4), slice(h + b, h + b + 4), slice(h + c, h + c + 4))
                num += w * hv[sl]
                den += w * hm[sl]
    den = den.reshape(-1)
    out = num.reshape(-1) / (torch.where(den > 0, den, torch.ones_like(den)) * (rho + 1e-3))
    return torch.where(den > 0, out, torch.zeros_like(out))


def oc_update(rho, dc, inside, volume_fraction, step_limit=0.2, min_density=0.0, container_cells=None, fixed=None,
              fixed_cell_density=1.0, exclude_fixed_cells=True, smart_oc=False):
    n = inside.sum().item() if container_cells is None else container_cells
    fixed = torch.zeros_like(inside) if fixed is None else (fixed & inside)
    if exclude_fixed_cells:
        n -= fixed.sum().item()
    target = volume_fraction * n
    shift = min(0.0, (-dc[inside]).min().item()) if smart_oc else 0.0
    neg = torch.clamp(-dc - shift, min=0.0)
    counted = inside & ~fixed if exclude_fixed_cells else inside

    def trial(mid):
        new = torch.clamp(rho * torch.sqrt(neg / mid), rho - step_limit, rho + step_limit)
        new = torch.clamp(new, min_density, 1.0)
        new = torch.where(fixed, torch.full_like(new, fixed_cell_density), new)
        return torch.where(inside, new, rho)

    lower, upper = 1e-20, 1e20
    while lower * (1 + 1e-7) < upper:
        mid = 0.5 * (lower + upper)
        if trial(mid)[counted].sum().item() > target:
            lower = mid
        else:
            upper = mid
    new = trial(mid)
    return new, (new - rho)[inside].abs().max().item(), mid
