"""spg_oc_trial_noalloc (the OC trial whose partials live in a fixed device
array instead of allocated memory) against spg_oc_trial, through the C ABI.

Both run the same kernels on the same grid, so volume, change and the new
densities must agree bit for bit.  Two calls in a row, the second on a
smaller input, check that the array's leftover partials from the first call
do not leak into the second.
"""
import pytest

torch = pytest.importorskip("torch")

from spgrid_topo_opt_amd import _lib  # noqa: E402

pytestmark = pytest.mark.gpu


def _case(n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    rho = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 0.8 + 0.1
    dc = -torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    r = torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    mask = (r < 0.85).to(torch.uint8) + 2 * ((r < 0.05).to(torch.uint8))  # inside; some fixed
    return mask.contiguous(), rho.contiguous(), dc.contiguous()


def _both(n, seed, mid, **kw):
    mask, rho, dc = _case(n, seed)
    out_ref = torch.empty_like(rho)
    vol, ch = _lib.oc_trial(n, mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), mid, 0.2, 0.01,
                            out_ptr=out_ref.data_ptr(), **kw)
    out = torch.full_like(rho, -1.0)
    res = torch.full((2,), -1.0, device="cuda", dtype=torch.float64)
    stream = torch.cuda.current_stream().cuda_stream
    _lib.oc_trial_noalloc(n, mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), mid, 0.2, 0.01, res.data_ptr(),
                          out_ptr=out.data_ptr(), stream=stream, **kw)
    torch.cuda.synchronize()
    return (vol, ch, out_ref), (res[0].item(), res[1].item(), out)


@pytest.mark.parametrize("kw", [dict(), dict(exclude_fixed=False, fixed_density=0.7, dc_shift=-0.05)])
def test_noalloc_oc_trial_matches_allocating_entry_point(kw):
    # 400,000 cells need more workgroups than the array holds partials for
    # (grid-stride loop); 5,000 cells use 20 of them on the second call
    for n, seed, mid in ((400_000, 1, 0.3), (5_000, 2, 0.05)):
        (v0, c0, o0), (v1, c1, o1) = _both(n, seed, mid, **kw)
        print(f"[no-alloc] n={n} volume {v0!r} vs {v1!r}, change {c0!r} vs {c1!r}", flush=True)
        assert v1 == v0 and c1 == c0
        assert torch.equal(o0, o1)


def test_noalloc_oc_trial_without_output_and_errors():
    mask, rho, dc = _case(3_000, 4)
    vol, ch = _lib.oc_trial(3_000, mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), 0.1, 0.2, 0.01)
    res = torch.zeros(2, device="cuda", dtype=torch.float64)
    _lib.oc_trial_noalloc(3_000, mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), 0.1, 0.2, 0.01, res.data_ptr())
    torch.cuda.synchronize()
    assert res[0].item() == vol and res[1].item() == ch
    # empty input: zero volume and change
    _lib.oc_trial_noalloc(0, mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), 0.1, 0.2, 0.01, res.data_ptr())
    torch.cuda.synchronize()
    assert res.tolist() == [0.0, 0.0]
    with pytest.raises(_lib.SpgError, match="negative cell count"):
        _lib.oc_trial_noalloc(-1, mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), 0.1, 0.2, 0.01, res.data_ptr())
    with pytest.raises(_lib.SpgError, match="multiplier"):
        _lib.oc_trial_noalloc(10, mask.data_ptr(), rho.data_ptr(), dc.data_ptr(), 0.0, 0.2, 0.01, res.data_ptr())
