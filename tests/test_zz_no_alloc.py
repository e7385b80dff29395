"""GPU: the allocation-free l1-ball entry point (``dtv_l1ball_noalloc``, its
state in a static device array) against ``dtv_l1ball`` with a caller
workspace, through the C ABI: two calls in a row, the second on fewer and
smaller vectors, each bitwise equal (same kernels, fixed-order reductions)."""
import ctypes

import pytest
import torch

from dtvproj import _capi, phantom, tv

pytestmark = pytest.mark.gpu


def _items(vs, outs, radii, sigma):
    items = (tv.L1Item * len(vs))()
    for k, (v, o, b) in enumerate(zip(vs, outs, radii)):
        items[k].v = v.data_ptr()
        items[k].out = o.data_ptr()
        items[k].n = v.numel()
        items[k].radius = float(b)
        items[k].sigma = float(sigma)
    return items


@pytest.mark.parametrize("mode", [1, 0])
def test_noalloc_l1ball_matches_the_workspace_entry_point(cuda, mode):
    print(f"[case] dtv_l1ball_noalloc mode={mode}", flush=True)
    L = _capi.lib()
    sigma = 1e-3 if mode == 1 else 1.0
    calls = [
        [phantom.gaussian((300000,), 71), phantom.gaussian((5000,), 72),
         phantom.gaussian((77,), 73)],
        [phantom.gaussian((1000,), 74)],
    ]
    stream = torch.cuda.current_stream(cuda).cuda_stream
    for vs in calls:
        vs = [(v * sigma).to(cuda) for v in vs]
        radii = [0.3 * float((v.double().abs() / sigma).sum()) for v in vs]
        ref = [torch.empty_like(v) for v in vs]
        tv._l1ball_device(vs, radii, ref, mode, sigma=sigma)
        got = [torch.full_like(v, float("nan")) for v in vs]
        items = _items(vs, got, radii, sigma)
        rc = L.dtv_l1ball_noalloc(ctypes.addressof(items), len(vs), mode, 1e-4, 100, stream)
        assert rc == 0
        torch.cuda.synchronize()
        for g, r in zip(got, ref):
            assert torch.equal(g, r)
    # beyond the scratch array's capacity: a clear error, nothing launched
    many = [torch.zeros(4, device=cuda) for _ in range(17)]
    items = _items(many, many, [1.0] * 17, 1.0)
    assert L.dtv_l1ball_noalloc(ctypes.addressof(items), 17, 0, 1e-4, 100, stream) == 3
    assert "large" in L.dtv_strerror(3).decode().lower()
