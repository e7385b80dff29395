"""GPU: the single-call operators captured in a HIP graph (torch.cuda.graph)
and replayed several times with other GPU work and synchronisations between
the replays.  Every launch they make is a kernel (the zero fills included),
so each replay must reproduce the eager result: bitwise for the partitioned
matched A^T (one part stores, two add into zeros, three go through the
workspace and the reduce kernel) and for the fixed-point forward, to FP32
rounding for the FP32-atomic forward."""
import pytest
import torch

from conftest import rel_l2

from dtvproj import geometry as G, ops, phantom

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _geom():
    return G.circular(G.uniform_angles(13) + 7.0, (45, 38, 29), 0.9, (64, 40), 1.3, 80.0, 160.0)


def _replays(fn, eager, exact, reps=3):
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        fn()  # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = fn()
    for rep in range(reps):
        torch.cuda.synchronize()
        junk = torch.randn(4096, device=eager.device).sum()  # eager work between replays
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        if exact:
            assert torch.equal(out, eager), f"replay {rep}"
        else:
            assert rel_l2(out, eager) <= 1e-6, f"replay {rep}"
        float(junk)


@pytest.mark.parametrize("parts", [1, 2, 3])
def test_partitioned_adjoint_graph_replays(cuda, parts):
    g = _geom()
    y = phantom.gaussian(g.sino_shape(2), 71).to(cuda)
    L = dict(algorithm=2, brick=(16, 16, 8), block=256, view_parts=parts)
    _case("graph adjoint", parts=parts)
    eager = ops.adjoint(y, g, L)
    _replays(lambda: ops.adjoint(y, g, L), eager, exact=True)


@pytest.mark.parametrize("deterministic", [False, True])
def test_forward_graph_replays(cuda, deterministic):
    g = _geom()
    x = phantom.gaussian(g.volume_shape(2), 72).to(cuda)
    _case("graph forward", deterministic=deterministic)
    eager = ops.forward(x, g, deterministic=deterministic)
    _replays(lambda: ops.forward(x, g, deterministic=deterministic), eager, exact=deterministic)
