"""Autotuner for the brick kernels' tile shape (brick edge lengths), ray
bundle size (lanes per workgroup) and view split (workgroups per brick).

For a geometry it times every candidate launch for A and for the matched
A^T on the device (median of ``reps`` timed calls after one warm-up, HIP
events on the current stream) and keeps the fastest.  Results are cached per
geometry signature in memory and, if ``path`` is given, in a JSON file.
All candidates compute the same result (to FP32 rounding), so tuning never
changes numerics beyond that.
"""
from __future__ import annotations

import json
import os
import threading

import torch

from . import ops
from .geometry import ConeGeometry

# (bx, by, bz, lanes); the adjoint needs bx*by*bz <= 32 * lanes.  A is timed
# on each shape once, at FORWARD_LANES: its brick kernel is built for
# 512-lane workgroups (six waves per SIMD), smaller ones run it below its
# register budget's occupancy.
FORWARD_LANES = 512
CANDIDATES = [
    (32, 16, 16, 256), (26, 26, 14, 256), (24, 24, 14, 256), (32, 32, 8, 256), (16, 16, 32, 256),
    (16, 16, 16, 256),
    (32, 16, 8, 256), (16, 16, 16, 128), (16, 16, 8, 128), (8, 8, 8, 64),
]

_cache = {}
_lock = threading.Lock()


def signature(geom: ConeGeometry) -> str:
    sc = geom.scheme
    return json.dumps([list(geom.volume_size), list(geom.detector_size), geom.n_views,
                       sc.axis_mode, round(float(sc.margin), 6)])


# view splits tried per shape: A the library's default for the shape (0) and
# none; A^T the library's workspace split for the shape (0), one part and two
# parts adding into the output
FORWARD_PARTS = (0, 1)
ADJOINT_PARTS = (0, 1, 2)


def launch_of(c):
    """(bx, by, bz, lanes[, view_parts]) -> launch dict (view_parts 0 or
    absent: the library's default split for the shape)."""
    L = dict(algorithm=2, brick=tuple(c[:3]), block=c[3])
    if len(c) > 4 and c[4]:
        L["view_parts"] = int(c[4])
    return L


def _time(fn, reps):
    fn()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    return times[len(times) // 2]


def _load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tune(geom: ConeGeometry, device=None, reps: int = 3, candidates=CANDIDATES,
         path: str | None = None, verbose: bool = False) -> dict:
    """Fastest ``dict(forward=launch, adjoint=launch, times=...)`` for geom."""
    key = signature(geom)
    with _lock:
        if key in _cache:
            return _cache[key]
        disk = _load(path)
        if key in disk:
            r = disk[key]
            res = dict(forward=launch_of(r["forward"]), adjoint=launch_of(r["adjoint"]),
                       times=r.get("times"))
            _cache[key] = res
            return res
    device = device or torch.device("cuda", torch.cuda.current_device())
    gen = torch.Generator().manual_seed(0)
    x = torch.rand(geom.volume_shape(1), generator=gen).to(device)
    y = torch.rand(geom.sino_shape(1), generator=gen).to(device)
    times = {"forward": {}, "adjoint": {}}
    shapes = set()
    for c in candidates:
        if tuple(c[:3]) not in shapes:
            shapes.add(tuple(c[:3]))
            for p in FORWARD_PARTS:
                cf = tuple(c[:3]) + (FORWARD_LANES, p)
                L = launch_of(cf)
                times["forward"][str(cf)] = _time(lambda: ops.forward(x, geom, L), reps)
        # the adjoint's lanes own at most 32 voxels each
        if c[0] * c[1] * c[2] <= 32 * c[3]:
            for p in ADJOINT_PARTS:
                ca = tuple(c[:4]) + (p,)
                L = launch_of(ca)
                times["adjoint"][str(ca)] = _time(lambda: ops.adjoint(y, geom, L), reps)
        if verbose:
            print(f"tune {c}: " + " ".join(f"{k} {v:.3f}" for d in times.values()
                                           for k, v in d.items() if k.startswith(str(c[:3])[:-1])),
                  flush=True)
    best_f = min(times["forward"], key=times["forward"].get)
    best_a = min(times["adjoint"], key=times["adjoint"].get)
    bf = tuple(json.loads(best_f.replace("(", "[").replace(")", "]")))
    ba = tuple(json.loads(best_a.replace("(", "[").replace(")", "]")))
    res = dict(forward=launch_of(bf), adjoint=launch_of(ba), times=times)
    with _lock:
        _cache[key] = res
        if path:
            disk = _load(path)
            disk[key] = dict(forward=list(bf), adjoint=list(ba), times=times)
            with open(path, "w") as f:
                json.dump(disk, f, indent=1)
    return res


def tune_multi(geoms, device=None, reps: int = 3, candidates=CANDIDATES,
               verbose: bool = False) -> dict:
    """One brick shape for A and one for A^T shared by every instance of a
    multi-instance launch (a single compiled shape keeps the launch on its
    constant-shape kernel), each the fastest over the candidates that fit
    every instance, timed on the multi launch itself with uniform random
    inputs (dense, like unroll residuals)."""
    from . import batched
    geoms = list(geoms)
    device = device or torch.device("cuda", torch.cuda.current_device())
    gen = torch.Generator().manual_seed(0)
    xs = [torch.rand(g.volume_shape(1), generator=gen).to(device) for g in geoms]
    ys = [torch.rand(g.sino_shape(1), generator=gen).to(device) for g in geoms]
    outs_s = [torch.empty_like(y) for y in ys]
    outs_v = [torch.empty_like(x) for x in xs]
    times = {"forward": {}, "adjoint": {}}
    for c in candidates:
        if c[3] != 256:
            continue
        shape = tuple(c[:3])
        adj_ok = c[0] * c[1] * c[2] <= 32 * c[3]   # at most 32 owned voxels per lane
        adj_shape = shape if adj_ok else (32, 16, 16)
        mp = batched.MultiProjector(geoms, [(shape, adj_shape)] * len(geoms), device=device)
        bf = mp.bind_forward(xs, outs_s)
        times["forward"][str(shape)] = _time(lambda: mp.run(bf, False), reps)
        if adj_ok:
            ba = mp.bind_adjoint(ys, outs_v)
            times["adjoint"][str(shape)] = _time(lambda: mp.run(ba, True), reps)
        if verbose:
            print(f"tune_multi {shape}: fwd {times['forward'][str(shape)]:.3f} ms  "
                  f"adj {times['adjoint'].get(str(shape), float('nan')):.3f} ms", flush=True)
    parse = lambda k: tuple(json.loads(k.replace("(", "[").replace(")", "]")))
    best_f = parse(min(times["forward"], key=times["forward"].get))
    best_a = parse(min(times["adjoint"], key=times["adjoint"].get))
    return dict(forward=best_f, adjoint=best_a, times=times)
