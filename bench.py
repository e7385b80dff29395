"""Benchmark: A + matched A^T throughput on one MI355X.

Workload (default): 256^3 FP32 volume at 0.5 mm, 60 views over 360 deg,
384x384 detector at 0.8 mm, SOD 600 mm, SDD 1100 mm.  One step = forward
projection of a seeded 20-ellipsoid phantom followed by the matched adjoint
of the resulting sinogram, inputs resident in HBM.  Metric = detector rays
per second of the pair (GRays/s) with the dominant kernel's fraction of the
HBM roofline (algorithmic bytes / kernel time / 8 TB/s).

Prints ONE JSON line.  ``--gpus N`` with N > 1 is refused: this build runs
on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dtv-net_amd"))
sys.path.insert(0, ROOT)

METRIC = "Grays/s A+A^T at 256^3/60 views (%HBM roofline)"


def _cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(views_sample, log):
    """FP32 CPU restatement (the oracle in FP32) on a bounded sample of the
    same workload: ``views_sample`` of the 60 views for A and for A^T.
    Views are independent, so rays/s on the sample is the workload's rate."""
    import torch
    from oracle import cone
    from dtvproj import geometry as G, phantom
    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    ang = G.uniform_angles(60)[views_sample]
    vol = phantom.phantom((256, 256, 256), 0.5, 20, seed=2)
    args = ((256, 256, 256), 0.5, (384, 384), 0.8, 600.0, 1100.0)
    t0 = time.perf_counter()
    sino = cone.circ_forward(vol, ang, *args, dtype=torch.float32)
    t1 = time.perf_counter()
    cone.circ_adjoint(sino, ang, *args, dtype=torch.float32)
    t2 = time.perf_counter()
    rays = len(views_sample) * 384 * 384
    log(f"cpu baseline: A {t1 - t0:.2f}s  A^T {t2 - t1:.2f}s on {len(views_sample)} view(s), "
        f"{threads} threads")
    return dict(value=rays / (t2 - t0) / 1e9, unit="Grays/s", cores=threads, kind="port",
                sample=f"{len(views_sample)} of 60 views (A then matched A^T), FP32 oracle "
                       f"restatement, torch CPU, {_cpu_model()}",
                seconds_forward=t1 - t0, seconds_adjoint=t2 - t1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=1)
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the projector "
              "pair is measured on one GPU's share and multi-GPU sharding is not part of "
              "this build", file=sys.stderr)
        sys.exit(2)

    import torch
    from dtvproj import configs, ops, phantom, roofline, _capi

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda", 0)
    _capi.lib()
    geom = configs.sparse_256()
    acct = roofline.pair_bytes(geom)
    log(f"workload sparse_256: samples={acct['samples']} rays={acct['rays']}")

    vol = phantom.phantom((256, 256, 256), 0.5, 20, seed=2).to(dev)
    torch.cuda.synchronize()

    # warmup
    for _ in range(max(args.warmup, 1)):
        s = ops.forward(vol, geom)
        ops.adjoint(s, geom)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        e0, e1, e2 = ev[i]
        e0.record(stream)
        s = ops.forward(vol, geom)
        e1.record(stream)
        ops.adjoint(s, geom)
        e2.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    t_fwd = sum(e0.elapsed_time(e1) for e0, e1, _ in ev) / args.steps * 1e-3
    t_adj = sum(e1.elapsed_time(e2) for _, e1, e2 in ev) / args.steps * 1e-3
    ms_per_step = elapsed / args.steps * 1e3
    value = acct["rays"] * args.steps / elapsed / 1e9
    log(f"forward {t_fwd * 1e3:.3f} ms  adjoint {t_adj * 1e3:.3f} ms  step {ms_per_step:.3f} ms")

    # dominant kernel = the slower of the two
    if t_adj >= t_fwd:
        dom, dom_t, dom_b = "k_adj (matched A^T)", t_adj, acct["adjoint"]
    else:
        dom, dom_t, dom_b = "k_fwd (A)", t_fwd, acct["forward"]
    achieved = dom_b / dom_t / 1e9
    pair_gbs = acct["pair"] / (t_fwd + t_adj) / 1e9
    roof = dict(bound="hbm", achieved=achieved, peak=roofline.HBM_PEAK_GBS, unit="GB/s",
                frac=achieved / roofline.HBM_PEAK_GBS, traffic=None, kernel=dom,
                pair_achieved=pair_gbs, pair_frac=pair_gbs / roofline.HBM_PEAK_GBS,
                forward_ms=t_fwd * 1e3, adjoint_ms=t_adj * 1e3,
                algorithmic_bytes=dict(forward=acct["forward"], adjoint=acct["adjoint"],
                                       samples=acct["samples"]))

    cpu = None
    if not args.no_cpu_baseline:
        views = list(range(0, 60, max(1, 60 // max(1, args.cpu_views))))[:args.cpu_views]
        cpu = cpu_baseline(views, log)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 20-ellipsoid phantom, seed 2)",
        "config": {"workload": "cone-beam A + matched A^T, 256^3 @0.5mm, 60 views, "
                               "384x384 @0.8mm, SOD 600, SDD 1100",
                   "batch": 1, "views": 60, "volume": [256, 256, 256], "detector": [384, 384]},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
