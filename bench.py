"""Benchmark: A + matched A^T throughput on one MI355X.

Workload (default): 256^3 FP32 volume at 0.5 mm, 60 views over 360 deg,
384x384 detector at 0.8 mm, SOD 600 mm, SDD 1100 mm.  One step = forward
projection of a seeded 20-ellipsoid phantom followed by the matched adjoint
of the resulting sinogram, inputs resident in HBM.  Metric = detector rays
per second of the pair (GRays/s), with the dominant kernel's roofline against
the resource that binds it (vector-instruction issue or HBM bandwidth).

Prints ONE JSON line.  ``--gpus N`` with N > 1 is refused: this build runs
on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
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


def cpu_baseline(views_sample, log, reps=3):
    """FP32 CPU restatement (the oracle in FP32) on a bounded sample of the
    same workload: ``views_sample`` of the 60 views for A and for A^T, the
    median of ``reps`` runs.  Views are independent, so rays/s on the sample
    is the workload's rate."""
    import torch
    from oracle import cone
    from dtvproj import geometry as G, phantom
    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    ang = G.uniform_angles(60)[views_sample]
    vol = phantom.phantom((256, 256, 256), 0.5, 20, seed=2)
    args = ((256, 256, 256), 0.5, (384, 384), 0.8, 600.0, 1100.0)
    runs = []
    for _ in range(max(1, reps)):
        t0 = time.perf_counter()
        sino = cone.circ_forward(vol, ang, *args, dtype=torch.float32)
        t1 = time.perf_counter()
        cone.circ_adjoint(sino, ang, *args, dtype=torch.float32)
        t2 = time.perf_counter()
        runs.append((t2 - t0, t1 - t0, t2 - t1))
    runs.sort()
    pair, fwd, adj = runs[len(runs) // 2]
    rays = len(views_sample) * 384 * 384
    log(f"cpu baseline: A {fwd:.2f}s  A^T {adj:.2f}s on {len(views_sample)} view(s), "
        f"{threads} threads, median of {len(runs)}")
    return dict(value=rays / pair / 1e9, unit="Grays/s", cores=threads, kind="port",
                sample=f"{len(views_sample)} of 60 views (A then matched A^T; views "
                       f"{list(views_sample)}), median of {len(runs)} runs, FP32 oracle "
                       f"restatement, torch CPU, {_cpu_model()}",
                seconds_forward=fwd, seconds_adjoint=adj)


def batched_unroll(log):
    """Second metric: 20-iteration FISTA-TV over 16 uneven instances (seeded
    sizes 64..256 x 64..160, 30..120 views), batched launches captured in one
    hipGraph vs the plain per-volume loop, both with the same step sizes."""
    import torch
    from dtvproj import batched, configs, ops, phantom, unroll
    dev = torch.device("cuda", 0)
    geoms = configs.uneven_16()
    b = []
    for i, g in enumerate(geoms):
        ph = phantom.phantom(g.volume_size, 1.0, 10, seed=100 + i).to(dev)
        b.append(ops.forward(ph, g))
    lam, eps, iters = 1e-3, 1e-3, 20
    steps = [1.0 / (unroll.power_norm(g, iters=8, device=dev) + lam * 12.0 / eps)
             for g in geoms]

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n

    t_plain = timed(lambda: batched.plain_fista_tv(geoms, b, iters, lam, eps, steps), 1)
    # one A and one A^T brick shape for the whole launch, picked by the
    # multi-launch autotuner (before the timed region)
    bf = batched.BatchedFistaTV(geoms, b, iters, lam, eps, steps, bricks="tune_shared")
    bf.capture()
    t_batched = timed(bf.replay, 3)
    plain = batched.plain_fista_tv(geoms, b, iters, lam, eps, steps)
    out = bf.replay()
    torch.cuda.synchronize()
    err = max(float((o - p).norm() / p.norm().clamp_min(1e-30)) for o, p in zip(out, plain))
    log(f"batched unroll: plain {t_plain:.3f} s/step, batched+graph {t_batched:.3f} s/step, "
        f"ratio {t_plain / t_batched:.2f}, max rel diff {err:.2e}")
    return dict(metric="batched vs plain 20-iteration FISTA-TV unroll, 16 uneven instances",
                plain_steps_per_s=1.0 / t_plain, batched_steps_per_s=1.0 / t_batched,
                ratio=t_plain / t_batched, max_rel_diff=err, instances=len(geoms),
                iterations=iters, tuned_bricks=dict(forward=list(bf.mp.tuned["forward"]),
                                                    adjoint=list(bf.mp.tuned["adjoint"])))


def ref_backprojector(sino, steps, log):
    """Reference backprojector B_ref (Standard scheme, backwardKernel.cu:22-53) at the same
    sizes: 60 views, 384x384 detector, 256^3 volume, SID 750 / SDD 1375 pixel units (the
    config's 600 / 1100 mm at a 0.8 mm pixel).  Voxel-driven, 16 B per voxel-view."""
    import torch
    from dtvproj import geometry as G, ops
    g = G.standard(G.uniform_angles(60), (256, 256, 256), (384, 384), 750.0, 1375.0)
    ops.bp_ref(sino, g)
    stream = torch.cuda.current_stream(sino.device)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        ops.bp_ref(sino, g)
    e1.record(stream)
    e1.synchronize()
    t = e0.elapsed_time(e1) / steps
    vv = 256 ** 3 * 60
    log(f"reference backprojector: {t:.3f} ms ({vv * 16 / t / 1e6:.0f} GB/s algorithmic)")
    return dict(ms=t, voxel_views=vv, algorithmic_gbs=vv * 16 / t / 1e6,
                geometry="Standard scheme, 60 views, 384x384, 256^3, SID 750 / SDD 1375 px")


def beijing_loop(steps, log):
    """The projector pair of the reference's Beijing unroll iteration
    (iterations/iteration.py:37-40): volume 256x256x64, 144 source angles x 21
    sub-detectors = 3024 views of 80x128, SID 708 / SDD 1143 (options.yaml).
    The scan's projection_matrix file does not ship, so the view vectors are the
    synthetic 21-plane layout with the same shapes.  Times A + B_ref (the pair
    the reference's iteration runs) and A + matched A^T, phantom input."""
    import numpy as np
    import torch
    from dtvproj import geometry as G, ops, phantom
    X, Y, Z, U, V = 256, 256, 64, 80, 128
    mag = 1143.0 / 708.0
    pitch = (1.2 * np.hypot(X, Y) * mag / (U * 21), 1.2 * Z * mag / V)
    g = G.beijing(G.beijing_default_vectors(144, 21, (U, V), 708.0, 1143.0, pitch),
                  (X, Y, Z), (U, V))
    dev = torch.device("cuda", 0)
    x = phantom.phantom((X, Y, Z), 1.0, 20, seed=2).to(dev)
    y = ops.forward(x, g)
    stream = torch.cuda.current_stream(dev)

    def timed(fn):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / steps

    ta = timed(lambda: ops.forward(x, g))
    tb = timed(lambda: ops.bp_ref(y, g))
    tt = timed(lambda: ops.adjoint(y, g))
    rays = g.n_views * U * V
    ref_ms = 168.0  # recorded: 400 iterations in about 67 s
    log(f"beijing loop pair: A {ta:.2f} ms, B_ref {tb:.2f} ms, matched A^T {tt:.2f} ms; "
        f"A + B_ref {ta + tb:.2f} ms vs the reference's ~{ref_ms:.0f} ms per iteration")
    return dict(forward_ms=ta, bp_ref_ms=tb, adjoint_ms=tt, pair_ref_ms=ta + tb,
                pair_matched_ms=ta + tt, rays=rays,
                grays_per_s_ref_pair=rays / (ta + tb) / 1e6,
                reference_iteration_ms=ref_ms,
                speedup_vs_reference_iteration=ref_ms / (ta + tb),
                geometry="Beijing, 256x256x64, 3024 views (144 x 21) of 80x128, "
                         "synthetic view vectors",
                note="the reference's 168 ms covers its whole iteration (A, B_ref and "
                     "elementwise work) on an unstated NVIDIA GPU")


def dense_inputs(geom, sino, acct, reps, log):
    """A on an i.i.d. N(0,1) volume (seed 20) and the matched A^T on an
    i.i.d. N(0,1) sinogram (seed 21), medians over ``reps`` event-timed
    launches.  The matched A^T drops rays whose value is zero, so on the
    phantom's projection (``sino``, where rays that miss every ellipsoid
    are zero) it does less than the algorithmic count; every ray of the
    N(0,1) sinogram is nonzero."""
    import torch
    from dtvproj import ops, phantom
    dev = sino.device
    x = phantom.gaussian(geom.volume_shape(1), 20).to(dev)
    y = phantom.gaussian(geom.sino_shape(1), 21).to(dev)
    stream = torch.cuda.current_stream(dev)

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)

    fwd = med(lambda: ops.forward(x, geom))
    adj = med(lambda: ops.adjoint(y, geom))
    nz = float((sino != 0).double().mean())
    log(f"dense inputs: forward {fwd:.3f} ms  adjoint {adj:.3f} ms; phantom sinogram "
        f"nonzero rays {nz:.3f}")
    return dict(forward_ms_median=fwd, adjoint_ms_median=adj,
                pair_grays=acct["rays"] / ((fwd + adj) * 1e-3) / 1e9,
                adjoint_achieved=acct["adjoint"] / (adj * 1e-3) / 1e9,
                phantom_sino_nonzero_ray_frac=nz, reps=reps,
                inputs="N(0,1) volume seed 20, N(0,1) sinogram seed 21")


def pdhg_512(iters, log):
    """Inference of the learned directional-TV PDHG unroll at 512x512x384 /
    120 views / 768x576 (40-ellipsoid phantom, Poisson I0 = 5e4 sinogram,
    weights 2e-3 / 2e-3 / 4e-3 for x / y / z): one warm-up pass, then one
    timed pass of ``iters`` iterations."""
    import time
    import torch
    from dtvproj import configs, ops, phantom, unroll
    dev = torch.device("cuda", 0)
    g = configs.large_512()
    truth = (phantom.phantom(g.volume_size, 0.4, 40, seed=5) * 0.02).to(dev)
    with torch.no_grad():
        y = phantom.poisson_noise(ops.forward(truth, g), 5e4, seed=8)
        norm = unroll.power_norm(g, iters=6, device=dev)
        net = unroll.PDHGDTVNet(g, iters=iters, lam=(2e-3, 2e-3, 4e-3), norm=norm).to(dev)
        net(y)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = net(y)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
    err = float((f - truth).norm() / truth.norm())
    log(f"pdhg/dtv 512: {iters} iterations in {t:.3f} s ({1e3 * t / iters:.1f} ms per "
        f"iteration), rel error {err:.4f}")
    return dict(iterations=iters, seconds=t, ms_per_iteration=1e3 * t / iters, rel_error=err,
                geometry="512x512x384 @0.4mm, 120 views, 768x576 @0.6mm, SOD 700, SDD 1200")


def volume_steps(reps, log):
    """The fused volume-update kernels on their own, event-timed over ``reps``
    launches after one warm-up: the FISTA-TV step (20 B of algorithmic HBM
    traffic per voxel: y, g, xprev read, x and y' written), the ASD-POCS TV
    descent step (8 B: f read, f' written) in both boundary modes, and the
    PDHG/DTV dual + primal pair (56 B: the dual reads f-bar and p and writes
    p; the primal reads f, A^T w and p and writes f and f-bar) at 256^3 and
    512x512x384 on uniform [0, 1) inputs; then one FISTA-TV step over the 16
    uneven instances of the batched unroll (one launch)."""
    import torch
    from dtvproj import batched, configs, tv
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    gen = torch.Generator(device=dev).manual_seed(40)

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def rec(name, shape, mode, ms, per_voxel, voxels):
        tbs = per_voxel * voxels / ms / 1e9
        log(f"{name} {shape} {mode}: {ms:.3f} ms, {tbs:.2f} TB/s ({per_voxel} B/voxel)")
        return dict(kernel=name, shape=list(shape), mode=mode, ms=ms, bytes_per_voxel=per_voxel,
                    voxels=voxels, achieved_tbs=tbs, frac=tbs * 1e3 / 8000.0)

    out = []
    for shape in ((256, 256, 256), (384, 512, 512)):
        full = (1, 1) + shape
        n = shape[0] * shape[1] * shape[2]
        u = lambda: torch.rand(full, device=dev, generator=gen)  # noqa: E731
        y, g, xp = u(), u(), u()
        xn, yn, fo = torch.empty_like(y), torch.empty_like(y), torch.empty_like(y)
        for mode in ("neumann", "reference"):
            t = timed(lambda: tv.fista_tv_step(y, g, xp, 0.1, 0.01, 0.01, 0.5, True, mode,
                                               xnew=xn, ynew=yn))
            out.append(rec("fista_tv_step", shape, mode, t, 20, n))
            t = timed(lambda: tv.tv_descent(y, 0.1, 1.0, mode, out=fo))
            out.append(rec("tv_descent", shape, mode, t, 8, n))
        del y, g, xp, xn, yn, fo
        f, fb, atw = u(), u(), u()
        p = torch.rand((3,) + full, device=dev, generator=gen)
        for mode in ("neumann", "reference"):
            t = timed(lambda: tv.pdhg_dtv_step(f, fb, atw, p, 0.1, 0.1, (0.3, 0.3, 0.3), mode))
            out.append(rec("pdhg_dtv_step", shape, mode, t, 56, n))
        del f, fb, atw, p
    geoms = configs.uneven_16()
    vox = sum(g.volume_size[0] * g.volume_size[1] * g.volume_size[2] for g in geoms)
    sinos = [torch.zeros(g.sino_shape(1), device=dev) for g in geoms]
    bf = batched.BatchedFistaTV(geoms, sinos, iters=2, steps=[0.1] * len(geoms))
    for t_ in bf.y[0] + bf.g + bf.x:
        t_.uniform_(generator=gen)
    t = timed(lambda: bf.tv_step(0))
    out.append(rec("fista_tv_step_multi", ("uneven_16",), "neumann", t, 20, vox))
    return out


def workloads(reps, log):
    """A and matched A^T with the default kernels on the other workload sizes
    (seeded ellipsoid phantoms; A^T on the phantom's own projection), medians
    of ``reps`` event-timed calls: 64^3 / 72 views, 4 x 128^3 / 90 views,
    256^3 / 60 views, 512x512x384 / 120 views, and the reference's Standard
    options geometry (256x256x64, 48 views of 288x28, SID 1000 / SDD 2000
    pixel units)."""
    import torch
    from dtvproj import configs, geometry as G, ops, phantom, roofline
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    cases = [("64^3 @1mm, 72 views, 128^2", configs.fista_64(), 1, 1.0, 10),
             ("4 x 128^3 @1mm, 90 views, 192^2", configs.unroll_128(), 4, 1.0, 10),
             ("256^3 @0.5mm, 60 views, 384^2", configs.sparse_256(), 1, 0.5, 20),
             ("512x512x384 @0.4mm, 120 views, 768x576", configs.large_512(), 1, 0.4, 40),
             ("Standard options 256x256x64, 48 views, 288x28",
              G.standard(G.uniform_angles(48), (256, 256, 64), (288, 28), 1000.0, 2000.0), 1,
              1.0, 20)]

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)

    out = []
    for name, g, batch, vx, nell in cases:
        vol = torch.cat([phantom.phantom(g.volume_size, vx, nell, seed=2 + b)
                         for b in range(batch)]).to(dev)
        sino = ops.forward(vol, g)
        ta = med(lambda: ops.forward(vol, g))
        tt = med(lambda: ops.adjoint(sino, g))
        acct = roofline.pair_bytes(g, batch=batch)
        rec = dict(workload=name, batch=batch, forward_ms=ta, adjoint_ms=tt, pair_ms=ta + tt,
                   grays_per_s=acct["rays"] / ((ta + tt) * 1e-3) / 1e9,
                   tap_gbs=acct["pair"] / ((ta + tt) * 1e-3) / 1e9)
        log(f"{name} (batch {batch}): A {ta:.2f} ms, A^T {tt:.2f} ms, "
            f"{rec['grays_per_s']:.2f} Grays/s")
        out.append(rec)
        del vol, sino
    return out


PMC_FILE = os.path.join(ROOT, "profiles", "pmc_sparse_256.json")


def roofline_record(dom, dom_t, dom_b, acct, t_fwd, t_adj, med_fwd, med_adj, log):
    """Roofline of the dominant kernel against the resource that binds it.

    Per-launch counts come from the committed rocprofv3 PMC summary of this
    workload (``profiles/pmc_sparse_256.json``): HBM bytes (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a copy kernel of
    known size) and VALU wave-instructions (SQ_INSTS_VALU).  They are divided
    by this run's event-timed launch duration.  Each resource's fraction is
    against its physical peak: 8 TB/s HBM, and one wave64 VALU instruction per
    compute unit per cycle (256 CUs at 2.4 GHz).  ``bound`` names the resource
    with the larger fraction.  ``tap_gbs`` is the algorithmic figure (32 B per
    in-support sample: the 8 trilinear taps), which the LDS-tiled kernels
    serve from LDS, so it is not a bound and exceeds the HBM peak."""
    from dtvproj import roofline
    # the headline A^T runs the partitioned kernel (k_adj_part); older
    # summaries hold the unpartitioned one (k_adj_brick)
    keys = ("k_adj_part", "k_adj_brick") if dom.startswith("k_adj") else ("k_fwd_brick",)
    rec = {}
    if os.path.exists(PMC_FILE):
        with open(PMC_FILE) as f:
            pmc = json.load(f)
        rec = next((pmc[k] for k in keys if k in pmc), {})
    traffic = rec.get("traffic_bytes")
    valu = rec.get("sq_insts_valu")
    conflict = rec.get("sq_lds_bank_conflict")
    cu_cycles = roofline.VALU_PEAK_WAVE_INSTS * dom_t
    hbm = None
    if traffic:
        g = traffic / dom_t / 1e9
        hbm = dict(achieved=g, peak=roofline.HBM_PEAK_GBS, unit="GB/s",
                   frac=g / roofline.HBM_PEAK_GBS)
    issue = None
    if valu:
        a = valu / dom_t / 1e9
        issue = dict(achieved=a, peak=roofline.VALU_PEAK_WAVE_INSTS / 1e9,
                     unit="G VALU wave64 instructions/s", frac=valu / cu_cycles)
    cands = [(n, r) for n, r in (("valu", issue), ("hbm", hbm)) if r is not None]
    if not cands:
        log("no PMC summary for the dominant kernel: roofline fractions unavailable")
        bound, best = "hbm", dict(achieved=None, peak=roofline.HBM_PEAK_GBS, unit="GB/s",
                                  frac=None)
    else:
        bound, best = max(cands, key=lambda c: c[1]["frac"])
    return dict(bound=bound, achieved=best["achieved"], peak=best["peak"], unit=best["unit"],
                frac=best["frac"], traffic=traffic, kernel=dom,
                kernel_template=rec.get("kernel"),
                hbm=hbm, valu_issue=issue,
                lds_bank_conflict_frac=conflict / cu_cycles if conflict else None,
                l2_hit_rate=rec.get("l2_hit_rate"),
                tap_gbs=dom_b / dom_t / 1e9,
                # each array read or written once: volume + sinogram, FP32
                compulsory_bytes=4 * (256 ** 3 + acct["rays"]),
                forward_ms=t_fwd * 1e3, adjoint_ms=t_adj * 1e3,
                forward_ms_median=med_fwd * 1e3, adjoint_ms_median=med_adj * 1e3,
                pair_tap_gbs=acct["pair"] / (t_fwd + t_adj) / 1e9,
                algorithmic_bytes=dict(forward=acct["forward"], adjoint=acct["adjoint"],
                                       samples=acct["samples"], per_sample=32),
                pmc_source="profiles/pmc_sparse_256.json")


LEGS = ("beijing", "batched", "pdhg512", "dense", "volume_steps", "workloads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=3)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-beijing", action="store_true")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense N(0,1)-input leg (profiles of the headline pair alone)")
    ap.add_argument("--leg", choices=LEGS, default=None,
                    help="run only this secondary workload (--steps = its repetitions or "
                         "iterations) and print its result as JSON, for profiling it alone")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="run one device copy of this size first (rocprof counter scale)")
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
    if args.leg is not None:
        def dense_leg():
            g = configs.sparse_256()
            vol = phantom.phantom((256, 256, 256), 0.5, 20, seed=2).to(dev)
            return dense_inputs(g, ops.forward(vol, g), roofline.pair_bytes(g),
                                max(1, args.steps), log)

        res = {"beijing": lambda: beijing_loop(max(1, args.steps), log),
               "batched": lambda: batched_unroll(log),
               "pdhg512": lambda: pdhg_512(max(1, args.steps), log),
               "dense": dense_leg,
               "volume_steps": lambda: volume_steps(max(1, args.steps), log),
               "workloads": lambda: workloads(max(1, args.steps), log)}[args.leg]()
        print(json.dumps({"leg": args.leg, "result": res}), flush=True)
        return
    _capi.lib()
    geom = configs.sparse_256()
    acct = roofline.pair_bytes(geom)
    log(f"workload sparse_256: samples={acct['samples']} rays={acct['rays']}")

    vol = phantom.phantom((256, 256, 256), 0.5, 20, seed=2).to(dev)
    torch.cuda.synchronize()
    if args.calibration_copy_mb:
        n = args.calibration_copy_mb * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {4 * n} bytes read, {4 * n} bytes written")
        del src, dst

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
    med_fwd = statistics.median(e0.elapsed_time(e1) for e0, e1, _ in ev) * 1e-3
    med_adj = statistics.median(e1.elapsed_time(e2) for _, e1, e2 in ev) * 1e-3
    ms_per_step = elapsed / args.steps * 1e3
    value = acct["rays"] * args.steps / elapsed / 1e9
    log(f"forward {t_fwd * 1e3:.3f} ms  adjoint {t_adj * 1e3:.3f} ms  step {ms_per_step:.3f} ms")

    # dominant kernel = the slower of the two
    if t_adj >= t_fwd:
        dom, dom_t, dom_b = "k_adj (matched A^T)", t_adj, acct["adjoint"]
    else:
        dom, dom_t, dom_b = "k_fwd (A)", t_fwd, acct["forward"]
    roof = roofline_record(dom, dom_t, dom_b, acct, t_fwd, t_adj, med_fwd, med_adj, log)

    dense = None if args.no_dense else dense_inputs(geom, s, acct, max(3, min(args.steps, 20)),
                                                    log)

    bref = ref_backprojector(s, max(1, min(args.steps, 10)), log)
    bj = None if args.no_beijing else beijing_loop(max(1, min(args.steps, 5)), log)

    cpu = None
    if not args.no_cpu_baseline:
        views = list(range(0, 60, max(1, 60 // max(1, args.cpu_views))))[:args.cpu_views]
        cpu = cpu_baseline(views, log, args.cpu_reps)

    bat = None if args.no_batched else batched_unroll(log)

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
        "dense_inputs": dense,
        "cpu_baseline": cpu,
        "reference_backprojector": bref,
        "beijing_loop": bj,
        "batched_unroll": bat,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
