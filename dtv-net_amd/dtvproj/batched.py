"""Many volumes of uneven size and view count in one launch per operator.

:class:`MultiProjector` builds the instance and work tables of
``dtv_cone_forward_multi`` / ``dtv_cone_adjoint_multi`` (include/dtvproj.h)
for a list of geometries and bound buffers.  The work list holds every
(instance, brick) pair sorted by estimated cost (views x brick voxels),
largest first, so the hardware's in-order workgroup dispatch (or the
persistent queue) balances uneven instances.

:class:`BatchedFistaTV` runs the FISTA-TV unroll over all instances with
three launches per iteration -- A (initialised with -b, so it yields the
residual A y - b directly), matched A^T, and the fused TV / non-negativity /
momentum step over a descriptor table -- with every table in static device
memory, so :meth:`BatchedFistaTV.capture` can record the whole unroll in one
hipGraph (torch.cuda.CUDAGraph on ROCm) and :meth:`replay` re-runs it.
:func:`plain_fista_tv` is the per-volume loop it replaces.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi, autotune, ops, tv, unroll


class Instance(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p), ("init", ctypes.c_void_p),
        ("views", ctypes.c_void_p),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("nu", ctypes.c_int32), ("nv", ctypes.c_int32), ("n_views", ctypes.c_int32),
        ("axis_mode", ctypes.c_int32),
        ("bx", ctypes.c_int32), ("by", ctypes.c_int32), ("bz", ctypes.c_int32),
        ("nbx", ctypes.c_int32), ("nby", ctypes.c_int32), ("nbz", ctypes.c_int32),
        ("view_parts", ctypes.c_int32),
        ("box_lo", ctypes.c_double * 3), ("box_hi", ctypes.c_double * 3),
        ("margin", ctypes.c_double), ("scale", ctypes.c_double),
    ]


class WorkItem(ctypes.Structure):
    _fields_ = [("instance", ctypes.c_int32), ("brick", ctypes.c_int32)]


class _Bound:
    """Host + device copies of one instance table and its work list."""

    def __init__(self, host_inst, host_work, device):
        self.host_inst = host_inst
        self.host_work = host_work
        self.dev_inst = torch.frombuffer(bytearray(host_inst), dtype=torch.uint8).to(device)
        self.dev_work = torch.frombuffer(bytearray(host_work), dtype=torch.uint8).to(device)
        self.n_inst = len(host_inst)
        self.n_work = len(host_work)
        self.ws = None  # partitioned A^T: halo workspace, allocated at bind time


class MultiProjector:
    """A / matched A^T over uneven instances, one kernel launch per call.

    ``bricks``: per-instance (forward, adjoint) brick shapes; default
    26x26x14 / 32x16x16 (the single-call defaults), ``"tune"`` to run the
    autotuner per instance, or ``"tune_shared"`` for the one A shape and one
    A^T shape that time fastest on this multi launch (one compiled shape
    keeps the launch on its constant-shape kernel).  ``persistent_blocks``: workgroups that pull
    (instance, brick) items from a device counter until the list is empty;
    ``"auto"`` (default) = as many as are resident at once for each kernel
    (the library caps the grid by the kernel's LDS and registers: three per
    compute unit for the default bricks); 0 = one workgroup per item.
    ``view_parts``: per-instance (forward, adjoint) splits of each brick's views
    over work items (forward at most 16, adjoint at most 2), a single int for
    every instance, or ``"auto"`` (default): bricks whose views would cost more
    than a sixth of a workgroup slot's share of the whole launch are split, so
    the launch has enough similar-sized items to balance (:func:`auto_parts`).
    ``partition`` (default True): A^T marches each sample in the one brick
    holding its floor cell and adds the bricks' upper-halo sums in a second
    kernel (``dtv_cone_adjoint_multi_ws``, equal to ``ops.adjoint`` per
    instance for the same brick and split); False runs the unpartitioned
    ``dtv_cone_adjoint_multi``."""

    def __init__(self, geoms, bricks=None, block=256, persistent_blocks="auto", device=None,
                 view_parts="auto", partition=True, forward_block=512):
        self.geoms = list(geoms)
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.block = block                  # A^T lanes per workgroup (at most 256)
        self.forward_block = forward_block  # A lanes per workgroup (at most 512)
        if persistent_blocks == "auto":
            # the library caps a persistent grid at the workgroups the device
            # holds at once for the kernel it launches (LDS, registers)
            cus = torch.cuda.get_device_properties(self.device).multi_processor_count
            persistent_blocks = 8 * cus
        self.persistent_blocks = int(persistent_blocks)
        self.tuned = None
        if bricks == "tune":
            bricks = []
            for g in self.geoms:
                t = autotune.tune(g, self.device)
                bricks.append((t["forward"]["brick"], t["adjoint"]["brick"]))
        elif bricks == "tune_shared":
            self.tuned = autotune.tune_multi(self.geoms, self.device)
            bricks = [(self.tuned["forward"], self.tuned["adjoint"])] * len(self.geoms)
        self.bricks = bricks or [((26, 26, 14), (32, 16, 16))] * len(self.geoms)
        if view_parts == "auto":
            cus = torch.cuda.get_device_properties(self.device).multi_processor_count
            view_parts = auto_parts(self.geoms, self.bricks, 3 * cus)
        elif isinstance(view_parts, int):
            view_parts = [(view_parts, view_parts)] * len(self.geoms)
        self.view_parts = [(int(f), int(a)) for f, a in view_parts]
        self.partition = bool(partition)
        self.counter = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.views = [g.device_views(self.device) for g in self.geoms]

    def _table(self, adjoint, ins, outs, inits):
        n = len(self.geoms)
        host = (Instance * n)()
        work = []
        for i, g in enumerate(self.geoms):
            e = host[i]
            b = self.bricks[i][1 if adjoint else 0]
            X, Y, Z = g.volume_size
            U, V = g.detector_size
            e.in_ = ins[i].data_ptr()
            e.out = outs[i].data_ptr()
            e.init = inits[i].data_ptr() if (inits is not None and inits[i] is not None) else None
            e.views = self.views[i].data_ptr()
            e.nx, e.ny, e.nz, e.nu, e.nv, e.n_views = X, Y, Z, U, V, g.n_views
            e.axis_mode = g.scheme.axis_mode
            e.bx, e.by, e.bz = b
            ext = 0 if adjoint else 1
            e.nbx = (X + ext + b[0] - 1) // b[0]
            e.nby = (Y + ext + b[1] - 1) // b[1]
            e.nbz = (Z + ext + b[2] - 1) // b[2]
            for c in range(3):
                e.box_lo[c] = g.scheme.box_lo[c]
                e.box_hi[c] = g.scheme.box_hi[c]
            e.margin = g.scheme.margin
            e.scale = g.scheme.scale
            parts = self.view_parts[i][1 if adjoint else 0]
            e.view_parts = parts
            nb = e.nbx * e.nby * e.nbz
            cost = float(g.n_views * b[0] * b[1] * b[2]) / parts
            work += [(cost, i, k) for k in range(nb * parts)]
        work.sort(key=lambda t: (-t[0], t[1], t[2]))
        wh = (WorkItem * len(work))()
        for j, (_, i, k) in enumerate(work):
            wh[j].instance = i
            wh[j].brick = k
        return _Bound(host, wh, self.device)

    def bind_forward(self, volumes, sinos, inits=None):
        """Bind buffers for out[i] = A volumes[i] - inits[i] (inits optional)."""
        return self._table(False, volumes, sinos, inits)

    def bind_adjoint(self, sinos, volumes):
        b = self._table(True, sinos, volumes, None)
        if self.partition:
            nbytes = ctypes.c_size_t(0)
            rc = _capi.lib().dtv_cone_adjoint_multi_ws_bytes(ctypes.addressof(b.host_inst),
                                                             b.n_inst, ctypes.byref(nbytes))
            _capi.check(rc, "dtv_cone_adjoint_multi_ws_bytes")
            b.ws = torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)
        return b

    def run(self, bound, adjoint: bool):
        L = _capi.lib()
        args = (ctypes.addressof(bound.host_inst), bound.dev_inst.data_ptr(), bound.n_inst,
                ctypes.addressof(bound.host_work), bound.dev_work.data_ptr(), bound.n_work,
                self.block if adjoint else self.forward_block,
                self.counter.data_ptr() if self.persistent_blocks else None,
                self.persistent_blocks)
        with torch.cuda.device(self.device):
            if adjoint and bound.ws is not None:
                name = "dtv_cone_adjoint_multi_ws"
                rc = L.dtv_cone_adjoint_multi_ws(*args, bound.ws.data_ptr(), bound.ws.numel(),
                                                 ops._stream(self.device))
            else:
                name = "dtv_cone_adjoint_multi" if adjoint else "dtv_cone_forward_multi"
                fn = L.dtv_cone_adjoint_multi if adjoint else L.dtv_cone_forward_multi
                rc = fn(*args, ops._stream(self.device))
        _capi.check(rc, name)

    def forward(self, volumes, inits=None):
        """Eager convenience: returns new sinograms (A v_i - init_i)."""
        sinos = [torch.empty(g.sino_shape(1), device=self.device) for g in self.geoms]
        self.run(self.bind_forward(volumes, sinos, inits), False)
        return sinos

    def adjoint(self, sinos):
        vols = [torch.empty(g.volume_shape(1), device=self.device) for g in self.geoms]
        self.run(self.bind_adjoint(sinos, vols), True)
        return vols


def auto_parts(geoms, bricks, slots):
    """Per-instance (forward, adjoint) view splits for a multi launch: with
    per-brick cost c_i = views x brick voxels and a target item cost of the
    launch's total over 6 x ``slots`` (about six items per resident
    workgroup), A splits a brick into ceil(c_i / target) parts (at most 16)
    and A^T into two when c_i exceeds twice the target."""
    def count(g, b, adjoint):
        e = 0 if adjoint else 1
        X, Y, Z = g.volume_size
        return ((X + e + b[0] - 1) // b[0]) * ((Y + e + b[1] - 1) // b[1]) * \
            ((Z + e + b[2] - 1) // b[2])

    out = []
    costs = []
    for adjoint in (False, True):
        c = [g.n_views * b[int(adjoint)][0] * b[int(adjoint)][1] * b[int(adjoint)][2]
             for g, b in zip(geoms, bricks)]
        total = sum(ci * count(g, b[int(adjoint)], adjoint) for ci, g, b in zip(c, geoms, bricks))
        costs.append((c, total / (6.0 * slots)))
    (cf, tf), (ca, ta) = costs
    for i in range(len(geoms)):
        pf = max(1, min(16, -(-cf[i] // max(tf, 1e-30))))
        pa = 2 if ca[i] > 2.0 * ta else 1
        out.append((int(pf), int(pa)))
    return out


class BatchedFistaTV:
    """``iters`` FISTA-TV iterations over uneven instances (3 launches per
    iteration, all tables static, graph-capturable)."""

    def __init__(self, geoms, sinos, iters=20, lam=1e-3, eps=1e-3, steps=None, bricks=None,
                 block=256, persistent_blocks="auto", mode="neumann", view_parts="auto"):
        self.mp = MultiProjector(geoms, bricks, block, persistent_blocks, view_parts=view_parts)
        dev = self.mp.device
        self.geoms = self.mp.geoms
        self.iters = iters
        self.mode = tv.MODES[mode]
        n = len(self.geoms)
        if steps is None:
            steps = [1.0 / (unroll.power_norm(g, device=dev) + lam * 12.0 / eps)
                     for g in self.geoms]
        self.b = [s.contiguous() for s in sinos]
        self.x = [torch.zeros(g.volume_shape(1), device=dev) for g in self.geoms]
        self.y = [[torch.zeros(g.volume_shape(1), device=dev) for g in self.geoms]
                  for _ in range(2)]
        self.r = [torch.empty(g.sino_shape(1), device=dev) for g in self.geoms]
        self.g = [torch.empty(g.volume_shape(1), device=dev) for g in self.geoms]
        self.fwd = [self.mp.bind_forward(self.y[p], self.r, self.b) for p in range(2)]
        self.adj = self.mp.bind_adjoint(self.r, self.g)
        betas = unroll.fista_momentum(iters)
        self.tv_tables = []
        for k in range(iters):
            p = k % 2
            items = (tv.TvItem * n)()
            for i, g in enumerate(self.geoms):
                X, Y, Z = g.volume_size
                it = items[i]
                it.y = self.y[p][i].data_ptr()
                it.g = self.g[i].data_ptr()
                it.xprev = self.x[i].data_ptr()
                it.xnew = self.x[i].data_ptr()
                it.ynew = self.y[1 - p][i].data_ptr()
                it.nx, it.ny, it.nz = X, Y, Z
                it.nonneg = 1
                it.step, it.lam, it.eps, it.beta = steps[i], lam, eps, betas[k]
            self.tv_tables.append(
                (items, torch.frombuffer(bytearray(items), dtype=torch.uint8).to(dev)))
        self.graph = None

    def reset(self):
        for t in self.x + self.y[0] + self.y[1]:
            t.zero_()

    def tv_step(self, k):
        """Iteration ``k``'s fused TV/FISTA update over every instance (one
        launch)."""
        host, dev_table = self.tv_tables[k]
        rc = _capi.lib().dtv_tv_fista_step_multi(ctypes.addressof(host), dev_table.data_ptr(),
                                                 len(self.geoms), self.mode,
                                                 ops._stream(self.mp.device))
        _capi.check(rc, "dtv_tv_fista_step_multi")

    def run(self):
        """Eager unroll (also what gets captured)."""
        for k in range(self.iters):
            p = k % 2
            self.mp.run(self.fwd[p], adjoint=False)
            self.mp.run(self.adj, adjoint=True)
            self.tv_step(k)
        return self.x

    def capture(self):
        """Record reset + the whole unroll in one graph (fill kernels only,
        no memset nodes)."""
        s = torch.cuda.Stream(self.mp.device)
        s.wait_stream(torch.cuda.current_stream(self.mp.device))
        with torch.cuda.stream(s):
            self.reset()
            self.run()
        torch.cuda.current_stream(self.mp.device).wait_stream(s)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.reset()
            self.run()
        return self.graph

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.x


def plain_fista_tv(geoms, sinos, iters=20, lam=1e-3, eps=1e-3, steps=None):
    """The per-volume path: a Python loop over instances, each running the
    unfused FISTA-TV iteration on the default single-instance kernels."""
    out = []
    for i, g in enumerate(geoms):
        step = None if steps is None else steps[i]
        out.append(unroll.fista_tv(sinos[i], g, iters, lam, eps, step=step, fused=False))
    return out
