"""Reconstruction unrolls around the projector pair.

* :func:`power_norm` -- ||A||^2 by power iteration on A^T A (step sizes for
  the matched pair; the reference hard-codes steps tuned for its FDK
  backprojector, e.g. L = 1e-4 at FISTA.py:14).
* :func:`fista_tv` -- FISTA on 1/2 ||A x - b||^2 + lam TV_eps(x), x >= 0,
  with the fused gfx950 update (``fused=True``) or the plain torch update
  (the per-volume reference-style path).  Same iteration as
  model/FISTA/FISTA.py:16-27 (gradient step, proximal/shrink step,
  t-momentum) with the smoothed TV gradient in place of the l1 shrink.
* :class:`PDHGDTVNet` -- learned primal-dual (Chambolle-Pock) unroll for
  directional TV, following model/PDHG/DTVCP.py:29-60 (data dual
  w = (w + sigma res)/(1 + sigma), directional duals, primal step,
  over-relaxation f_bar = 2 f - f_old) with learned per-iteration sigma,
  tau and per-axis TV weights; the exact dual prox of lam ||D_a f||_1 (a
  clip to [-lam, lam]) replaces the reference's Newton l1-ball projection.
  Each iteration runs under torch.utils.checkpoint to cap activation memory.
* :func:`asd_pocs` -- ASD-POCS as model/PDHG/ASDPOCS.py:28-47 runs it
  (data step, then ``ng`` TV descent steps on the fused gfx950 kernel).
* :func:`dtv_cp` -- the directional-TV Chambolle-Pock iteration of
  model/PDHG/DTVCP.py:29-60 with the reference's l1-ball dual constraints
  (the Newton l1-ball projection on the fused gfx950 kernel).
* :func:`fista_l1` -- the soft-threshold FISTA of model/FISTA/FISTA.py:16-27
  (fused shrink + momentum kernel).
* :func:`fdpg_dtv` -- the directional-TV dual-gradient iteration of
  model/FISTA/DTVFISTA.py:23-43 (fused kernel per step).
* :func:`power_method` -- HyperParams.py:6-13 for an operator pair.
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F
from torch.utils.checkpoint import checkpoint

from . import autograd as _ag
from . import ops, tv
from .geometry import ConeGeometry


@torch.no_grad()
def power_norm(geom: ConeGeometry, iters: int = 20, seed: int = 0, device="cuda") -> float:
    """Largest eigenvalue of A^T A (= ||A||^2) by power iteration."""
    gen = torch.Generator().manual_seed(seed)
    x = torch.rand(geom.volume_shape(1), generator=gen).to(device)
    x /= x.norm()
    lam = 0.0
    for _ in range(iters):
        y = ops.adjoint(ops.forward(x, geom), geom)
        lam = float(y.norm())
        x = y / lam
    return lam


def fista_momentum(n: int):
    """beta_k = (t_k - 1) / t_{k+1} with t_1 = 1, t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2
    (FISTA.py:23-25)."""
    t = 1.0
    out = []
    for _ in range(n):
        tn = (1.0 + math.sqrt(1.0 + 4.0 * t * t)) / 2.0
        out.append((t - 1.0) / tn)
        t = tn
    return out


@torch.no_grad()
def fista_tv(b, geom: ConeGeometry, iters: int, lam: float, eps: float, step: float = None,
             x0=None, fused: bool = True, mode: str = "neumann", lip: float = None):
    """FISTA-TV reconstruction of sinogram ``b`` [B,1,A,V,U]; returns x."""
    B = b.shape[0]
    if step is None:
        if lip is None:
            lip = power_norm(geom, device=b.device)
        # gradient Lipschitz: ||A||^2 + lam ||D||^2 / eps, ||D||^2 <= 12 in 3-D
        step = 1.0 / (lip + lam * 12.0 / eps)
    x = torch.zeros(geom.volume_shape(B), device=b.device) if x0 is None else x0.clone()
    y = x.clone()
    ybuf = torch.empty_like(x)
    for beta in fista_momentum(iters):
        r = ops.forward(y, geom)
        r -= b
        g = ops.adjoint(r, geom)
        if fused:
            x, ynew = tv.fista_tv_step(y, g, x, step, lam, eps, beta, True, mode,
                                       xnew=x, ynew=ybuf)
            y, ybuf = ynew, y
        else:
            x, y = tv.fista_tv_step_torch(y, g, x, step, lam, eps, beta, True, mode)
    return x


class PDHGDTVNet(torch.nn.Module):
    """Learned ``iters``-iteration PDHG for directional TV with per-iteration
    sigma_k, tau_k and per-axis weights lam_{a,k} (softplus-positive).

    Without autograd (``torch.no_grad``, inference) on GPU tensors the
    volume updates of each iteration run as the two fused launches of
    :func:`tv.pdhg_dtv_step` in preallocated buffers (``fused=False`` keeps
    the torch ops)."""

    def __init__(self, geom: ConeGeometry, iters: int = 10, sigma: float = 0.5,
                 tau: float = 0.5, lam=(1e-2, 1e-2, 1e-2), norm: float = 1.0,
                 mode: str = "neumann", checkpointing: bool = True, project=None,
                 backproject=None, fused: bool = True):
        super().__init__()
        self.fused = fused
        self.geom = geom
        # differentiable A and A^T; defaults are the gfx950 matched pair
        self.A = project or (lambda x: _ag.project(x, geom))
        self.AT = backproject or (lambda y: _ag.backproject(y, geom))
        self.iters = iters
        self.mode = mode
        self.checkpointing = checkpointing
        # sigma * tau * ||K||^2 < 1 with K = [A; D]; ||D||^2 <= 12
        self.knorm = math.sqrt(norm + 12.0)
        inv = lambda v: math.log(math.expm1(v))  # softplus^-1
        self.sigma = torch.nn.Parameter(torch.full((iters,), inv(sigma)))
        self.tau = torch.nn.Parameter(torch.full((iters,), inv(tau)))
        self.lam = torch.nn.Parameter(torch.tensor([[inv(l) for l in lam]] * iters))

    def _iteration(self, k, f, fbar, w, px, py, pz):
        s = F.softplus(self.sigma[k]) / self.knorm
        t = F.softplus(self.tau[k]) / self.knorm
        lam = F.softplus(self.lam[k])
        res = self.A(fbar) - self._b
        w = (w + s * res) / (1 + s)
        fn, fbar, (px, py, pz) = tv.pdhg_dtv_step_torch(f, fbar, self.AT(w), (px, py, pz), s,
                                                        t, lam, self.mode)
        return fn, fbar, w, px, py, pz

    @torch.no_grad()
    def _forward_fused(self, b, f0):
        f = torch.zeros(self.geom.volume_shape(b.shape[0]), device=b.device, dtype=b.dtype) \
            if f0 is None else f0.detach().clone().contiguous()
        fbar = f.clone()
        w = torch.zeros_like(b)
        p = torch.zeros((3,) + tuple(f.shape), device=f.device, dtype=f.dtype)
        # the step sizes as the torch iteration forms them (float32), one sync
        s = (F.softplus(self.sigma) / self.knorm).tolist()
        s1 = (1 + F.softplus(self.sigma) / self.knorm).tolist()
        t = (F.softplus(self.tau) / self.knorm).tolist()
        lam = F.softplus(self.lam).tolist()
        for k in range(self.iters):
            r = self.A(fbar)
            w.add_(r.sub_(b), alpha=s[k]).div_(s1[k])
            tv.pdhg_dtv_step(f, fbar, self.AT(w), p, s[k], t[k], lam[k], self.mode)
        return f

    def forward(self, b, f0=None):
        if self.fused and b.is_cuda and not torch.is_grad_enabled():
            return self._forward_fused(b, f0)
        self._b = b
        B = b.shape[0]
        f = torch.zeros(self.geom.volume_shape(B), device=b.device, dtype=b.dtype) \
            if f0 is None else f0
        w = torch.zeros_like(b)
        px = torch.zeros_like(f)
        py = torch.zeros_like(f)
        pz = torch.zeros_like(f)
        fbar = f
        for k in range(self.iters):
            if self.checkpointing and torch.is_grad_enabled():
                f, fbar, w, px, py, pz = checkpoint(self._iteration, k, f, fbar, w, px, py, pz,
                                                    use_reentrant=False)
            else:
                f, fbar, w, px, py, pz = self._iteration(k, f, fbar, w, px, py, pz)
        return f


@torch.no_grad()
def power_method(op, op_t, shape, iters: int = 10, seed: int = 0, device="cpu",
                 dtype=torch.float32) -> float:
    """||K|| as model/PDHG/Utils/HyperParams.py:6-13 estimates it: from a
    seeded uniform start, ``iters`` steps x = K^T K x / ||K^T K x||, then
    ||K x||."""
    gen = torch.Generator().manual_seed(seed)
    x = torch.rand(shape, generator=gen, dtype=torch.float64).to(device=device, dtype=dtype)
    s = 0.0
    for _ in range(iters):
        x = op_t(op(x))
        x = x / x.norm()
        s = float(op(x).norm())
    return s


@torch.no_grad()
def asd_pocs(image, sino, project, backproject, cascades: int = 30, beta: float = 5e-5,
             beta_red: float = 0.9999, ng: int = 10, alpha: float = 1e-3, rmax: float = 0.95,
             alpha_red: float = 0.95, eps: float = 1.0, mode: str = "reference",
             fused: bool = True, adaptive: bool = False):
    """ASD-POCS of model/PDHG/ASDPOCS.py:28-47.

    Per cascade: f = f - beta B(A f - sino); dp = mean|f0 - f|; on the first
    cascade the TV step is fixed to alpha * dp; ``ng`` steps
    f -= step * sum_a D_a^T (D_a f / (|D f| + eps)) (:37-43, eps = 1); f0 = f;
    beta *= beta_red; the result is relu(f).  In the reference f0 is the same
    tensor as f, whose TV steps run in place, so its dg = mean|f - f0| is 0 and
    the step is never reduced; ``adaptive=True`` measures dg against f before
    the TV steps and multiplies the step by ``alpha_red`` when dg > rmax * dp
    (the rule as published).  ``project``/``backproject`` are the pair the
    caller runs (the reference: the Beijing module's ForwardProjection and
    BackProjection); the TV steps use the fused kernel when ``fused`` (FP32
    GPU volumes) and the torch formulation otherwise."""
    f = image
    f0 = image
    step = None
    for i in range(cascades):
        f = f - beta * backproject(project(f) - sino)
        if i == 0 or adaptive:
            dp = float((f0 - f).abs().mean())
        if i == 0:
            step = alpha * dp
        f_data = f
        if fused:
            bufs = (f, torch.empty_like(f))
            for j in range(ng):
                tv.tv_descent(bufs[j % 2], step, eps, mode, out=bufs[(j + 1) % 2])
            f = bufs[ng % 2] if ng else f
        else:
            for _ in range(ng):
                f = tv.tv_descent_torch(f, step, eps, mode)
        if adaptive:
            dg = float((f - f_data).abs().mean())
            if dg > rmax * dp:
                step *= alpha_red
        f0 = f
        beta *= beta_red
    return F.relu(f)


@torch.no_grad()
def dtv_cp(image, sino, project, backproject, cascades: int = 30, v=(1.0, 1.0, 1.0),
           mu: float = 1.0, sigma: float = 1e-7, tau: float = 1e-7, mode: str = "reference",
           fused: bool = True, refresh: int = 10):
    """Directional-TV Chambolle-Pock of model/PDHG/DTVCP.py:29-60.

    Every ``refresh`` cascades (10, as the reference) the bounds
    t_a = sum|D_a f| are refreshed (:34-37).
    Per cascade: res = A f_bar - sino; w = (w + sigma res)/(1 + sigma);
    r = B w; for each axis p_a' = p_a + v_a sigma D_a f_bar and
    p_a = p_a' - sigma sign(p_a') l1ball(|p_a'| / sigma, v_a t_a) (:41-46);
    c = -relu(-c - sigma mu f_bar) (:47); f_new = f - r - tau (sum_a v_a
    D_a^T p_a + mu c) (:48-52 -- the reference does not scale r by tau);
    f_bar = 2 f_new - f.  ``v`` = (v1, v2, v3): the reference sets
    v_a = 1e5 / ||D_a||, mu = 1e5 / ||I||, sigma = tau = 1e-7 (:18-27).
    With ``fused`` the three l1-ball projections and dual updates run as one
    ``dtv_l1ball`` call (the Newton iterations on the GPU, no host
    synchronisation); otherwise as torch ops (:func:`tv.l1ball_torch`)."""
    w = torch.zeros_like(sino)
    p = [torch.zeros_like(image) for _ in range(3)]
    c = torch.zeros_like(image)
    f = image
    fbar = image
    t = [0.0, 0.0, 0.0]
    axes = ("x", "y", "z")
    for k in range(cascades):
        if k % refresh == 0:
            t = [float(tv.D(f, a, mode).abs().sum()) for a in axes]
        res = project(fbar) - sino
        w = (w + sigma * res) / (1 + sigma)
        recon = backproject(w)
        pas = [p[i] + v[i] * sigma * tv.D(fbar, a, mode) for i, a in enumerate(axes)]
        if fused:
            p = tv.l1ball_dual([pa.contiguous() for pa in pas], sigma,
                               [v[i] * t[i] for i in range(3)])
        else:
            p = [pa - sigma * torch.sign(pa) * tv.l1ball_torch(pa.abs() / sigma, v[i] * t[i])
                 for i, pa in enumerate(pas)]
        c = -F.relu(-c - sigma * mu * fbar)
        reg = sum(v[i] * tv.DT(p[i], a, mode) for i, a in enumerate(axes)) + mu * c
        fnew = f - recon - tau * reg
        fbar = 2 * fnew - f
        f = fnew
    return f


@torch.no_grad()
def fista_l1(image, sino, project, backproject, cascades: int = 30, lamb: float = 10e3,
             L: float = 1e-4, fused: bool = True):
    """Soft-threshold FISTA of model/FISTA/FISTA.py:16-27 (its constants:
    lamb = 10e3, L = 1e-4, tuned for the reference's FDK backprojector).

    Per cascade: d = y - L B(A y - sino); I = sign(d) relu(|d| - lamb L);
    t' = (1 + sqrt(1 + 4 t^2)) / 2; y = I + (t - 1)/t' (I - I_prev); returns I.
    With ``fused`` the shrink and momentum are one ``dtv_shrink_step``
    launch; otherwise torch ops."""
    t = 1.0
    x = image.clone()          # I = Ip = image
    y = image.clone()
    for _ in range(cascades):
        g = backproject(project(y) - sino)
        tp = (1.0 + math.sqrt(1.0 + 4.0 * t * t)) / 2.0
        beta = (t - 1.0) / tp
        if fused:
            x, y = tv.shrink_step(y.contiguous(), g.contiguous(), x, L, lamb * L, beta,
                                  xnew=x, ynew=torch.empty_like(y))
        else:
            x, y = tv.shrink_step_torch(y, g, x, L, lamb * L, beta)
        t = tp
    return x


@torch.no_grad()
def fdpg_dtv(image, sino, project, backproject, v, cascades: int = 30, sigma: float = 1e-3,
             t=(1e-3, 1e-3, 1e-3), mode: str = "reference", fused: bool = True):
    """Directional-TV dual-gradient iteration of model/FISTA/DTVFISTA.py:23-43.

    Per cascade: z = relu(f - sigma B(A f - sino)); per axis
    p_a = sign(D_a z) max(|D_a z| - v_a) (torch.max without a dimension: one
    scalar over the whole tensor); f_new = sum_a t_a D_a^T p_a + z;
    f = f_new + (t - 1)/t' (f_new - f).  The reference computes t' from t but
    never assigns t, so t stays 1 and the momentum factor is 0; that is
    kept.  ``v`` = (v1, v2, v3): the reference sets v_a = 0.88 ||D_a|| /
    1e-3 (``power_method`` of its sparse matrices, :16-18).  With ``fused``
    each step after A and A^T is one ``dtv_fdpg_step`` (three launches);
    otherwise torch ops."""
    f = image
    tm = 1.0
    for _ in range(cascades):
        tp = (1.0 + math.sqrt(1.0 + 4.0 * tm * tm)) / 2.0
        beta = (tm - 1.0) / tp
        g = backproject(project(f) - sino)
        if fused:
            f = tv.fdpg_step(f.contiguous(), g.contiguous(), sigma, v, t, beta, mode)
        else:
            f = tv.fdpg_step_torch(f, g, sigma, v, t, beta, mode)
    return f
