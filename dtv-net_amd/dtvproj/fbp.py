"""FBP front end: cosine weighting fused with the per-row Ram-Lak filtering
(``dtv_row_filter``), the step before the backprojection in the reference's
``*WithFBP`` modules (StandardGeometry.py:55-65, BeijingGeometry.py:60-63).
"""
from __future__ import annotations

import torch

from . import _capi
from .ops import _stream


def row_filter(x: torch.Tensor, taps: torch.Tensor, weight: torch.Tensor = None) -> torch.Tensor:
    """``conv2d(x * weight, taps, padding=(0, W/2))`` over the last dimension.

    x: float32 device tensor [..., W] (rows = everything before the last
    dimension); weight: [Vw, W], applied to row r as weight[r % Vw], or None;
    taps: 1-D float32, odd length (the output keeps width W).  Runs on the
    gfx950 kernel; raises RuntimeError when the HIP library is missing."""
    if x.dtype != torch.float32 or not x.is_cuda:
        raise RuntimeError("row_filter: x must be a float32 GPU tensor")
    W = x.shape[-1]
    x = x.contiguous()
    t = taps.reshape(-1).to(device=x.device, dtype=torch.float32).contiguous()
    if t.numel() % 2 == 0:
        raise RuntimeError(f"row_filter: the kernel length must be odd, got {t.numel()}")
    w = None
    wrows = 0
    if weight is not None:
        w = weight.reshape(-1, W).to(device=x.device, dtype=torch.float32).contiguous()
        wrows = w.shape[0]
    out = torch.empty_like(x)
    rows = x.numel() // W
    with torch.cuda.device(x.device):
        rc = _capi.lib().dtv_row_filter(x.data_ptr(), out.data_ptr(), rows, W,
                                        w.data_ptr() if w is not None else None, wrows,
                                        t.data_ptr(), t.numel(), _stream(x.device))
    _capi.check(rc, "dtv_row_filter")
    return out
