"""Reference difference operators as explicit sparse matrices -- TEST
INFRASTRUCTURE ONLY.

Restates ``model/PDHG/Utils/Gradients.py:8-47``: for a volume flattened in C
order (x fastest), each operator has -1 on the diagonal and +1 at column i+s
for the rows listed below (rows without a +1 give -f_i):

* D_x: s = 1, rows i with (i + 1) % X != 0                  (:8-16)
* D_y: s = volumeSize[1] (= Y), rows i < N - Y               (:18-25)
  -- the reference strides by Y, not X, and does not stop at slice ends;
* D_z: s = X*Y, rows i < N - X*Y                             (:27-34)

The transposes are the matrices' transposes (:37-45).  ``neumann_*`` give the
conventional forward differences with zero at the far boundary of each axis.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def reference_matrix(volume_size, axis: str) -> sp.csr_matrix:
    X, Y, Z = volume_size
    N = X * Y * Z
    rows = list(range(N))
    cols = list(range(N))
    data = [-1.0] * N
    if axis == "x":
        for i in range(N):
            if (i + 1) % X != 0:
                rows.append(i)
                cols.append(i + 1)
                data.append(1.0)
    elif axis == "y":
        for i in range(N - Y):
            rows.append(i)
            cols.append(i + Y)
            data.append(1.0)
    elif axis == "z":
        for i in range(N - X * Y):
            rows.append(i)
            cols.append(i + X * Y)
            data.append(1.0)
    else:
        raise ValueError(axis)
    return sp.csr_matrix((np.array(data), (np.array(rows), np.array(cols))), shape=(N, N))


def neumann_matrix(volume_size, axis: str) -> sp.csr_matrix:
    X, Y, Z = volume_size
    N = X * Y * Z
    idx = np.arange(N).reshape(Z, Y, X)
    rows, cols, data = [], [], []
    if axis == "x":
        a, b = idx[:, :, :-1], idx[:, :, 1:]
    elif axis == "y":
        a, b = idx[:, :-1, :], idx[:, 1:, :]
    elif axis == "z":
        a, b = idx[:-1], idx[1:]
    else:
        raise ValueError(axis)
    a = a.reshape(-1)
    b = b.reshape(-1)
    rows = np.concatenate([a, a])
    cols = np.concatenate([a, b])
    data = np.concatenate([-np.ones(a.size), np.ones(a.size)])
    return sp.csr_matrix((data, (rows, cols)), shape=(N, N))


def apply(M, vol):
    """M @ vec(vol) for vol [..., Z, Y, X] (numpy, FP64)."""
    v = np.asarray(vol, dtype=np.float64)
    shp = v.shape
    flat = v.reshape(-1, shp[-3] * shp[-2] * shp[-1])
    return np.stack([M @ f for f in flat]).reshape(shp)


def tv_eps_grad(vol, eps, volume_size, mode="neumann"):
    """Gradient of sum_c sqrt(sum_a (D_a f)_c^2 + eps^2), FP64:
    sum_a D_a^T ((D_a f) / n)."""
    mats = [(neumann_matrix if mode == "neumann" else reference_matrix)(volume_size, a)
            for a in "xyz"]
    v = np.asarray(vol, dtype=np.float64).reshape(-1)
    d = [M @ v for M in mats]
    n = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + eps ** 2)
    g = sum(M.T @ (di / n) for M, di in zip(mats, d))
    return g.reshape(np.asarray(vol).shape)


def asd_tv_grad(vol, eps, volume_size, mode="reference"):
    """gradTV of model/PDHG/ASDPOCS.py:38-42 in FP64:
    sum_a D_a^T ((D_a f) / (sqrt(sum_a (D_a f)^2) + eps)), eps = 1 there."""
    mats = [(neumann_matrix if mode == "neumann" else reference_matrix)(volume_size, a)
            for a in "xyz"]
    vol = np.asarray(vol, dtype=np.float64)
    out = []
    for v in vol.reshape(-1, mats[0].shape[0]):  # one volume at a time
        d = [M @ v for M in mats]
        n = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2) + eps
        out.append(sum(M.T @ (di / n) for M, di in zip(mats, d)))
    return np.stack(out).reshape(vol.shape)


def asd_potential(vol, eps, volume_size, mode="reference"):
    """sum_c [r_c - eps log(1 + r_c / eps)], r = |D f|_c (FP64): the
    functional whose gradient :func:`asd_tv_grad` is."""
    mats = [(neumann_matrix if mode == "neumann" else reference_matrix)(volume_size, a)
            for a in "xyz"]
    total = 0.0
    for v in np.asarray(vol, dtype=np.float64).reshape(-1, mats[0].shape[0]):
        r = np.sqrt(sum((M @ v) ** 2 for M in mats))
        total += float(np.sum(r - eps * np.log1p(r / eps)))
    return total


def l1ball(v, b, tol=1e-4, max_iter=100):
    """model/PDHG/Utils/L1ball.py:4-17 in FP64 (numpy): identity when
    sum|v| < b (:5-6); otherwise Newton steps on lam from 0 with slope
    -count(|v| - lam > 0) + 0.001 while |objective| > tol and fewer than
    max_iter steps (:7-14); ``np.max(lam, 0)`` of a scalar is lam (:15);
    sign(v) relu(|v| - lam) (:16)."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    if a.sum() < b:
        return v.copy()
    lam = 0.0
    obj = np.maximum(a - lam, 0).sum() - b
    it = 0
    while abs(obj) > tol and it < max_iter:
        obj = np.maximum(a - lam, 0).sum() - b
        slope = -np.count_nonzero(a - lam > 0) + 0.001
        lam -= obj / slope
        it += 1
    return np.sign(v) * np.maximum(a - lam, 0)
