"""Literal scalar FP64 restatements of the reference kernels, for tiny inputs.

TEST INFRASTRUCTURE ONLY.  These follow the reference kernels statement by
statement (one pixel or one voxel at a time, Python ``while`` loops, explicit
box tests) so that the vectorised oracle in ``cone.py`` -- its masking,
windowing, chunking and index arithmetic -- is checked against an
independent, obviously-sequential implementation.  The only deliberate
difference from the reference is that alpha is advanced as
alpha_start + k*step (exact) rather than accumulated.
"""
from __future__ import annotations

import math

import numpy as np

PI_REF = 3.14159265359


def tri(vol, x, y, z):
    """Trilinear sample of vol [Z, Y, X] at texel-centred coordinates,
    zero outside (``tex3D(x + 0.5, y + 0.5, z + 0.5)``, border mode)."""
    Z, Y, X = vol.shape
    x0, y0, z0 = math.floor(x), math.floor(y), math.floor(z)
    fx, fy, fz = x - x0, y - y0, z - z0
    acc = 0.0
    for k, wz in ((z0, 1 - fz), (z0 + 1, fz)):
        if not 0 <= k < Z:
            continue
        for j, wy in ((y0, 1 - fy), (y0 + 1, fy)):
            if not 0 <= j < Y:
                continue
            for i, wx in ((x0, 1 - fx), (x0 + 1, fx)):
                if not 0 <= i < X:
                    continue
                acc += wx * wy * wz * float(vol[k, j, i])
    return acc


def bil(img, x, y):
    """Bilinear sample of img [V, U] at texel-centred coords, zero outside."""
    V, U = img.shape
    x0, y0 = math.floor(x), math.floor(y)
    fx, fy = x - x0, y - y0
    acc = 0.0
    for j, wy in ((y0, 1 - fy), (y0 + 1, fy)):
        if not 0 <= j < V:
            continue
        for i, wx in ((x0, 1 - fx), (x0 + 1, fx)):
            if not 0 <= i < U:
                continue
            acc += wx * wy * float(img[j, i])
    return acc


def _normalize(v):
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return (v[0] / n, v[1] / n, v[2] / n)


def std_forward(vol, angles, volume_size, detector_size, sid, sdd):
    """Standard forwardKernel.cu:22-86, one pixel at a time.  vol [Z, Y, X]."""
    X, Y, Z = volume_size
    U, V = detector_size
    out = np.zeros((len(angles), V, U))
    s = sid / sdd
    for a, ang in enumerate(angles):
        th = ang / 180 * PI_REF - PI_REF / 2
        ex = (math.cos(th), math.sin(th))
        ey = (-ex[1], ex[0])
        for iv in range(V):
            for iu in range(U):
                detx = iu + (-U / 2)
                dety = iv + (-V / 2)
                w = (ex[0] * sdd + ey[0] * detx, ex[1] * sdd + ey[1] * detx)
                d = _normalize((w[0], w[1], dety))
                src = (-ex[0] * sid, -ex[1] * sid, 0.0)
                if abs(d[0]) >= abs(d[1]):
                    a0 = (-X / 2 - src[0]) / d[0]
                    a1 = (X / 2 - src[0]) / d[0]
                else:
                    a0 = (-Y / 2 - src[1]) / d[1]
                    a1 = (Y / 2 - src[1]) / d[1]
                a_lo = min(a0, a1) - 3
                a_hi = max(a0, a1) + 3
                acc = 0.0
                k = 0
                while a_lo + k * s < a_hi:
                    al = a_lo + k * s
                    px = (src[0] + al * d[0]) / s + X / 2
                    py = (src[1] + al * d[1]) / s + Y / 2
                    pz = (src[2] + al * d[2]) / s + Z / 2
                    acc += tri(vol, px, py, pz)
                    k += 1
                out[a, iv, iu] = acc * s
    return out


def bj_forward(vol, P, volume_size, detector_size):
    """Beijing forwardKernel.cu:17-76, one pixel at a time (no
    sampleInterval factor).  vol [Z, Y, X], P float32 [n, 12]."""
    X, Y, Z = volume_size
    U, V = detector_size
    P = np.asarray(P, dtype=np.float32).astype(np.float64)
    out = np.zeros((P.shape[0], V, U))
    N = (X, Y, Z)
    for j in range(P.shape[0]):
        S = (P[j, 0], P[j, 1], -P[j, 2])
        D = (P[j, 3], P[j, 4], -P[j, 5])
        v = P[j, 6:9]
        u = P[j, 9:12]
        for iv in range(V):
            for iu in range(U):
                tx = 0.5 + iu + (-U / 2)
                ty = 0.5 + iv + (-V / 2)
                pix = tuple(D[c] + tx * u[c] + ty * v[c] for c in range(3))
                d = _normalize(tuple(pix[c] - S[c] for c in range(3)))
                m = max(abs(d[0]), max(abs(d[2]), abs(d[1])))
                ax = 0 if abs(d[0]) == m else (1 if abs(d[1]) == m else 2)
                a0 = (-N[ax] / 2 - S[ax]) / d[ax]
                a1 = (N[ax] / 2 - S[ax]) / d[ax]
                a_lo = min(a0, a1) - 3
                a_hi = max(a0, a1) + 3
                acc = 0.0
                k = 0
                while a_lo + k < a_hi:
                    al = a_lo + k
                    p = [S[c] + al * d[c] + N[c] / 2 for c in range(3)]
                    acc += tri(vol, p[0], p[1], p[2])
                    k += 1
                out[j, iv, iu] = acc
    return out


def std_bp_ref(sino, angles, volume_size, detector_size, sid, sdd):
    """Standard backwardKernel.cu:22-53 one voxel column at a time.
    sino [A, V, U] -> vol [Z, Y, X]."""
    X, Y, Z = volume_size
    U, V = detector_size
    A = len(angles)
    s = sid / sdd
    vol = np.zeros((Z, Y, X))
    for a, ang in enumerate(angles):
        th = ang / 180 * PI_REF - PI_REF / 2
        ex = (math.cos(th), math.sin(th))
        ey = (-ex[1], ex[0])
        src = (ex[0] * sid, ex[1] * sid)
        det = (ex[0] * (sdd - sid), ex[1] * (sdd - sid))
        for iy in range(Y):
            for ix in range(X):
                c = ((ix - X / 2) * s, (iy - Y / 2) * s)
                # the source-through-voxel line meets the detector line where
                # its component along ex equals SDD - SID
                lam = (sdd - sid + sid) / (sid + c[0] * ex[0] + c[1] * ex[1])
                pt = (-src[0] + lam * (c[0] + src[0]), -src[1] + lam * (c[1] + src[1]))
                x = pt[0] * ey[0] + pt[1] * ey[1] + U / 2
                L = (src[0] + c[0]) * ex[0] + (src[1] + c[1]) * ex[1]
                dz = sdd / L * s
                coff = dz / s
                for k in range(Z):
                    z = (k - Z / 2) * dz + V / 2
                    vol[k, iy, ix] += bil(sino[a], x, z) * coff * coff * PI_REF / A
    return vol


def bj_bp_ref(sino, P, volume_size, detector_size):
    """Beijing backwardKernel.cu:18-45 one voxel at a time.
    sino [n, V, U] -> vol [Z, Y, X]."""
    X, Y, Z = volume_size
    U, V = detector_size
    P = np.asarray(P, dtype=np.float32).astype(np.float64)
    n = P.shape[0]
    angles_num = n // 21

    def det3(a, b, c):
        return float(np.dot(a, np.cross(b, c)))

    vol = np.zeros((Z, Y, X))
    for k in range(Z):
        for iy in range(Y):
            for ix in range(X):
                c = np.array([ix - X / 2, iy - Y / 2, k - Z / 2])
                acc = 0.0
                for j in range(n):
                    S = np.array([P[j, 0], P[j, 1], -P[j, 2]])
                    D = np.array([P[j, 3], P[j, 4], -P[j, 5]])
                    v = P[j, 6:9]
                    u = P[j, 9:12]
                    f = 1.0 / det3(u, v, S - c)
                    sx = f * det3(c - S, v, S - D) + U / 2
                    sy = f * det3(u, c - S, S - D) + V / 2
                    fr = f * det3(u, v, S - D)
                    if sx < -1 or sx > X + 1 or sy < -1 or sy > Y + 1:
                        continue
                    acc += fr * bil(sino[j], sx - 0.5, sy - 0.5)
                vol[k, iy, ix] = acc * 2 * PI_REF / angles_num
    return vol
