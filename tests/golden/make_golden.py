"""Writes tests/golden/oracle_small.npz: small seeded cases of every projector
scheme and of the TV operators, computed by the FP64 oracle.

The reference ships no tests, fixtures or data files, so these fixtures are
self-generated: they freeze the oracle (tests/test_golden.py re-runs it on the
stored inputs and checks the stored outputs) and give the GPU path fixed
input/output pairs to meet (same file, ``-m gpu`` tests).  Inputs are stored
with the outputs, so nothing depends on re-generating them bit for bit.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dtv-net_amd"))

from oracle import cone, tv  # noqa: E402
from dtvproj import geometry as G  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "oracle_small.npz")

# Standard (reference world units: one detector pixel), includes the cardinal
# angles whose central detector column runs parallel to a voxel axis
STD = dict(vol=(10, 8, 6), det=(12, 7), angles=[0.0, 90.0, 37.5, 200.0, 315.0],
           sid=60.0, sdd=120.0)
# Beijing vectors: 2 source angles x 21 planes (B_ref groups views by 21)
BJ = dict(vol=(8, 8, 6), det=(6, 5), n_angles=2, planes=21, sid=40.0, sdd=70.0,
          pitch=(0.6, 0.7))
# physical circular orbit (mm)
CIRC = dict(vol=(8, 8, 8), voxel=1.0, det=(12, 10), pitch=1.2, n_views=6, sod=40.0,
            sdd=80.0)
TV = dict(vol=(7, 6, 5), eps=0.05)


def gaussian(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(shape, generator=g, dtype=torch.float64)


def build():
    f = {}
    # --- Standard
    X, Y, Z = STD["vol"]
    U, V = STD["det"]
    A = len(STD["angles"])
    vol = gaussian((1, 1, Z, Y, X), 101)
    sino = gaussian((1, 1, A, V, U), 102)
    ang = np.asarray(STD["angles"], dtype=np.float64)
    args = (ang, STD["vol"], STD["det"], STD["sid"], STD["sdd"])
    f["std_vol"] = vol.numpy()
    f["std_sino"] = sino.numpy()
    f["std_angles"] = ang
    f["std_forward"] = cone.std_forward(vol, *args).numpy()
    f["std_adjoint"] = cone.std_adjoint(sino, *args).numpy()
    f["std_bp_ref"] = cone.std_bp_ref(sino, *args).numpy()
    # --- Beijing
    X, Y, Z = BJ["vol"]
    U, V = BJ["det"]
    P = G.beijing_default_vectors(BJ["n_angles"], BJ["planes"], BJ["det"], BJ["sid"],
                                  BJ["sdd"], BJ["pitch"])
    n = P.shape[0]
    vol = gaussian((1, 1, Z, Y, X), 111)
    sino = gaussian((1, 1, n, V, U), 112)
    args = (P, BJ["vol"], BJ["det"])
    f["bj_vectors"] = P
    f["bj_vol"] = vol.numpy()
    f["bj_sino"] = sino.numpy()
    f["bj_forward"] = cone.bj_forward(vol, *args).numpy()
    f["bj_adjoint"] = cone.bj_adjoint(sino, *args).numpy()
    f["bj_bp_ref"] = cone.bj_bp_ref(sino, *args).numpy()
    # --- physical circular
    X, Y, Z = CIRC["vol"]
    U, V = CIRC["det"]
    ang = G.uniform_angles(CIRC["n_views"])
    vol = gaussian((1, 1, Z, Y, X), 121)
    sino = gaussian((1, 1, len(ang), V, U), 122)
    args = (ang, CIRC["vol"], CIRC["voxel"], CIRC["det"], CIRC["pitch"], CIRC["sod"],
            CIRC["sdd"])
    f["circ_angles"] = np.asarray(ang, dtype=np.float64)
    f["circ_vol"] = vol.numpy()
    f["circ_sino"] = sino.numpy()
    f["circ_forward"] = cone.circ_forward(vol, *args).numpy()
    f["circ_adjoint"] = cone.circ_adjoint(sino, *args).numpy()
    # --- TV operators (explicit sparse matrices, both boundary modes)
    X, Y, Z = TV["vol"]
    v = gaussian((Z, Y, X), 131).numpy()
    f["tv_vol"] = v
    for mode, mk in (("neumann", tv.neumann_matrix), ("reference", tv.reference_matrix)):
        for a in "xyz":
            M = mk(TV["vol"], a)
            f[f"tv_{mode}_D{a}"] = tv.apply(M, v)
            f[f"tv_{mode}_DT{a}"] = tv.apply(M.T.tocsr(), v)
        f[f"tv_{mode}_grad"] = tv.tv_eps_grad(v, TV["eps"], TV["vol"], mode)
    return f


if __name__ == "__main__":
    fx = build()
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT}: {len(fx)} arrays, {os.path.getsize(OUT)} bytes")
