"""Golden fixtures (tests/golden/oracle_small.npz, written by
tests/golden/make_golden.py from the FP64 oracle).

CPU: the oracle still reproduces every stored output from the stored inputs
(regression freeze), and the stored pairs satisfy the identities they must:
<A x, y> = <x, A^T y> for every scheme, and the reference backprojectors are
*not* the adjoint (the reference wires its FDK backprojector as A's
gradient).  GPU (``-m gpu``): the HIP path meets the stored outputs to the
parity bar, for both kernel families."""
import os

import numpy as np
import pytest
import torch

from conftest import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "oracle_small.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX) as f:
        return {k: f[k] for k in f.files}


def _gen():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_oracle_reproduces_fixtures(fx):
    fresh = _gen().build()
    assert set(fresh) == set(fx)
    for k, v in fresh.items():
        ref = fx[k]
        assert v.shape == ref.shape, k
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(v - ref).max() <= 1e-12 * scale, k


@pytest.mark.parametrize("p", ["std", "bj", "circ"])
def test_fixture_adjoint_identity(fx, p):
    lhs = np.vdot(fx[f"{p}_forward"], fx[f"{p}_sino"])
    rhs = np.vdot(fx[f"{p}_vol"], fx[f"{p}_adjoint"])
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("p", ["std", "bj"])
def test_fixture_reference_backprojector_is_not_the_adjoint(fx, p):
    a, b = fx[f"{p}_adjoint"], fx[f"{p}_bp_ref"]
    assert np.linalg.norm(a - b) > 0.1 * np.linalg.norm(a)


def test_fixture_tv_transpose_pairs(fx):
    v = fx["tv_vol"]
    for mode in ("neumann", "reference"):
        for a in "xyz":
            # <D v, v> = <v, D^T v>
            assert np.isclose(np.vdot(fx[f"tv_{mode}_D{a}"], v), np.vdot(v, fx[f"tv_{mode}_DT{a}"]),
                              rtol=1e-12, atol=1e-12)


# ---- GPU: the HIP path against the fixtures --------------------------------

TOL = 1e-5
LAUNCHES = [None, dict(algorithm=1), dict(algorithm=2, brick=(8, 8, 8), block=64)]


def _r(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev, torch.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("launch", LAUNCHES)
def test_gpu_standard_fixture(cuda, fx, launch):
    import JITStandardGeometry as J
    from dtvproj import geometry as G, ops
    STD = _gen().STD
    print(f"[case] golden standard launch={launch}", flush=True)
    vs, ds = STD["vol"], STD["det"]
    ang = fx["std_angles"]
    g = G.standard(ang, vs, ds, STD["sid"], STD["sdd"])
    assert rel_l2(ops.forward(_t(fx["std_vol"], cuda), g, launch), _r(fx["std_forward"])) <= TOL
    assert rel_l2(ops.adjoint(_t(fx["std_sino"], cuda), g, launch), _r(fx["std_adjoint"])) <= TOL
    args = (torch.from_numpy(ang).to(cuda), torch.tensor(vs).to(cuda), torch.tensor(ds).to(cuda),
            STD["sid"], STD["sdd"], 0)
    assert rel_l2(J.forward(_t(fx["std_vol"], cuda), *args), _r(fx["std_forward"])) <= TOL
    assert rel_l2(J.backward(_t(fx["std_sino"], cuda), *args), _r(fx["std_bp_ref"])) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("launch", LAUNCHES)
def test_gpu_beijing_fixture(cuda, fx, launch):
    import JITBeijingGeometry as J
    from dtvproj import geometry as G, ops
    BJ = _gen().BJ
    print(f"[case] golden beijing launch={launch}", flush=True)
    vs, ds = BJ["vol"], BJ["det"]
    P = fx["bj_vectors"]
    g = G.beijing(P, vs, ds)
    assert rel_l2(ops.forward(_t(fx["bj_vol"], cuda), g, launch), _r(fx["bj_forward"])) <= TOL
    assert rel_l2(ops.adjoint(_t(fx["bj_sino"], cuda), g, launch), _r(fx["bj_adjoint"])) <= TOL
    args = (torch.tensor(vs, dtype=torch.int32).to(cuda), torch.tensor(ds, dtype=torch.int32).to(cuda),
            torch.from_numpy(P).to(cuda), 0)
    assert rel_l2(J.backward(_t(fx["bj_sino"], cuda), *args), _r(fx["bj_bp_ref"])) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("launch", LAUNCHES)
def test_gpu_circular_fixture(cuda, fx, launch):
    from dtvproj import geometry as G, ops
    CIRC = _gen().CIRC
    print(f"[case] golden circular launch={launch}", flush=True)
    g = G.circular(fx["circ_angles"], CIRC["vol"], CIRC["voxel"], CIRC["det"], CIRC["pitch"],
                   CIRC["sod"], CIRC["sdd"])
    assert rel_l2(ops.forward(_t(fx["circ_vol"], cuda), g, launch), _r(fx["circ_forward"])) <= TOL
    assert rel_l2(ops.adjoint(_t(fx["circ_sino"], cuda), g, launch), _r(fx["circ_adjoint"])) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["neumann", "reference"])
def test_gpu_tv_fixture(cuda, fx, mode):
    from dtvproj import tv
    TV = _gen().TV
    print(f"[case] golden tv mode={mode}", flush=True)
    v = _t(fx["tv_vol"], cuda)
    for a in "xyz":
        assert rel_l2(tv.D(v, a, mode), _r(fx[f"tv_{mode}_D{a}"])) <= TOL
        assert rel_l2(tv.DT(v, a, mode), _r(fx[f"tv_{mode}_DT{a}"])) <= TOL
    assert rel_l2(tv.tv_eps_grad(v, TV["eps"], mode), _r(fx[f"tv_{mode}_grad"])) <= TOL


# ---- full-size workload fixtures (tests/golden/oracle_workloads.npz) --------

WFIX = os.path.join(ROOT, "tests", "golden", "oracle_workloads.npz")


def _wgen():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_workloads", os.path.join(ROOT, "tests", "golden", "make_workloads.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_workload_fixture_sparse256_reproduces_on_a_voxel_subset():
    """The oracle still gives the stored full-size A^T at 48 of the stored
    voxels (the sinogram regenerated from its seed, checked by checksum)."""
    from oracle import cone
    from dtvproj import phantom
    m = _wgen()
    c = m.S256
    with np.load(WFIX) as f:
        vox, ref = f["s256_voxels"], f["s256_adjoint"]
        ssum, ssq = float(f["s256_sino_sum"]), float(f["s256_sino_sumsq"])
    assert np.array_equal(vox, m.s256_voxels())
    y = phantom.gaussian((1, 1, c["n_views"], c["det"][1], c["det"][0]), c["sino_seed"]).double()
    assert float(y.sum()) == ssum and float((y * y).sum()) == ssq
    pick = np.r_[0:24, len(vox) - 24:len(vox)]
    got = cone.circ_adjoint_at(y, m.angles(c["n_views"]), c["vol"], c["voxel"], c["det"],
                               c["pitch"], c["sod"], c["sdd"], vox[pick])[0].numpy()
    assert np.abs(got - ref[pick]).max() <= 1e-12 * np.abs(ref).max()


def test_workload_fixture_uneven16_is_consistent():
    m = _wgen()
    with np.load(WFIX) as f:
        idx, vals = f["u16_voxels"], f["u16_values"]
        step, lip = float(f["u16_step"]), float(f["u16_lip"])
        peak, total = float(f["u16_peak"]), float(f["u16_sum"])
    c = m.U16
    assert np.array_equal(idx, m.u16_voxels())
    assert abs(step - 1.0 / (lip + c["lam"] * 12.0 / c["eps"])) <= 1e-15 * step
    assert np.isfinite(vals).all() and vals.min() >= 0.0 and vals.max() <= peak
    assert total > 0 and peak > 0


@pytest.mark.slow
def test_workload_fixture_large512_reproduces_on_a_subset():
    """The oracle still gives the stored 512x512x384 / 120-view A at 32 of the
    stored rays and A^T at 32 of the stored voxels (inputs regenerated from
    their seeds, checked by checksum)."""
    from oracle import cone
    from dtvproj import phantom
    m = _wgen()
    c = m.L512
    X, Y, Z = c["vol"]
    U, V = c["det"]
    args = (m.angles(c["n_views"]), c["vol"], c["voxel"], c["det"], c["pitch"], c["sod"], c["sdd"])
    with np.load(WFIX) as f:
        fx = {k: f[k] for k in f.files if k.startswith("l512_")}
    assert np.array_equal(fx["l512_rays"], m.l512_rays())
    assert np.array_equal(fx["l512_voxels"], m.l512_voxels())
    x = phantom.gaussian((1, 1, Z, Y, X), c["vol_seed"]).double()
    assert float(x.sum()) == float(fx["l512_vol_sum"])
    pick = np.r_[0:16, len(fx["l512_rays"]) - 16:len(fx["l512_rays"])]
    got = cone.circ_forward_at(x, *args, fx["l512_rays"][pick])[0].numpy()
    assert np.abs(got - fx["l512_forward"][pick]).max() <= 1e-12 * np.abs(fx["l512_forward"]).max()
    del x
    y = phantom.gaussian((1, 1, c["n_views"], V, U), c["sino_seed"]).double()
    assert float(y.sum()) == float(fx["l512_sino_sum"])
    vpick = np.r_[0:16, len(fx["l512_voxels"]) - 16:len(fx["l512_voxels"])]
    got = cone.circ_adjoint_at(y, *args, fx["l512_voxels"][vpick])[0].numpy()
    assert np.abs(got - fx["l512_adjoint"][vpick]).max() <= 1e-12 * np.abs(fx["l512_adjoint"]).max()


@pytest.mark.slow
def test_workload_fixture_sparse256_forward_reproduces_on_a_subset():
    from oracle import cone
    from dtvproj import phantom
    m = _wgen()
    c = m.S256
    X, Y, Z = c["vol"]
    with np.load(WFIX) as f:
        rays, ref, vsum = f["s256f_rays"], f["s256f_forward"], float(f["s256f_vol_sum"])
    assert np.array_equal(rays, m.s256_rays())
    x = phantom.gaussian((1, 1, Z, Y, X), 20).double()
    assert float(x.sum()) == vsum
    pick = np.r_[0:16, len(rays) - 16:len(rays)]
    got = cone.circ_forward_at(x, m.angles(c["n_views"]), c["vol"], c["voxel"], c["det"],
                               c["pitch"], c["sod"], c["sdd"], rays[pick])[0].numpy()
    assert np.abs(got - ref[pick]).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.slow
def test_workload_fixture_beijing_reproduces_on_a_subset():
    """The oracle still gives the stored 3024-view Beijing A^T at 32 of the
    stored voxels (the sinogram regenerated from its seed, checked by sum)."""
    from oracle import cone
    from dtvproj import phantom
    m = _wgen()
    with np.load(WFIX) as f:
        fx = {k: f[k] for k in f.files if k.startswith("bj_")}
    assert np.array_equal(fx["bj_voxels"], m.bj_voxels())
    g = m.bj_geometry()
    assert g.n_views == 3024
    y = phantom.gaussian(g.sino_shape(1), m.BJ["sino_seed"]).double()
    assert abs(float(y.sum()) - float(fx["bj_sino_sum"])) <= 1e-12 * float(fx["bj_sino_sumsq"])
    pick = np.r_[0:16, len(fx["bj_voxels"]) - 16:len(fx["bj_voxels"])]
    sc = g.scheme
    got = cone.table_adjoint_at(y, g.views, g.volume_size, g.detector_size, sc.axis_mode,
                                sc.box_lo, sc.box_hi, sc.margin, sc.scale,
                                fx["bj_voxels"][pick])[0].numpy()
    assert np.abs(got - fx["bj_adjoint"][pick]).max() <= 1e-12 * np.abs(fx["bj_adjoint"]).max()


def test_workload_fixture_standard_options_reproduces_on_a_subset():
    """The oracle still gives the stored Standard-options A^T (all 48 views)
    at 32 of the stored voxels."""
    from oracle import cone
    from dtvproj import phantom
    m = _wgen()
    with np.load(WFIX) as f:
        fx = {k: f[k] for k in f.files if k.startswith("std_")}
    assert np.array_equal(fx["std_voxels"], m.std_voxels())
    g = m.std_geometry()
    y = phantom.gaussian(g.sino_shape(1), m.STD["sino_seed"]).double()
    assert abs(float(y.sum()) - float(fx["std_sino_sum"])) <= 1e-12 * float(fx["std_sino_sumsq"])
    pick = np.r_[0:16, len(fx["std_voxels"]) - 16:len(fx["std_voxels"])]
    sc = g.scheme
    got = cone.table_adjoint_at(y, g.views, g.volume_size, g.detector_size, sc.axis_mode,
                                sc.box_lo, sc.box_hi, sc.margin, sc.scale,
                                fx["std_voxels"][pick])[0].numpy()
    assert np.abs(got - fx["std_adjoint"][pick]).max() <= 1e-12 * np.abs(fx["std_adjoint"]).max()


def test_workload_fixture_table_forwards_reproduce_on_a_subset():
    """The oracle still gives the stored A of the Beijing scan and of the
    Standard options geometry at 32 of the stored rays each, bitwise."""
    from oracle import cone
    from dtvproj import phantom
    m = _wgen()
    with np.load(WFIX) as f:
        fx = {k: f[k] for k in f.files if k.startswith(("bjf_", "stdf_"))}
    for key, g, seed in (("bjf", m.bj_geometry(), 2043), ("stdf", m.std_geometry(), 2044)):
        assert np.array_equal(fx[f"{key}_rays"], m.table_rays_pick(g, 4096, seed))
        X, Y, Z = g.volume_size
        x = phantom.gaussian((1, 1, Z, Y, X), 40).double()
        assert abs(float(x.sum()) - float(fx[f"{key}_vol_sum"])) <= 1e-12 * float(
            fx[f"{key}_vol_sumsq"])
        pick = fx[f"{key}_rays"][np.r_[0:16, 4080:4096]]
        sc = g.scheme
        got = cone.table_forward_at(x, g.views, g.volume_size, g.detector_size, sc.axis_mode,
                                    sc.box_lo, sc.box_hi, sc.margin, sc.scale, pick)[0].numpy()
        assert np.array_equal(got, fx[f"{key}_forward"][np.r_[0:16, 4080:4096]]), key
