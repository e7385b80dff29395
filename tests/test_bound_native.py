"""Host build of the brick kernels' per-(brick, view) staging
(``dtv::box_view_stage``): the matched adjoint's fixed-point overflow bound
and the candidate-pixel rectangle against brute force.

The int32 LDS accumulators of the matched A^T cannot overflow only if the
staged weight bound is at least the largest sum of trilinear weights any
owned voxel collects from one view; tests/native/bound_check.cpp enumerates
every ray of the detector with the kernels' own set-up (ray_core, 32.32
fixed point, exact_window) over circular orbits at magnification 1.1 to 10,
tilted detectors with an off-plane source and sources next to the volume,
on bricks from 8^3 to 32x32x8 (edge bricks included), and checks the bound
and that no ray outside the rectangle reaches the brick.  Each brick is also
checked as the partitioned adjoint stages it (samples whose floor cell is the
brick's own, the upper-halo voxels read back too)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "bound_check.cpp")
INC = os.path.join(ROOT, "dtv-net_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fixed_point_bound_and_rectangle_native(tmp_path):
    exe = str(tmp_path / "bound_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", INC, "-o", exe, SRC], check=True)
    out = subprocess.run([exe], stdout=subprocess.PIPE, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert out.stdout.startswith("ok"), out.stdout
    checks, inf_views = map(int, out.stdout.split()[1:3])
    ratio = float(out.stdout.split()[3])
    # every staged finite bound was checked, and some views took the FP32 path
    assert checks > 1000 and inf_views > 0
    assert 0.0 < ratio <= 1.0
