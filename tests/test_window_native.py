"""Host build of the shared geometry header: the exact fixed-point sample
window (``dtv::exact_window``) against brute force and an int64 bisection.

The brick kernels march only the samples of this window, so an off-by-one
here would drop or double-count samples at brick faces; the checker covers
the cardinal Standard angles (direction components far below 2^-33 that
round to a zero fixed-point step) and far-away synthetic rays with up to
2^20 samples (the re-centred solve)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "window_check.cpp")
INC = os.path.join(ROOT, "dtv-net_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_exact_window_native(tmp_path):
    exe = str(tmp_path / "window_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", INC, "-o", exe, SRC], check=True)
    out = subprocess.run([exe], stdout=subprocess.PIPE, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    assert out.stdout.startswith("ok"), out.stdout
    cases, hits = map(int, out.stdout.split()[1:3])
    assert cases > 200000 and hits > 100000
