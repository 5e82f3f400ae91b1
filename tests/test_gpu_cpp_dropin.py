"""The C++ gFusion drop-in (csrc/gfusion.hpp) driving the reference frame
loop (kinect.cpp:654-695) through the C ABI: built by build(), run here."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oglfusionkinectv2_amd",
                   "kf_gfusion_example")


def test_cpp_frame_loop_tracks():
    print("case: C++ drop-in 20 frames 128^3", flush=True)
    assert os.path.exists(EXE), "C++ example not built"
    r = subprocess.run([EXE, "20", "128", "1"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    rows = np.array([[float(v) for v in l.split()] for l in r.stdout.strip().splitlines()])
    assert rows.shape == (20, 6)
    assert rows[0, 1] == 0 and (rows[1:, 1] == 1).all()
    assert rows[:, 5].max() < 50.0  # mm
