"""The raycast's exact empty-space fast-forward (csrc/kf_march.hpp) against
plain float stepping, on the host: random starts, steps and targets,
power-of-two steps, round-half-even ties and runs that cross binades."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None and not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("perturb", [False, True], ids=["exact_rcp", "rcp_4ulp"])
def test_march_to_equals_plain_stepping(tmp_path, perturb):
    """perturb: the step-count estimates use a reciprocal off by up to 4 ulps
    (the device uses the hardware reciprocal); results must not change."""
    exe = str(tmp_path / "march_check")
    subprocess.run([HIPCC, "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "oglfusionkinectv2_amd", "csrc")]
                   + (["-DKF_MARCH_RCP_PERTURB"] if perturb else [])
                   + [os.path.join(ROOT, "tests", "march_check.cpp"), "-o", exe], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    r = subprocess.run([exe, "3000000"], stdout=subprocess.PIPE, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
