"""FP16 / FP32 TSDF against the FP64 oracle (tests/precision.py): one frame
from pinned FP64 state must stay within the per-frame tolerances (pose
<= 0.1 mm and 0.1 mrad, TSDF |d| <= 1e-3 of truncation outside <= 0.1 % of
voxels, raycast vertex p99 <= 0.5 mm); the free-running differences are
reported and bounded loosely (they include the feedback loop's growth, which
the FP32 oracle shows too)."""
import json

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f16", "f32"])
def test_precision_vs_fp64_oracle(storage):
    from precision import evaluate

    print("precision: 128^3 %s, 12 frames, pinned 4/8/11" % storage, flush=True)
    r = evaluate(volume=128, frames=12, storage=storage, pinned=(4, 8, 11), with_f32_oracle=True)
    print(json.dumps(r))
    for p in r["pinned"]:
        assert p["pose_mm"] <= 0.1 and p["pose_mrad"] <= 0.1, p
        t = p["tsdf"]
        assert t["frac_gt_1e3"] <= 1e-3 and t["weight_differs"] <= 1e-3, p
        assert p["raycast"]["p99_mm"] <= 0.5 and p["raycast"]["hit_flips"] <= 1e-3, p
    fr = r["free_running"]
    assert fr["ate_max_mm"] <= 10 * max(1.0, r["f32_oracle_vs_f64_oracle"]["ate_max_mm"])
