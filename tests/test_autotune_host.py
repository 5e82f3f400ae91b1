"""Autotuner host logic (no GPU): the selection rule, the generated C table
(round trip, and the committed table's rows use only tried values)."""
import os

import pytest

from oglfusionkinectv2_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "oglfusionkinectv2_amd", "csrc", "kf_tuned.inc")


def test_choose_keeps_the_current_value_within_noise():
    assert autotune.choose({1: 1.00, 2: 0.995, 4: 1.2}, 1, tol=0.01) == 1
    assert autotune.choose({1: 1.00, 2: 0.95, 4: 1.2}, 1, tol=0.01) == 2
    assert autotune.choose({1: 1.00, 2: 0.95}, 7, tol=0.01) == 2  # current not measured


def test_candidates_fit_the_volume():
    assert max(autotune.candidates("integrate_zlen", 128)) == 128
    assert 0 in autotune.candidates("integrate_zlen", 64)
    for k in autotune.DEFAULTS:
        assert autotune.DEFAULTS[k] in autotune.candidates(k, 512)


def test_inc_round_trip(tmp_path):
    rows = [dict(storage="f32", volume=512, stage_ms={"integrate": 0.1, "raycast": 0.2},
                 chosen=dict(integrate_zlen=64, integrate_rows=2, raycast_xcd_group=8, raycast_lanes=1,
                             brick_rebuild_period=32, integrate_columns=0, raycast_hybrid=0)),
            dict(storage="f16", volume=128, stage_ms={"integrate": 0.01, "raycast": 0.05},
                 chosen=dict(autotune.DEFAULTS))]
    p = tmp_path / "t.inc"
    autotune.write_inc(rows, str(p))
    back = autotune.read_inc(str(p))
    assert [(r["storage"], r["volume"]) for r in back] == [("f32", 512), ("f16", 128)]
    for r, b in zip(rows, back):
        for k, v in r["chosen"].items():
            assert b[k] == v


def test_committed_table_uses_tried_values():
    rows = autotune.read_inc(INC)
    for r in rows:
        for k in autotune.DEFAULTS:
            assert r[k] in autotune.candidates(k, r["volume"]), (r, k)
