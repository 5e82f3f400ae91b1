"""bench.py's frame accounting under the arguments the benchmark is run with
(no GPU): the sequence it generates covers every leg, the traffic of at
least one frame is counted, and a failing leg is recorded instead of ending
the run."""
import json
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = [
    (["--gpus", "1", "--steps", "20", "--warmup", "5"]),
    (["--gpus", "1", "--steps", "1", "--warmup", "0"]),
    ([]),
    (["--steps", "400", "--warmup", "20", "--cpu-frames", "500"]),
    (["--steps", "20", "--warmup", "5", "--no-cpu"]),
]


@pytest.mark.parametrize("argv", CASES)
def test_generated_frames_cover_every_leg(argv):
    a = bench.parse(argv)
    plan = bench.frame_plan(a.warmup, a.steps, a.cpu_frames, a.cpu_warm, not a.no_cpu)
    # the GPU legs index frames 0 .. warmup + steps - 1
    assert plan["generate"] >= a.warmup + a.steps
    # the CPU oracle indexes frames 0 .. cpu_warm + cpu_frames - 1
    if not a.no_cpu:
        assert plan["generate"] >= a.cpu_warm + a.cpu_frames
    sched = bench.count_schedule(a.steps)
    assert sched, "no frame would have its traffic counted"
    assert all(0 <= i < a.steps for i in sched)


def test_cpu_baseline_refuses_a_short_sequence():
    import numpy as np
    with pytest.raises(ValueError):
        bench.cpu_baseline(np.zeros((3, 424, 512), np.uint16), np.tile(np.eye(4), (3, 1, 1)), 64, 5)


def test_run_leg_records_errors_and_continues():
    out = {}

    def boom():
        raise IndexError("index 25 is out of bounds")

    r = bench.run_leg(out, "cpu_baseline", boom)
    assert "IndexError" in r["error"] and out["cpu_baseline"] is r
    assert bench.run_leg(out, "ok", lambda: 3) == 3


def test_plan_mode_with_the_benchmark_arguments():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "1", "--steps", "20",
                        "--warmup", "5", "--plan"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["plan"] == {"gpu_frames": 25, "cpu_frames": 62, "generate": 62}
    assert line["count_frames"] == [19]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_cpu_baseline_runs_on_a_small_volume(seq20):
    depth, poses = seq20
    fps, threads, dt = bench.cpu_baseline(depth[:4], poses[:4], 32, 2, warm=2)
    assert fps > 0 and threads >= 1 and dt > 0
