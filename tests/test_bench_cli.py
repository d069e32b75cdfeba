"""bench.py's contract where it needs no GPU: one GPU only, the whole-step
CPU baseline leg and its JSON line, and the workload arithmetic the roofline
object uses (tree depth and compulsory bytes at the benchmark's sizes)."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def run(args, timeout=300):
    return subprocess.run([sys.executable, str(ROOT / "bench.py")] + args, capture_output=True,
                          text=True, timeout=timeout)


def test_more_than_one_gpu_is_refused():
    r = run(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpu_leg_times_whole_steps():
    r = run(["--only", "cpu", "--n", "3000", "--cpu-steps", "2", "--config1-n", "1500",
             "--config1-steps", "1", "--cpu-threads", "2", "--cpu-repeats", "3"])
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert set(out) == {"n3000_steps2", "n1500_steps1"}
    for key, (n, steps) in (("n3000_steps2", (3000, 2)), ("n1500_steps1", (1500, 1))):
        c = out[key]
        assert c["kind"] in ("reference", "port")
        assert (c["n"], c["steps"], c["repeats"]) == (n, steps, 3)
        assert len(c["values"]) == 3 and c["value"] == sorted(c["values"])[1]  # the median
        assert c["unit"] == "particle-steps/s" and c["value"] > 0
        assert f"{steps} whole step(s) of all {n} bodies" in c["sample"]


def test_workload_arithmetic():
    # the floor(n/2) recursion: depth 21 at 10 M bodies and 18 at 1 M (leaf 7),
    # 14 at 100 k; compulsory bytes 104 + 16 D + 20 per particle-step
    assert bench.tree_depth(10_000_000, 7) == 21
    assert bench.tree_depth(1_000_000, 7) == 18
    assert bench.tree_depth(100_000, 7) == 14
    assert bench.tree_depth(7, 7) == 0 and bench.tree_depth(8, 7) == 1
    assert bench.compulsory_bytes(10_000_000, 7) == 104 + 16 * 21 + 20
