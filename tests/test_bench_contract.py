"""The bench line's contract, checked on the host (no GPU): the one-GPU
guard of --gpus and the keys, types and fixed fields of every JSON line."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

KEYS = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
        "vs_baseline", "dtype", "data", "config")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line for a run that did not happen


def test_line_carries_the_contract_keys():
    line = bench._line("m", 1.23456789, "GB/s", 2.5, 200, 20, "f32", {"workload": "w"})
    assert tuple(line) == KEYS
    assert line["n_gpus"] == 1 and line["data"] == "synthetic" and line["vs_baseline"] is None
    assert line["value"] == 1.2346 and line["ms_per_step"] == 2.5
    assert line["higher_is_better"] is True and line["scaling"] == "weak"
    assert json.loads(json.dumps(line)) == line
    t = bench._line("t", 3.0, "s", 3000.0, 1, 0, "f64", {"workload": "w"}, higher=False, scaling="strong")
    assert t["higher_is_better"] is False and t["scaling"] == "strong"



def test_headline_traffic_comes_from_a_committed_profile():
    # roofline.traffic of the default line is the measured HBM bytes per
    # launch recorded in profiles/ (FETCH_SIZE x 2 + WRITE_SIZE)
    for role in ("fp32_fused", "fp64_fused"):
        t, src = bench.profile_traffic(bench.DEFAULT, role)
        assert isinstance(t, int) and t > 10 ** 9, (role, t)
        assert src.startswith("profiles" + os.sep) and os.path.isfile(os.path.join(ROOT, src))
