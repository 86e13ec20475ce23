"""CPU checks of bench.py's contract: the conv2 HBM traffic it reports comes
from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, scaled to the
patches of the launch it describes, and a request for more than one GPU
exits with a message before any device work."""
import importlib.util
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_conv2_traffic_from_committed_profile():
    bench = _bench()
    text = open(os.path.join(ROOT, "profiles", "hbm_traffic_descriptor_pipeline_fp32.md")).read()
    m = re.search(r"conv2 \(`[^`]*`\): (\d+) MB read and (\d+) MB written per (\d+)-patch launch", text)
    assert m, "the summary states conv2's read and written MB per launch"
    read_mb, write_mb, patches = (int(g) for g in m.groups())
    # the fp32 conv2 launch reads its 28^3 x 64-channel input about once
    input_mb = patches * 28 ** 3 * 64 * 4 / 1e6
    assert 1.0 <= read_mb / input_mb < 1.2
    per_launch = bench.measured_traffic("fp32", patches)
    assert per_launch == (read_mb + write_mb) * 1e6
    # scaled linearly to the average launch of the benchmark (1000 patches)
    assert abs(bench.measured_traffic("fp32", 1000) - per_launch * 1000 / patches) < 1.0
    # no summary for a precision -> no traffic figure
    assert bench.measured_traffic("bf16", 1000) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
