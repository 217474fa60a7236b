"""bench.py pieces that run without a GPU: the one-GPU refusal, the option
parser, and the HBM-traffic lookup in the committed counter summaries."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _bench(*args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=120, env=e)


@pytest.mark.parametrize("args,env", [(["--gpus", "2"], {}), (["--gpus", "8", "--steps", "1"], {}),
                                      ([], {"WORLD_SIZE": "4"})])
def test_more_than_one_gpu_is_refused(args, env):
    r = _bench(*args, env=env)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_help_lists_the_contract_options():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--batch-shots", "--batch-graph", "--pass-steps",
                "--batch-real-steps", "--fwi-scale"):
        assert opt in r.stdout, opt


def test_measured_traffic_of_the_single_shot_rtm_shape():
    """The default RTM shape's per-launch traffic is the six-launch weighted
    mean of its deferred-imaging instances, read from the committed summary;
    an unknown shape has none."""
    import bench
    key = "step_kernel<BWD_RTM,REF>_BY=16_RPT=1_SB=1"
    traffic, src = bench.measured_traffic(key)
    assert src is not None and src.startswith("profiles" + os.sep)
    found = {}
    for line in open(os.path.join(ROOT, src)):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic":
            found[parts[1]] = int(parts[2])
    per = [found[key + "_IMG=%d" % i] for i in (0, 3, 4, 5)]
    assert traffic == (2 * per[0] + 2 * per[1] + per[2] + per[3]) // 6
    # the measured bytes lie between the compulsory 28 B (no update) and 44 B
    # (both updates) per cell of the 493 x 493 x 293 updated region
    cells = 493 * 493 * 293
    assert 28 * cells < traffic < 44 * cells * 1.3
    assert bench.measured_traffic("step_kernel<BWD_RTM,REF>_BY=3_RPT=7_SB=1") == (None, None)
