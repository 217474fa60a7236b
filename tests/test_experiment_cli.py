"""The FWI time-to-misfit experiment's command line (no GPU): which phase
needs which inputs, and how baselines from separate runs are merged."""
import json

import pytest

from fwi3d_amd import experiment


def _write(tmp_path, name, obj):
    p = tmp_path / name
    p.write_text(json.dumps(obj))
    return str(p)


@pytest.mark.parametrize("argv", [["--phase", "learned"], ["--phase", "fwi"],
                                  ["--phase", "learned", "--load-net", "net.npy"]])
def test_phases_that_need_inputs(argv, capsys):
    with pytest.raises(SystemExit) as e:
        experiment.main(argv)
    assert e.value.code == 2
    assert "needs" in capsys.readouterr().err


def test_baselines_of_another_scale_or_seed_are_refused(tmp_path, capsys):
    other = _write(tmp_path, "b.json", {"scale": "tiny", "seed": 0, "plain": {}, "guarded_plain": {}})
    with pytest.raises(SystemExit):
        experiment.main(["--phase", "learned", "--load-net", "n.npy", "--baselines", other])
    assert "another scale or seed" in capsys.readouterr().err
    seed = _write(tmp_path, "c.json", {"scale": "small", "seed": 3, "plain": {}, "guarded_plain": {}})
    with pytest.raises(SystemExit):
        experiment.main(["--phase", "learned", "--load-net", "n.npy", "--baselines", seed])


def test_baselines_must_hold_both_runs(tmp_path, capsys):
    only = _write(tmp_path, "p.json", {"scale": "small", "seed": 0, "plain": {"misfit": [1.0]}})
    with pytest.raises(SystemExit):
        experiment.main(["--phase", "learned", "--load-net", "n.npy", "--baselines", only])
    assert "both the plain and the guarded plain" in capsys.readouterr().err


def test_baselines_of_two_runs_are_merged(tmp_path, monkeypatch, capsys):
    """A plain run and a guarded-plain run made separately reach the learned
    phase as one baselines dict; the other keys come from the first file."""
    seen = {}

    def fake_run(scale, seed, log=None, phase="all", baselines=None, **kw):
        seen.update(scale=scale, seed=seed, phase=phase, baselines=baselines, **kw)
        return {"ok": True}

    monkeypatch.setattr(experiment, "run", fake_run)
    a = _write(tmp_path, "a.json", {"scale": "small", "seed": 0, "grid": [96, 96, 48], "plain": {"seconds": 3.5}})
    b = _write(tmp_path, "b.json", {"scale": "small", "seed": 0, "grid": [1, 1, 1],
                                    "guarded_plain": {"seconds": 5.8}, "plain": {"seconds": 99.0}})
    assert experiment.main(["--phase", "learned", "--load-net", "n.npy", "--baselines", a + "," + b]) == 0
    base = seen["baselines"]
    assert base["guarded_plain"] == {"seconds": 5.8}
    assert base["plain"] == {"seconds": 99.0}  # a later file's run replaces an earlier one's
    assert base["grid"] == [96, 96, 48]
    assert seen["phase"] == "learned" and seen["tune"] is True
    assert json.loads(capsys.readouterr().out) == {"ok": True}
