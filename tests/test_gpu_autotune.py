"""The launch autotuner on the device (GPU): coordinate descent over pool size, jump cap, relaunch
batching and chain length on the slab-with-inclusion scene (2e5-packet sample), stored in a scratch
cache.  Bars, written here: the stored setting is the fastest candidate timed, it is read back by
`tuned`, it is no slower than the engine defaults it started from, and a run with it tallies the
full budget."""
import json

import numpy as np
import pytest

from ymc3d_amd import autotune, configs

pytestmark = pytest.mark.gpu


def test_tune_stores_the_fastest_setting(engine, tmp_path, capsys):
    scene = configs.skin_slab_sphere(photons=200_000)
    cache = str(tmp_path / "tune.json")
    best = autotune.tune(engine, scene, repeats=1, max_sweeps=2, cache=cache, verbose=True)
    lines = [json.loads(l) for l in capsys.readouterr().out.splitlines() if l.startswith("{")]
    assert best["trials"] == len(lines) >= sum(len(v) for v in autotune.SEARCH.values()) - len(autotune.SEARCH)
    fastest = min(l["kernel_ms"] for l in lines)
    default = lines[0]["kernel_ms"]  # the descent starts at the defaults
    assert best["kernel_ms"] <= default
    assert abs(best["kernel_ms"] - fastest) <= 0.003 * default + 1e-9  # within the descent's gain threshold
    stored = autotune.tuned(scene, scene.params, cache=cache)
    keys = ("pool_lanes", "jump_max_crossings", "relaunch_lanes", "relaunch_wait", "chain_lanes")
    assert {k: best[k] for k in keys} == stored
    res = engine.simulate(**scene.params, **stored)
    assert int(res.totals["photons"].sum()) == 200_000 and np.all(res.R_w >= 0)
    # A second tune of the same workload also descends from the stored setting: plant one away from
    # the defaults' path, and the planted setting is among the timed candidates.
    planted = dict(stored, relaunch_lanes=20, relaunch_wait=8)
    db = json.load(open(cache))
    db[autotune.device_name()][autotune.signature(scene, scene.params)]["params"] = planted
    json.dump(db, open(cache, "w"))
    again = autotune.tune(engine, scene, repeats=1, max_sweeps=1, cache=cache, verbose=True)
    timed = [json.loads(l) for l in capsys.readouterr().out.splitlines() if l.startswith("{")]
    assert any(all(t[k] == planted[k] for k in keys) for t in timed)
    assert again["kernel_ms"] <= min(t["kernel_ms"] for t in timed) * 1.003 + 1e-9
