"""Autotuner bookkeeping on the host: keys, cache file round trip, damaged cache."""

from __future__ import annotations

import json

from xla_cco.minimize import tuning


def test_cache_round_trip_and_damaged_file(tmp_path, monkeypatch) -> None:
    path = tmp_path / "t.json"
    monkeypatch.setenv(tuning.CACHE_ENV, str(path))
    tuning.forget()
    assert tuning.cached_choice("lbfgs/d2/n10/m5/b8/fp64/mem100") is None
    choice = tuning.TuneChoice(32, 2, 8, 1.5)
    with tuning._lock:
        tuning._choices["k"] = choice
        tuning._store()
    tuning.forget()
    assert tuning.cached_choice("k") == choice
    assert json.loads(path.read_text())["k"]["shards"] == 2
    path.write_text("{not json")
    tuning.forget()
    assert tuning.cached_choice("k") is None  # ignored, not raised


def test_lane_bucket_is_a_power_of_two() -> None:
    assert [tuning._pow2(x) for x in (1, 2, 3, 64, 65, 1000)] == [1, 2, 4, 64, 128, 1024]
