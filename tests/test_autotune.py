"""CPU: the schedule autotuner's selection logic (no kernels run: the
candidates are host callables and the timer is injected)."""

import pytest

from subsetix_amd import autotune


def _fake_timer(costs):
    """Timer that reports costs[k] for the k-th timed candidate."""
    calls = []

    def timer(fn, reps, bound=None):
        fn()
        calls.append(fn)
        return costs[len(calls) - 1]

    return timer, calls


def test_bucket_half_octaves():
    assert autotune.bucket(0) == 0.0
    assert autotune.bucket(1) == 0.0
    assert autotune.bucket(1024) == 10.0
    assert autotune.bucket(1448) == 10.5  # 2^10.5 = 1448.2
    assert autotune.bucket(1500) == 10.5
    assert autotune.bucket(1800) == 11.0


def test_tune_picks_fastest_and_caches():
    timer, calls = _fake_timer([5.0, 2.0, 9.0])
    t = autotune.Autotuner(warmup=1, timer=timer)
    ran = []
    cands = {"seg": lambda: ran.append("seg"), 4: lambda: ran.append(4), 16: lambda: ran.append(16)}
    t.enabled = True
    assert t.pick(("upwind", 1), lambda: cands) == 4
    # warm-up run and timed run of each candidate, in the factory's order
    assert ran == ["seg", "seg", 4, 4, 16, 16]
    assert t.table[("upwind", 1)].times_us == {"seg": 5.0, "4": 2.0, "16": 9.0}
    # later calls answer from the table without running anything
    t.enabled = False
    assert t.pick(("upwind", 1), lambda: pytest.fail("factory called for a cached key")) == 4
    assert len(calls) == 3
    # a second key of the family: no warm-up runs (each candidate ran before)
    timer2, _ = _fake_timer([1.0, 2.0, 3.0])
    t._timer = timer2
    t.enabled = True
    ran.clear()
    assert t.pick(("upwind", 2), lambda: cands) == "seg"
    assert ran == ["seg", 4, 16]


def test_nearest_bucket_when_not_tuning():
    t = autotune.Autotuner(timer=lambda fn, reps, bound: 0.0)
    t.table[("grad_tag", 4, 20.0, 5.0)] = autotune.Entry(16)
    t.table[("grad_tag", 4, 24.0, 5.0)] = autotune.Entry(64)
    t.table[("grad_tag", 8, 20.5, 5.0)] = autotune.Entry(4)
    factory = lambda: pytest.fail("tuned while disabled")
    assert t.pick(("grad_tag", 4, 20.5, 5.5), factory) == 16   # distance 1.0
    assert t.pick(("grad_tag", 4, 23.0, 5.0), factory) == 64   # distance 1.0 vs 3.0
    assert t.pick(("grad_tag", 4, 22.0, 6.5), factory) is None  # nothing within two octaves
    assert t.pick(("grad_tag", 2, 20.0, 5.0), factory) is None  # discrete components must agree


def test_event_timer_bound_cuts_reps(monkeypatch):
    calls = []
    t = autotune.Autotuner(reps=5, warmup=0, timer=lambda fn, reps, bound: calls.append(bound) or 1.0)
    t.enabled = True
    t.pick(("k", 1), lambda: {"a": lambda: None, "b": lambda: None})
    assert calls[0] == float("inf") and calls[1] == 2.0


def test_ties_keep_first_candidate():
    timer, _ = _fake_timer([3.0, 3.0, 3.0])
    t = autotune.Autotuner(warmup=0, timer=timer)
    t.enabled = True
    assert t.pick(("k",), lambda: {-1: lambda: None, -2: lambda: None, 64: lambda: None}) == -1


def test_disabled_tuner_returns_none():
    t = autotune.Autotuner(timer=lambda fn, reps, bound: pytest.fail("timed while disabled"))
    assert t.pick(("setop", 0), lambda: {4: lambda: None}) is None


def test_empty_candidates_rejected():
    t = autotune.Autotuner(timer=lambda fn, reps, bound: 0.0)
    t.enabled = True
    with pytest.raises(ValueError):
        t.pick(("x",), lambda: {})


def test_tuning_context_restores_state():
    was = autotune.enabled()
    with autotune.tuning() as t:
        assert t is autotune.default() and autotune.enabled()
        with autotune.tuning(False):
            assert not autotune.enabled()
        assert autotune.enabled()
    assert autotune.enabled() == was


def test_save_load_round_trip(tmp_path):
    timer, _ = _fake_timer([4.0, 1.0])
    t = autotune.Autotuner(warmup=0, timer=timer)
    t.enabled = True
    key = ("transfer", 1, 4, 1, 2, True, 18.5, 21.0)
    t.pick(key, lambda: {"seg": lambda: None, 16: lambda: None})
    path = str(tmp_path / "tune.json")
    t.save(path)
    u = autotune.Autotuner()
    assert u.load(path) == 1
    assert u.lookup(key) == 16
    assert u.table[key].times_us == {"seg": 4.0, "16": 1.0}
