"""Autotuner decisions without a GPU: deterministic choice (fastest, earliest
on ties), cached decisions reused without timing, stable workload keys."""
import json

import pytest
import torch


class _S:
    def __init__(self, n, shape):
        self.lors = torch.zeros((n, 7))
        self.shape = shape


def test_choose_fastest_first_on_ties():
    from srf_amd.autotune import CANDIDATES, choose
    t = {i: 1.0 for i in range(len(CANDIDATES))}
    assert choose(t) == 0
    t[5] = 0.5
    t[7] = 0.5
    assert choose(t) == 5
    with pytest.raises(ValueError):
        choose({})


def test_tune_is_deterministic_and_cached(tmp_path):
    from srf_amd.autotune import Autotuner, CANDIDATES
    scans = [_S(1000, (96, 96, 64)), _S(50_000, (256, 256, 128))]
    fake = {c: 1.0 + 0.01 * ((7 * i) % 5) for i, c in enumerate(CANDIDATES)}
    calls = []

    def timer(c):
        calls.append(c)
        return fake[c]

    path = str(tmp_path / "cache.json")
    a = Autotuner(path).tune(scans, 8, 3.66, 600.0, True, timer=timer)
    assert len(calls) == len(CANDIDATES)
    want = min(range(len(CANDIDATES)), key=lambda i: (fake[CANDIDATES[i]], i))
    assert a == CANDIDATES[want]
    b = Autotuner(path).tune(scans, 8, 3.66, 600.0, True, timer=lambda c: 1 / 0)
    assert b == a
    d = json.load(open(path))
    assert len(d) == 1 and list(d.values())[0]["config"] == a.__dict__
    # a different workload is a different key
    c = Autotuner(path).tune(scans[:1], 8, 3.66, 600.0, True, timer=lambda c: 1.0)
    assert c == CANDIDATES[0]
    assert len(json.load(open(path))) == 2


def test_signature_ignores_scan_order_and_rounds_grids():
    from srf_amd.autotune import signature
    a = [_S(1000, (96, 96, 64)), _S(3000, (256, 256, 128))]
    b = [_S(3000, (250, 256, 128)), _S(1000, (96, 96, 64))]
    assert signature(a, True, "dev") == signature(b, True, "dev")
    assert signature(a, False, "dev") != signature(a, True, "dev")


@pytest.mark.parametrize("content", ["", "{\"a\": ", "[1, 2]", "not json at all"])
def test_corrupt_cache_is_ignored_and_replaced(tmp_path, content):
    """An unreadable cache file (empty, truncated, not an object, not JSON) is
    ignored with a warning; the workload is timed again and the file rewritten
    whole."""
    from srf_amd.autotune import Autotuner, CANDIDATES
    path = tmp_path / "cache.json"
    path.write_text(content)
    with pytest.warns(UserWarning, match="unreadable"):
        t = Autotuner(str(path))
    assert t.cache == {} and t.cache_error
    scans = [_S(1000, (96, 96, 64))]
    got = t.tune(scans, 8, 3.66, 600.0, True, timer=lambda c: 2.0 - 0.1 * (c == CANDIDATES[3]))
    assert got == CANDIDATES[3]
    d = json.load(open(path))
    assert list(d.values())[0]["config"] == CANDIDATES[3].__dict__


def test_malformed_entry_is_retimed(tmp_path):
    from srf_amd.autotune import Autotuner, CANDIDATES, signature
    scans = [_S(1000, (96, 96, 64))]
    key = signature(scans, True)
    path = tmp_path / "cache.json"
    for entry in ({"config": {"item_records": 777}}, {"config": {"bogus": 1}}, {"nothing": 0}):
        path.write_text(json.dumps({key: entry}))
        with pytest.warns(UserWarning, match="timing again"):
            got = Autotuner(str(path)).tune(scans, 8, 3.66, 600.0, True, timer=lambda c: 1.0)
        assert got == CANDIDATES[0]
        assert json.load(open(path))[key]["config"] == CANDIDATES[0].__dict__


def test_timing_failure_falls_back_to_default(tmp_path):
    from srf_amd.autotune import Autotuner
    from srf_amd.batch import BatchConfig

    def broken(c):
        raise RuntimeError("HIP error")

    with pytest.warns(UserWarning, match="default schedule"):
        got = Autotuner(str(tmp_path / "c.json")).tune([_S(10, (96, 96, 64))], 8, 3.66, 600.0, True, timer=broken)
    assert got == BatchConfig()
    assert not (tmp_path / "c.json").exists()
