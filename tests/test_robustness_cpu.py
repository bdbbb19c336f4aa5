"""Failure handling without a GPU: checkpoint files (round trip, workload key,
atomic write, safe loading) and the batch's fallback from HIP-graph replay to
eager launches when capture fails."""
import os
import warnings

import pytest
import torch


def test_checkpoint_round_trip_and_key(tmp_path):
    from srf_amd import checkpoint as ckpt
    path = str(tmp_path / "run" / "osem.ckpt")
    key = {"kind": "osem", "n_sub": 4, "shape": [8, 8, 4], "perm": "abc"}
    x = [torch.arange(12, dtype=torch.float32).reshape(3, 4), torch.full((5,), float("nan"))]
    y = [torch.ones(7)]
    assert ckpt.load(path, key) is None                      # nothing yet
    ckpt.save(path, key, 9, x, y)
    assert sorted(os.listdir(tmp_path / "run")) == ["osem.ckpt"]   # no temporary file left
    st = ckpt.load(path, key)
    assert st["done"] == 9
    assert torch.equal(st["images"][0], x[0]) and torch.isnan(st["images"][1]).all()
    assert torch.equal(st["values"][0], y[0])
    with pytest.raises(ValueError, match="another workload"):
        ckpt.load(path, dict(key, n_sub=8))
    # a later save replaces the file whole
    ckpt.save(path, key, 12, x)
    st = ckpt.load(path, key)
    assert st["done"] == 12 and "values" not in st


def test_checkpoint_refuses_foreign_files(tmp_path):
    from srf_amd import checkpoint as ckpt
    p = tmp_path / "other.pt"
    torch.save({"weights": torch.ones(3)}, p)
    with pytest.raises(ValueError, match="not an OSEM checkpoint"):
        ckpt.load(str(p), {"kind": "osem"})
    bad = tmp_path / "garbage.pt"
    bad.write_bytes(b"\x00\x01 not a torch file")
    with pytest.raises(Exception):
        ckpt.load(str(bad), {"kind": "osem"})


def test_tensor_digest_names_content():
    from srf_amd.checkpoint import tensor_digest
    a = torch.randperm(1000, generator=torch.Generator().manual_seed(1))
    b = torch.randperm(1000, generator=torch.Generator().manual_seed(2))
    assert tensor_digest(a) == tensor_digest(a.clone())
    assert tensor_digest(a) != tensor_digest(b)


def _fake_batch():
    from srf_amd.batch import BatchOsem
    b = BatchOsem.__new__(BatchOsem)
    b.graph, b.graph_error, b.done = None, None, 0
    calls = []
    b.iteration = lambda: calls.append("eager")
    return b, calls


def test_capture_failure_falls_back_to_eager():
    b, calls = _fake_batch()
    attempts = []

    def failing_capture():
        attempts.append(1)
        b.graph_error = "forced: capture unsupported"
        raise RuntimeError("forced: capture unsupported")

    b.capture = failing_capture
    with pytest.warns(UserWarning, match="capture failed"):
        b.advance(graph=True)
    assert calls == ["eager"] and b.done == 1
    with warnings.catch_warnings():
        warnings.simplefilter("error")          # no second attempt, no second warning
        b.advance(graph=True)
    assert calls == ["eager", "eager"] and attempts == [1] and b.done == 2


def test_graph_replayed_when_capture_works():
    b, calls = _fake_batch()
    replays = []

    class G:
        def replay(self):
            replays.append(1)

    def capture():
        b.graph = G()
        return b.graph

    b.capture = capture
    b.advance(graph=True)
    b.advance(graph=True)
    b.advance(graph=False)
    assert replays == [1, 1] and calls == ["eager"] and b.done == 3


def test_event_order_refuses_host_tensors():
    """geometric_subset_order runs on the GPU only (srf_event_order): host
    tensors are refused before the library is called, and an empty index
    list comes back as it is."""
    from srf_amd.functions import geometric_subset_order
    with pytest.raises(ValueError, match="on one GPU"):
        geometric_subset_order(torch.zeros(10, 6), torch.arange(10), 2)
    empty = geometric_subset_order(torch.zeros(0, 6), torch.arange(0), 2)
    assert empty.numel() == 0 and empty.dtype == torch.int64
