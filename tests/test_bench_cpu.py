"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message and exit status 2 before any GPU work; --help lists the
workload switches."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    """bench.py as a module (its top level only defines functions)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_help_lists_switches():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--no-batch", "--no-scatter", "--no-emap", "--tof-window"):
        assert flag in r.stdout


def test_traffic_counter_csv(tmp_path):
    """The live traffic figure's parsing: per-kernel means of one counter and
    the first device-copy dispatch (the 1 GiB probe), in dispatch order."""
    bench = _bench()
    rows = [("Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"),
            ("7", "__amd_rocclr_copyBuffer", "FETCH_SIZE", "999"),       # a later copy
            ("2", "__amd_rocclr_copyBuffer", "FETCH_SIZE", "524288"),    # the probe
            ("11", "srf::k_bp_tiled<true, false, false>(...)", "FETCH_SIZE", "100"),
            ("12", "srf::k_bp_tiled<true, false, false>(...)", "FETCH_SIZE", "300"),
            ("13", "srf::k_bp_tiled<true, false, false>(...)", "WRITE_SIZE", "5")]
    import csv
    p = tmp_path / "c.csv"
    with open(p, "w", newline="") as f:
        csv.writer(f).writerows(rows)
    means, first_copy = bench._counter_means(str(p), "FETCH_SIZE")
    assert first_copy == 524288.0
    assert means["srf::k_bp_tiled<true, false, false>(...)"] == 200.0
    # 1 GiB / (524288 KiB * 1024) = 2 bytes moved per counted byte
    assert bench.TRAFFIC_COPY / (first_copy * 1024.0) == 2.0
