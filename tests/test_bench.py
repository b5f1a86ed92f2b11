"""bench.py's command-line contract (CPU only: nothing here touches a GPU)."""
import importlib.util
import subprocess
import sys

from conftest import REPO


def _bench():
    spec = importlib.util.spec_from_file_location("bench", REPO / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_reference_flags_follow_the_mode():
    """The CPU baseline runs the reference with the flags of the workload the GPU line measures: -g 0.05 only with
    resolve (tn93's default -g otherwise), -t 0.015 -l 500 and the CSV format always."""
    b = _bench()
    assert b.mode_flags("resolve") == ["-t", "0.015", "-a", "resolve", "-g", "0.05", "-l", "500", "-f", "csv"]
    for m in ("average", "skip", "gapmm"):
        f = b.mode_flags(m)
        assert f == ["-t", "0.015", "-a", m, "-l", "500", "-f", "csv"]
        assert "-g" not in f


def test_cpu_baseline_record(tmp_path):
    """The baseline times the reference on the affinity set's threads (and OMP_NUM_THREADS when it differs) and on 1
    thread, names the build variant and the host, and keeps the sha256 of the reference's outputs for parity."""
    import os

    import pytest
    b = _bench()
    exe, kind = b.ref_exe()
    if exe is None:
        pytest.skip("neither the reference build nor the oracle is built")
    sys.path.insert(0, str(REPO))
    from swifttn93_amd import gen
    fa = tmp_path / "x.fa"
    gen.write_fasta(fa, gen.c1s(40, seed=3))
    cb, keep = b.cpu_baseline([["-t", "0.05", "-l", "500", str(fa)]], 780, "test sample")
    assert cb["cores"] in b.thread_counts() and b.thread_counts()[0] == len(os.sched_getaffinity(0))
    assert cb["one_thread"]["cores"] == 1 and cb["one_thread"]["value"] > 0
    assert cb["kind"] == kind and cb["variant"] and cb["host"]["nproc"] >= 1
    assert len(keep) == 1 and len(keep[0][0]) == 64
