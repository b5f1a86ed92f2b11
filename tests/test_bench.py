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
