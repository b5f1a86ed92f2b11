"""The engine's logarithm (tn93g_math.h, built for the host) against the running libm log(), CPU only.

Outside the uncertain band around rounding midpoints its rounded value must equal libm's, and inside the band libm
must return one of the two candidates.  The widest misrounding of libm seen on the sample (distance of the exact
value from the midpoint, in ulps) must stay well inside the band.
"""
import subprocess

import pytest

from conftest import REPO


@pytest.fixture(scope="module")
def logcheck(tmp_path_factory):
    exe = tmp_path_factory.mktemp("log") / "logcheck"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-o", str(exe),
                    str(REPO / "tests" / "native" / "logcheck.cpp")], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_log_matches_libm(logcheck, seed):
    out = subprocess.run([str(logcheck), "3000000", str(seed), "0.05"], capture_output=True, text=True)
    n, flagged, bad_unflagged, bad_flagged, max_mis = out.stdout.split()
    assert out.returncode == 0, out.stderr
    assert int(bad_unflagged) == 0 and int(bad_flagged) == 0
    assert float(max_mis) < 0.025          # libm's worst misrounding seen: ~0.0175 ulp from the midpoint
    assert int(flagged) < 0.12 * int(n)    # the band is 2 x 0.05 ulp wide


def test_window_is_needed(logcheck):
    """With no band at all, libm's value differs from the correctly rounded one on some inputs."""
    out = subprocess.run([str(logcheck), "3000000", "3", "0.0"], capture_output=True, text=True)
    assert out.returncode == 1
    assert int(out.stdout.split()[2]) > 0


@pytest.fixture(scope="module")
def logfast(tmp_path_factory):
    exe = tmp_path_factory.mktemp("logf") / "logfast"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-o", str(exe),
                    str(REPO / "tests" / "native" / "logfast.cpp")], check=True)
    return exe


@pytest.mark.parametrize("seed", [4, 5])
def test_fast_log_encloses_libm(logfast, seed):
    """The fast path's logarithm: libm's value always lies in its enclosure; its error stays under the proven
    8 u |v| (measured against the double-double logarithm) and the enclosure is 16 u |v| either side (the rounding of
    its ends adds at most one ulp, 2 u |v|)."""
    out = subprocess.run([str(logfast), "2000000", str(seed)], capture_output=True, text=True)
    n, outside, max_err, max_hw = out.stdout.split()
    assert out.returncode == 0, out.stderr
    assert int(outside) == 0
    assert float(max_err) < 8.0
    assert 15.0 <= float(max_hw) <= 18.0
