"""The replay's repeated addition (tn93g_add_repeat, tn93g_math.h built for the host) against the same additions made
one after the other, bit for bit, CPU only."""
import subprocess

import pytest

from conftest import REPO


@pytest.fixture(scope="module")
def addrep(tmp_path_factory):
    exe = tmp_path_factory.mktemp("addrep") / "addrep"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-o", str(exe),
                    str(REPO / "tests" / "native" / "addrep.cpp")], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_add_repeat_equals_sequential_additions(addrep, seed):
    out = subprocess.run([str(addrep), "400000", str(seed), "20000"], capture_output=True, text=True)
    n, bad = (int(x) for x in out.stdout.split())
    assert n == 400000 and bad == 0, out.stderr
    assert out.returncode == 0


MISTAKES = [
    # ties rounded away from zero instead of to even
    ("F = (F >> t & 1ull) ? F + (1ull << (t - 1)) : F - (1ull << (t - 1));", "F = F + (1ull << (t - 1));"),
    # one rounding taken for three powers of two passed (it is right for one or two only)
    ("(int)(tn93g_bits(x) >> 52) > 2) return false;", "(int)(tn93g_bits(x) >> 52) > 3) return false;"),
]


@pytest.mark.parametrize("k", range(len(MISTAKES)))
def test_planted_mistake_is_caught(tmp_path, k):
    """A closed form with a planted mistake must fail the check."""
    src = (REPO / "tests" / "native" / "addrep.cpp").read_text()
    hdr = (REPO / "swifttn93_amd" / "csrc" / "tn93g_math.h").read_text()
    bad = hdr.replace(*MISTAKES[k])
    assert bad != hdr
    (tmp_path / "tn93g_math.h").write_text(bad)
    (tmp_path / "tn93g_logtab.h").write_text((REPO / "swifttn93_amd" / "csrc" / "tn93g_logtab.h").read_text())
    (tmp_path / "addrep.cpp").write_text(src.replace('"../../swifttn93_amd/csrc/tn93g_math.h"', '"tn93g_math.h"'))
    exe = tmp_path / "addrep_bad"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-o", str(exe), str(tmp_path / "addrep.cpp")],
                   check=True)
    out = subprocess.run([str(exe), "200000", "5", "20000"], capture_output=True, text=True)
    assert out.returncode == 1 and int(out.stdout.split()[1]) > 0
