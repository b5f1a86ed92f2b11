"""The fast tail's shared-reciprocal quotients against the compiler's IEEE division on the GPU, bit for bit, and the
whole fast tail against the same formula with plain divisions (tests/native/divcheck.hip, built by build())."""
import subprocess

import pytest

from conftest import REPO

EXE = REPO / "tests" / "native" / "_build" / "divcheck"


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
def test_shared_reciprocal_division_is_ieee(seed):
    assert EXE.exists(), "tests/native/_build/divcheck is built by __graft_entry__.build()"
    out = subprocess.run([str(EXE), str(seed), "16384"], capture_output=True, text=True, timeout=120)
    nq, bq, nt, bt = (int(x) for x in out.stdout.split())
    assert nq == nt == 16384 * 256 * 64
    assert bq == 0 and bt == 0, out.stdout
    assert out.returncode == 0
