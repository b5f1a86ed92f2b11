"""The dense replay's repeated addition on the GPU (the header's exact and two-crossing cases and closed form,
composed as the replay composes them) against the m FP64 additions made one after the other on the device, bit for
bit (tests/native/addrep_gpu.hip, built by build())."""
import subprocess

import pytest

from conftest import REPO

EXE = REPO / "tests" / "native" / "_build" / "addrep_gpu"


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
def test_repeated_addition_on_gpu(seed):
    assert EXE.exists(), "tests/native/_build/addrep_gpu is built by __graft_entry__.build()"
    out = subprocess.run([str(EXE), str(seed), "1024", "3000"], capture_output=True, text=True, timeout=120)
    n, bad = (int(x) for x in out.stdout.split())
    assert n == 1024 * 256 * 8
    assert bad == 0, out.stdout
    assert out.returncode == 0
