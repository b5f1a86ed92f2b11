"""tn93g_run_noalloc (every work buffer in the device's fixed scratch array, links written into the caller's arrays)
against tn93g_run on the same inputs: two calls in a row, the second smaller, then the refusals.  Needs an MI355X."""
import sys

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(REPO))
import swifttn93_amd as T  # noqa: E402
from swifttn93_amd import gen  # noqa: E402


def _same(want, got):
    """Integer outputs and link distances bit for bit (the same kernels); the mean's enclosure may differ in its
    last bits (row sums are added with atomics), so both must contain each other's print."""
    assert (want["hist"] == got["hist"]).all()
    for k in ("link_row", "link_col", "link_dist", "link_lo", "link_hi"):
        assert np.array_equal(want[k], got[k]), k
    for k in ("found_links", "skipped", "weighted", "max_lo", "max_hi", "n_replayed", "n_uncertain"):
        assert want[k] == got[k], k
    assert got["mean_lo"][0] <= want["mean_hi"][0] and want["mean_lo"][0] <= got["mean_hi"][0]
    assert "%.6g" % want["mean_lo"][0] == "%.6g" % got["mean_lo"][0]
    assert "%.6g" % want["mean_hi"][0] == "%.6g" % got["mean_hi"][0]


def test_noalloc_matches_run():
    eng = T.Engine(0)
    try:
        cases = [
            # 500 x 1,300 HIV-pol-shaped (B/D/H/V codes: replayed pairs), -a resolve -g 0.05
            (gen.c1s(500, seed=151), dict(mode=T.RESOLVE, threshold=0.03, min_overlap=500, resolve_fraction=0.05)),
            # then a smaller set with every pair written, -a average
            (gen.c0(seed=152, N=200, L=700), dict(mode=T.AVERAGE, threshold=1.0, min_overlap=300)),
        ]
        for k, (codes, kw) in enumerate(cases):
            print(f"noalloc call {k}: {codes.shape}", flush=True)
            s = eng.upload(codes)
            want = eng.run(s, **kw)
            got = eng.run_noalloc(s, link_capacity=1 << 16, **kw)
            assert got["n_replayed"] > 0 or k == 1
            _same(want, got)
            s.free()
        # -s bipartite and appended ranges through the same scratch
        codes = gen.c1s(420, seed=153)
        s = eng.upload(codes)
        for kw in (dict(kind=T.CROSS, n1=120), dict(kind=T.APPEND, n1=360)):
            print(f"noalloc {kw}", flush=True)
            kw.update(mode=T.RESOLVE, threshold=0.05, min_overlap=500, resolve_fraction=0.05)
            _same(eng.run(s, **kw), eng.run_noalloc(s, **kw))
        # refusals, each with the reason, and the scratch still serves a run afterwards
        with pytest.raises(RuntimeError, match="link_capacity"):
            eng.run_noalloc(s, link_capacity=10, mode=T.RESOLVE, threshold=1.0, min_overlap=500)
        with pytest.raises(RuntimeError, match="exact-path"):
            eng.run_noalloc(s, flags=T.FLAG_EXACT, mode=T.RESOLVE, threshold=0.05, min_overlap=500)
        kw = dict(mode=T.SKIP, threshold=0.05, min_overlap=500)
        _same(eng.run(s, **kw), eng.run_noalloc(s, **kw))
        s.free()
        big = eng.upload(gen.c1s(2000, seed=154))  # about 80 MB of run buffers: beyond the 16 MiB scratch
        with pytest.raises(RuntimeError, match="scratch"):
            eng.run_noalloc(big, mode=T.RESOLVE, threshold=0.015, min_overlap=500)
        big.free()
    finally:
        eng.close()
