"""GPU parity of constant-column folding: columns where every sequence holds the same clean base are left out of the
device arrays by tn93g_set_upload and their sites added to every pair's counts (swifttn93.cpp:591-660 folds them the
same way).

* the real SARS-CoV-2 sets of the reference (two thirds of their 29,903 columns are constant): the reference build's
  sha256 of file and JSON under eight option sets, folded and with TN93G_NOFOLD=1;
* seeded sets with most columns forced constant, every option set, against the oracle (which never folds), and the
  C ABI's full results folded and unfolded, for both backends, every tile height and the no-allocation run.

Needs an MI355X: run with ``pytest -m gpu``."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO, sars2_argv

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(REPO))
import swifttn93_amd as T  # noqa: E402
from swifttn93_amd import gen  # noqa: E402

SEP = "\x1f"


def _cli():
    if not T.CLI_PATH.exists() or not T.LIB_PATH.exists():
        pytest.fail("the engine is not built (python3 -m swifttn93_amd.build)")
    return T.CLI_PATH


def _oracle_cli():
    cli = REPO / "oracle" / "_build" / "tn93_oracle"
    if not cli.exists():
        subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "oracle"], check=True)
    return cli


def run_batch(cases, tmp_path, env=None, timeout=600):
    """cases: list of (dir, argv without program name), run by one `tn93g --batch` process."""
    bf = tmp_path / "batch.txt"
    bf.write_text("".join(SEP.join([str(d)] + list(a)) + "\n" for d, a in cases))
    print(f"tn93g --batch: {len(cases)} cases, env {env}", flush=True)
    p = subprocess.run([str(_cli()), "--batch", str(bf)], capture_output=True, timeout=timeout,
                       env=dict(os.environ, **(env or {})))
    assert p.returncode == 0, p.stderr[-2000:]
    res = []
    for d, _ in cases:
        rc = int((d / "rc.txt").read_text())
        out = (d / "out.txt").read_bytes() if (d / "out.txt").exists() else None
        res.append((rc, out, (d / "stdout.bin").read_bytes()))
    return res


def with_constant_columns(codes, frac, seed):
    """A copy of codes where a fraction `frac` of the columns (chosen by seed) hold one clean base in every sequence:
    sequence 0's base there, or A where sequence 0 has none."""
    rng = np.random.default_rng(seed)
    out = codes.copy()
    L = codes.shape[1]
    cols = np.sort(rng.choice(L, int(frac * L), replace=False))
    base = np.where(codes[0, cols] < 4, codes[0, cols], 0).astype(np.uint8)
    out[:, cols] = base[None, :]
    return out


@pytest.mark.parametrize("nofold", ["0", "1"])
def test_sars2_pins(sars2, tmp_path, nofold):
    """The real SARS-CoV-2 set against the reference build's hashes: folded (20,673 of 29,903 columns) and not."""
    paths, pins = sars2
    cases = []
    for k, pin in enumerate(pins):
        d = tmp_path / f"p{k}"
        d.mkdir()
        cases.append((d, ["-q"] + sars2_argv(pin["flags"], paths) + ["-o", "out.txt", str(paths["in"])]))
    res = run_batch(cases, tmp_path, env={"TN93G_NOFOLD": nofold})
    bad = []
    for pin, (rc, out, so) in zip(pins, res):
        if rc != 0 or hashlib.sha256(out or b"").hexdigest() != pin["ref_file_sha256"] or \
                hashlib.sha256(so).hexdigest() != pin["ref_stdout_sha256"]:
            bad.append((pin["flags"], rc, len(out or b""), pin["file_bytes"], so[:300]))
    assert not bad, bad


def test_sars2_folded_count(sars2):
    """The upload folds exactly the constant columns of the real set (20,673), and the results agree unfolded."""
    paths, _ = sars2
    names, codes = [], []
    lut = np.full(256, 255, dtype=np.uint8)
    for k, ch in enumerate("ACGTURYSWKMBDHVN?-"):
        lut[ord(ch)] = lut[ord(ch.lower())] = k
    for block in paths["in"].read_text().split(">")[1:]:
        lines = block.splitlines()
        names.append(lines[0])
        codes.append(lut[np.frombuffer("".join(lines[1:]).encode(), dtype=np.uint8)])
    codes = np.stack(codes)
    want = int(((codes == codes[0]).all(0) & (codes[0] < 4)).sum())
    assert want == 20673  # the SARS-CoV-2 dev set's constant columns (29,903 - 9,230 that vary)
    eng = T.Engine(0)
    try:
        s = eng.upload(codes)
        r = eng.run(s, mode=T.RESOLVE, resolve_fraction=0.05, min_overlap=20000, threshold=0.0005, certify=True)
        s.free()
        os.environ["TN93G_NOFOLD"] = "1"
        try:
            s = eng.upload(codes)
        finally:
            del os.environ["TN93G_NOFOLD"]
        u = eng.run(s, mode=T.RESOLVE, resolve_fraction=0.05, min_overlap=20000, threshold=0.0005, certify=True)
        s.free()
    finally:
        eng.close()
    assert r["folded_columns"] == 20673 and u["folded_columns"] == 0
    _same(r, u)


CONFIGS = [
    ("-a", "resolve", "-g", "0.05", "-l", "500"),
    ("-a", "average", "-l", "500"),
    ("-a", "skip", "-l", "500"),
    ("-a", "gapmm", "-l", "500"),
    ("-a", "RYN", "-g", "0.5", "-l", "500"),
    ("-a", "resolve", "-l", "1250", "-0"),
]


@pytest.mark.parametrize("workload", ["c1s", "c2s"])
def test_constant_columns_vs_oracle(workload, tmp_path):
    """Most columns constant (60 % of a 1,300-column set with B/D/H/V codes, 50 % of a 29,903-column set with dense
    N blocks): every option set at -t 0.015 and at -t 1, and -s / -s -m, byte for byte against the oracle."""
    if workload == "c1s":
        codes = with_constant_columns(gen.c1s(400, seed=21), 0.6, 1)
        cfgs = CONFIGS
    else:
        codes = with_constant_columns(gen.c2s(160, seed=22), 0.5, 2)
        cfgs = [("-a", "resolve", "-g", "0.05", "-l", "5000"), ("-a", "average", "-l", "5000"),
                ("-a", "gapmm", "-l", "5000")]
    fa, fb = tmp_path / "in.fa", tmp_path / "second.fa"
    n1 = codes.shape[0] * 2 // 5
    gen.write_fasta(fa, codes)
    gen.write_fasta(fb, codes[n1:], [f"r{i}" for i in range(codes.shape[0] - n1)])
    fq = tmp_path / "first.fa"
    gen.write_fasta(fq, codes[:n1], [f"q{i}" for i in range(n1)])
    runs = []
    for ci, cfg in enumerate(cfgs):
        for t in ["0.015", "1"]:
            runs.append((f"k{ci}_{t}", ["-t", t, *cfg], fa))
    runs.append(("s", ["-t", "0.05", "-s", str(fb), *cfgs[0]], fq))
    runs.append(("sm", ["-t", "0.05", "-s", str(fb), "-m", *cfgs[1]], fq))
    cases, want = [], []
    for name, argv, src in runs:
        d = tmp_path / name
        d.mkdir()
        p = subprocess.run([str(_oracle_cli()), "-q", *argv, "-o", str(d / "want.txt"), str(src)],
                           capture_output=True, check=True)
        want.append(((d / "want.txt").read_bytes(), p.stdout))
        cases.append((d, ["-q", *argv, "-o", "out.txt", str(src)]))
    for env in [{"TN93G_NOFOLD": "0"}, {"TN93G_NOFOLD": "1"}]:
        res = run_batch(cases, tmp_path, env=env)
        for (d, argv), (rc, out, so), (wo, ws) in zip(cases, res, want):
            assert rc == 0, (env, argv)
            assert out == wo, (env, argv, len(out or b""), len(wo))
            assert so == ws, (env, argv, so[:400], ws[:400])


def _same(a, b):
    for k in ["hist", "link_row", "link_col", "link_dist", "link_lo", "link_hi"]:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    for k in ["weighted", "max_lo", "max_hi", "found_links", "skipped", "exact"]:
        assert a[k] == b[k], (k, a[k], b[k])
    # the fast path adds its row sums by atomics, in no fixed order: the mean's enclosure may differ in its last bits
    # from run to run (exact results are tn93's own sums and must be equal)
    lo, hi, lo1, hi1 = a["mean_lo"][0], a["mean_hi"][0], b["mean_lo"][0], b["mean_hi"][0]
    assert lo <= hi1 and lo1 <= hi and abs(lo - lo1) <= 1e-12 * abs(lo1) and abs(hi - hi1) <= 1e-12 * abs(hi1)
    if a["exact"]:
        assert a["mean_lo"] == b["mean_lo"] and a["mean_hi"] == b["mean_hi"]


def test_folded_backends_and_tiles():
    """The C ABI's whole result on a folded set equals the unfolded one for each backend and tile height (popcount), under
    RESOLVE with the -g fallback (replayed B/D/H/V pairs included), and the no-allocation run agrees."""
    codes = with_constant_columns(gen.c1s(300, seed=23), 0.7, 3)
    eng = T.Engine(0)
    try:
        folded = eng.upload(codes)
        os.environ["TN93G_NOFOLD"] = "1"
        try:
            plain = eng.upload(codes)
        finally:
            del os.environ["TN93G_NOFOLD"]
        kw = dict(mode=T.RESOLVE, resolve_fraction=0.05, min_overlap=300, threshold=0.03)
        base = eng.run(plain, backend=T.BACKEND_POPC, tile_rows=16, **kw)
        assert base["folded_columns"] == 0 and base["n_replayed"] > 0
        for backend in [T.BACKEND_POPC, T.BACKEND_MFMA]:
            for rows in ([8, 16, 32] if backend == T.BACKEND_POPC else [16]):  # the MFMA tiles are 16 rows
                r = eng.run(folded, backend=backend, tile_rows=rows, **kw)
                assert r["folded_columns"] >= int(0.7 * codes.shape[1]), r["folded_columns"]
                assert r["backend_used"] == backend and r["tile_rows_used"] == rows
                _same(r, base)
        ex = eng.run(folded, flags=T.FLAG_EXACT, **kw)
        ex0 = eng.run(plain, flags=T.FLAG_EXACT, **kw)
        _same(ex, ex0)
        na = eng.run_noalloc(folded, link_capacity=1 << 16, **kw)
        assert na["folded_columns"] > 0
        for k in ["hist", "link_row", "link_col", "link_dist"]:
            assert np.array_equal(np.asarray(na[k]), np.asarray(base[k])), k
        folded.free()
        plain.free()
    finally:
        eng.close()
