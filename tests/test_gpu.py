"""GPU parity: the product (libtn93g.so through the tn93g CLI and the C ABI) against the golden fixtures, the
published tn93 pins and the oracle, byte for byte.  Needs an MI355X: run with ``pytest -m gpu``."""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, REPO

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(REPO))
import swifttn93_amd as T  # noqa: E402
from swifttn93_amd import gen  # noqa: E402

SEP = "\x1f"


def _cli():
    if not T.CLI_PATH.exists() or not T.LIB_PATH.exists():
        pytest.fail("the engine is not built (python3 -m swifttn93_amd.build)")
    return T.CLI_PATH


def run_batch(cases, tmp_path, timeout=600):
    """cases: list of (dir, argv list without program name).  Runs them in one tn93g process."""
    bf = tmp_path / "batch.txt"
    bf.write_text("".join(SEP.join([str(d)] + list(a)) + "\n" for d, a in cases))
    print(f"tn93g --batch: {len(cases)} cases", flush=True)
    p = subprocess.run([str(_cli()), "--batch", str(bf)], capture_output=True, timeout=timeout)
    assert p.returncode == 0, p.stderr[-2000:]
    res = []
    for d, _ in cases:
        rc = int((d / "rc.txt").read_text())
        out = (d / "out.txt").read_bytes() if (d / "out.txt").exists() else None
        res.append((rc, out, (d / "stdout.bin").read_bytes()))
    return res


def test_a_fuzz_fixtures(fuzz_cases, tmp_path):
    cases = []
    for k, c in enumerate(fuzz_cases):
        d = tmp_path / f"c{k}"
        d.mkdir()
        for key, v in c["files"].items():
            (d / f"{key}.fa").write_bytes(v.encode("latin-1"))
        argv = ["-q"] + [f"{f.strip('{}')}.fa" if f.startswith("{") else f for f in c["flags"]] + ["-o", "out.txt",
                                                                                                   "in.fa"]
        cases.append((d, argv))
    res = run_batch(cases, tmp_path)
    bad = []
    for k, (c, (rc, out, so)) in enumerate(zip(fuzz_cases, res)):
        if c["rc"] != 0:
            if rc == 0:
                bad.append((k, c["flags"], "expected failure"))
            continue
        exp = None if c["out"] is None else c["out"].encode("latin-1")
        if rc != 0 or out != exp or so != c["stdout"].encode("latin-1"):
            bad.append((k, c["flags"], rc, (out or b"")[:200], (exp or b"")[:200], so[:300]))
    assert not bad, bad[:3]


def _oracle_cli():
    cli = REPO / "oracle" / "_build" / "tn93_oracle"
    if not cli.exists():
        subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "oracle"], check=True)
    return cli


CONFIGS = [
    ("-a", "resolve", "-g", "0.05", "-l", "500"),
    ("-a", "average", "-l", "500"),
    ("-a", "skip", "-l", "500"),
    ("-a", "gapmm", "-l", "500"),
    ("-a", "RYN", "-g", "0.5", "-l", "500"),
    ("-a", "resolve", "-l", "1250", "-0"),
]


@pytest.mark.parametrize("workload", ["c1s", "c0"])
def test_synthetic_vs_oracle(workload, tmp_path):
    """Every option set at -t 0.015 and at -t 1 (every pair written) on a seeded synthetic set."""
    codes = gen.c1s(500) if workload == "c1s" else gen.c0()
    fa = tmp_path / "in.fa"
    gen.write_fasta(fa, codes)
    cases, want = [], []
    for ci, cfg in enumerate(CONFIGS):
        for t in ["0.015", "1"]:
            for fmt in (["csv", "csvn"] if t == "1" else ["csv"]):
                d = tmp_path / f"k{ci}_{t}_{fmt}"
                d.mkdir()
                argv = ["-q", "-t", t, "-f", fmt, *cfg, "-o", "out.txt", str(fa)]
                p = subprocess.run([str(_oracle_cli())] + argv[:-3] + ["-o", str(d / "want.txt"), str(fa)],
                                   capture_output=True, check=True)
                want.append(((d / "want.txt").read_bytes(), p.stdout))
                cases.append((d, argv))
    res = run_batch(cases, tmp_path)
    for (d, argv), (rc, out, so), (wo, ws) in zip(cases, res, want):
        assert rc == 0, argv
        assert out == wo, (argv, len(out or b""), len(wo))
        assert so == ws, (argv, so[:400], ws[:400])


def test_two_files_and_fst(tmp_path):
    codes = gen.c1s(300, seed=11)
    a, b = tmp_path / "a.fa", tmp_path / "b.fa"
    gen.write_fasta(a, codes[:120], [f"q{i}:{1 + i % 3}" for i in range(120)])
    gen.write_fasta(b, codes[120:], [f"r{i}" for i in range(180)])
    cases, want = [], []
    for k, extra in enumerate([[], ["-m"], ["-f", "csvn"], ["-f", "hyphy"]]):
        d = tmp_path / f"k{k}"
        d.mkdir()
        argv = ["-q", "-t", "0.05", "-s", str(b), *extra, "-o", "out.txt", str(a)]
        p = subprocess.run([str(_oracle_cli())] + argv[:-3] + ["-o", str(d / "want.txt"), str(a)],
                           capture_output=True, check=True)
        want.append(((d / "want.txt").read_bytes(), p.stdout))
        cases.append((d, argv))
    for (d, argv), (rc, out, so), (wo, ws) in zip(cases, run_batch(cases, tmp_path), want):
        assert rc == 0 and out == wo and so == ws, argv


def test_network_pins(hiv8766, tmp_path):
    """The ten published tn93 sha256 pins on the real 8,766-sequence HIV-1 pol set (38,416,995 pairs per run)."""
    P = json.loads((GOLDEN / "real" / "network_pins.json").read_text())
    cases = []
    for pin in P["pins"]:
        d = tmp_path / ("t" + pin["t_arg"])
        d.mkdir()
        flags = list(P["flags"])
        cases.append((d, flags[:1] + ["-o", "out.txt", "-t", pin["t_arg"]] + flags[1:] + [str(hiv8766)]))
    res = run_batch(cases, tmp_path)
    for pin, (rc, out, so) in zip(P["pins"], res):
        assert rc == 0
        assert hashlib.sha256(out).hexdigest() == pin["ref_file_sha256"], pin["t_arg"]
        assert hashlib.sha256(so).hexdigest() == pin["ref_stdout_sha256"], pin["t_arg"]


@pytest.fixture(scope="module")
def engine():
    e = T.Engine(0)
    yield e
    e.close()


def _same(r1, r2):
    assert (r1["hist"] == r2["hist"]).all()
    for k in ["link_row", "link_col"]:
        assert (r1[k] == r2[k]).all()
    assert r1["found_links"] == r2["found_links"] and r1["skipped"] == r2["skipped"]
    assert r1["weighted"] == r2["weighted"]
    # distances: the enclosures must overlap and print the same
    for a, b in zip(r1["link_dist"], r2["link_dist"]):
        assert "%g" % a == "%g" % b


@pytest.mark.parametrize("mode", [T.RESOLVE, T.AVERAGE, T.SKIP, T.SUBSET, T.GAPMM])
def test_fast_path_equals_replay(engine, mode):
    """The popcount + event-table path and the site-order replay kernel agree on every pair (all written)."""
    codes = gen.c1s(400, seed=5)
    s = engine.upload(codes)
    kw = dict(mode=mode, threshold=1.0, min_overlap=500, resolve_fraction=0.05 if mode == T.RESOLVE else 1.0,
              subset_mask=(1 << 5) | (1 << 6) | (1 << 15))
    fast = engine.run(s, **kw)
    slow = engine.run(s, flags=T.FLAG_ALL_SLOW, **kw)
    exact = engine.run(s, flags=T.FLAG_EXACT, **kw)
    _same(fast, slow)
    _same(fast, exact)
    assert exact["exact"]
    assert slow["n_replayed"] == 400 * 399 // 2
    for r in (fast, slow):
        assert (r["link_lo"] <= exact["link_dist"]).all() and (exact["link_dist"] <= r["link_hi"]).all()
        # the enclosures of the row-order mean and of the maximum contain the exact values
        assert r["mean_lo"][0] <= exact["mean_lo"][0] == exact["mean_hi"][0] <= r["mean_hi"][0]
        assert r["max_lo"][0] <= exact["max_lo"][0] <= r["max_hi"][0]
        assert (r["mean_hi"][0] - r["mean_lo"][0]) <= 1e-9 * abs(exact["mean_lo"][0])
    s.free()


@pytest.mark.parametrize("mode", [T.RESOLVE, T.AVERAGE, T.GAPMM, T.SUBSET])
def test_long_alignment_count_tables(engine, mode):
    """Long alignments (5,000 columns): the replay takes the clean/clean counts of order-dependent pairs from the
    special-sequence tables, rows against columns and (cross, append) columns against rows; every kind must match the
    per-pair counts of the all-replay run and the exact mode pair for pair."""
    codes = gen.c1(seed=9, n=260, L=5000, clusters=8, bdhv=0.002)  # 24 sequences with B/D/H/V codes
    s = engine.upload(codes)
    base = dict(mode=mode, threshold=1.0, min_overlap=500, resolve_fraction=0.05 if mode == T.RESOLVE else 1.0,
                subset_mask=(1 << 5) | (1 << 6) | (1 << 11) | (1 << 15))
    for kw in [dict(), dict(kind=T.CROSS, n1=100), dict(kind=T.APPEND, n1=200)]:
        kw = dict(base, **kw)
        fast = engine.run(s, **kw)
        slow = engine.run(s, flags=T.FLAG_ALL_SLOW, **kw)
        exact = engine.run(s, flags=T.FLAG_EXACT, **kw)
        assert fast["n_replayed"] > 0, kw
        _same(fast, slow)
        _same(fast, exact)
        assert (fast["link_lo"] <= exact["link_dist"]).all() and (exact["link_dist"] <= fast["link_hi"]).all()
    s.free()


def test_full_size_c1_vs_oracle(tmp_path):
    """The benchmark workload itself (6,700 x 1,300, 22.4 M pairs) against the oracle, byte for byte."""
    fa = tmp_path / "c1.fa"
    gen.write_fasta(fa, gen.c1())
    d = tmp_path / "run"
    d.mkdir()
    argv = ["-q", "-t", "0.015", "-a", "resolve", "-g", "0.05", "-l", "500", "-o", "out.txt", str(fa)]
    p = subprocess.run([str(_oracle_cli())] + argv[:-3] + ["-o", str(tmp_path / "want.txt"), str(fa)],
                       capture_output=True, check=True, timeout=600)
    (rc, out, so), = run_batch([(d, argv)], tmp_path)
    assert rc == 0
    assert out == (tmp_path / "want.txt").read_bytes()
    assert so == p.stdout


@pytest.mark.parametrize("kind", ["triangle", "cross"])
def test_backends_agree(engine, kind):
    """The int8 MFMA one-hot contraction and the VALU popcount give the same counts, hence the same outputs."""
    codes = gen.c1s(333, seed=9)
    s = engine.upload(codes)
    kw = dict(mode=T.RESOLVE, threshold=1.0, min_overlap=500, resolve_fraction=0.05)
    if kind == "cross":
        kw.update(kind=T.CROSS, n1=101)
    a = engine.run(s, backend=T.BACKEND_POPC, **kw)
    b = engine.run(s, backend=T.BACKEND_MFMA, **kw)
    _same(a, b)
    assert (a["link_dist"] == b["link_dist"]).all()
    s.free()


C2_CONFIGS = [
    ("-t", "0.0005", "-a", "resolve", "-g", "0.05", "-l", "25000"),
    ("-t", "1", "-a", "resolve", "-g", "0.05", "-l", "25000"),
    ("-t", "1", "-a", "average", "-l", "20000"),
    ("-t", "0.001", "-a", "skip", "-l", "25000", "-f", "csvn"),
]


@pytest.mark.parametrize("backend", ["popc", "mfma", "gemm"])
def test_long_alignment_c2s_vs_oracle(backend, tmp_path, monkeypatch):
    """SARS-CoV-2-shaped long alignment (300 x 29,903, N-masked blocks of 100-2,000 columns) against the oracle,
    on the popcount backend, on the int8 MFMA backend (clean/clean counts on the matrix unit, N from runs) and with
    the clean/clean counts as rocBLAS int8 GEMMs (TN93G_COUNTS=gemm)."""
    if backend == "gemm":
        monkeypatch.setenv("TN93G_COUNTS", "gemm")
    else:
        monkeypatch.setenv("TN93G_BACKEND", backend)
    codes = gen.c2s(300)
    fa = tmp_path / "in.fa"
    gen.write_fasta(fa, codes)
    cases, want = [], []
    for ci, cfg in enumerate(C2_CONFIGS):
        d = tmp_path / f"k{ci}"
        d.mkdir()
        argv = ["-q", *cfg, "-o", "out.txt", str(fa)]
        p = subprocess.run([str(_oracle_cli())] + argv[:-3] + ["-o", str(d / "want.txt"), str(fa)],
                           capture_output=True, check=True, timeout=600)
        want.append(((d / "want.txt").read_bytes(), p.stdout))
        cases.append((d, argv))
    for (d, argv), (rc, out, so), (wo, ws) in zip(cases, run_batch(cases, tmp_path), want):
        assert rc == 0, argv
        assert out == wo, (argv, len(out or b""), len(wo))
        assert so == ws, (argv, so[:400], ws[:400])


def _same_summary(r1, r2):
    _same(r1, r2)
    assert (r1["link_dist"] == r2["link_dist"]).all()
    assert r1["max_lo"][0] == r2["max_lo"][0] and r1["max_hi"][0] == r2["max_hi"][0]
    # the row sums are atomics in either order: the mean enclosures must overlap and print alike
    assert r1["mean_lo"][0] <= r2["mean_hi"][0] and r2["mean_lo"][0] <= r1["mean_hi"][0]
    assert "%.6g" % r1["mean_lo"][0] == "%.6g" % r2["mean_lo"][0]


@pytest.mark.parametrize("kind", ["triangle", "cross", "cross_cols"])
def test_gemm_tables_equal_kernel_tables(kind):
    """A single run's special-sequence tables as int8 GEMMs (rocBLAS, TN93G_TABLES=blas) against the same tables from
    k_spec16 (TN93G_TABLES=kernel), the GEMM run first in a fresh context (its tables are not left over from another
    run): triangle, -s with special sequences on both sides, and -s with special sequences among the columns only."""
    codes = gen.c2s(200, seed=5)
    kw = dict(mode=T.RESOLVE, threshold=0.05, min_overlap=5000, resolve_fraction=0.05)
    if kind == "cross":
        kw.update(kind=T.CROSS, n1=80)
    elif kind == "cross_cols":  # the rows are the last 40 sequences, none of them special
        spec = [k for k in range(codes.shape[0]) if any(c in (11, 12, 13, 14) for c in codes[k].tolist())]
        plain = [k for k in range(codes.shape[0]) if k not in set(spec)]
        order = plain[-40:] + [k for k in range(codes.shape[0]) if k not in set(plain[-40:])]
        codes = codes[order]
        kw.update(kind=T.CROSS, n1=40)
    eng = T.Engine(0)
    try:
        s = eng.upload(codes)
        res = {}
        for how in ["blas", "kernel"]:
            os.environ["TN93G_TABLES"] = how
            try:
                res[how] = eng.run(s, **kw)
            finally:
                del os.environ["TN93G_TABLES"]
        got, want = res["blas"], res["kernel"]
        s.free()
    finally:
        eng.close()
    print(kind, "replayed", want["n_replayed"], flush=True)
    assert want["n_replayed"] > 0
    for f in ["hist", "link_row", "link_col", "link_dist"]:
        assert np.array_equal(np.asarray(got[f]), np.asarray(want[f])), f
    for f in ["weighted", "max_lo", "max_hi", "found_links", "skipped", "n_replayed"]:
        assert got[f] == want[f], (f, got[f], want[f])
    assert abs(got["mean_lo"][0] - want["mean_lo"][0]) <= 1e-12 * abs(want["mean_lo"][0])


@pytest.mark.parametrize("kind", ["triangle", "cross", "cross_wide"])
@pytest.mark.parametrize("mode", ["resolve", "average"])
def test_gemm_counts_equal_popcount(kind, mode):
    """A single run's clean/clean counts as rocBLAS int8 GEMMs (TN93G_COUNTS=gemm) against the popcount pair kernel
    (TN93G_COUNTS=popc) on a dense-N long alignment, the GEMM run first in a fresh context: triangle (its products in
    three blocks of rows), -s with fewer rows than columns, and -s with more rows than columns (the products' operands
    span both ranges)."""
    codes = gen.c2s(180, seed=7)
    kw = dict(mode=T.RESOLVE if mode == "resolve" else T.AVERAGE, threshold=0.05 if mode == "resolve" else 1.0,
              min_overlap=5000, resolve_fraction=0.05 if mode == "resolve" else 1.0)
    if kind == "cross":
        kw.update(kind=T.CROSS, n1=50)
    elif kind == "cross_wide":
        kw.update(kind=T.CROSS, n1=130)
    eng = T.Engine(0)
    try:
        s = eng.upload(codes)
        res = {}
        for how in ["gemm", "popc"]:
            os.environ["TN93G_COUNTS"] = how
            os.environ["TN93G_GEMM_BLOCKS"] = "3"  # a triangle's products in three blocks of rows
            try:
                res[how] = eng.run(s, **kw)
            finally:
                del os.environ["TN93G_COUNTS"], os.environ["TN93G_GEMM_BLOCKS"]
        s.free()
    finally:
        eng.close()
    got, want = res["gemm"], res["popc"]
    assert got["backend_used"] == T.BACKEND_GEMM and want["backend_used"] == T.BACKEND_POPC
    _same_summary(got, want)
    for f in ["hist", "link_row", "link_col", "link_dist"]:
        assert np.array_equal(np.asarray(got[f]), np.asarray(want[f])), f
    print(kind, mode, "links", len(want["link_dist"]), flush=True)
    assert len(want["link_dist"]) > 0


def test_link_blocks_beyond_the_pinned_pool():
    """Single fast-path runs write their links straight into pinned link blocks of the library (at most 8): twelve
    results kept alive at once (the last ones on host blocks and the host's copy), then freed and the run repeated on
    the kept blocks, and a run with more links than the previous one read back (its block too small): every result's
    links and summary equal the first's, in a fresh context."""
    codes = gen.c0(seed=31, N=700, L=900)
    kw = dict(mode=T.RESOLVE, threshold=0.05, min_overlap=300, resolve_fraction=0.05)
    eng = T.Engine(0)
    try:
        s = eng.upload(codes)
        first = eng.run(s, **kw)
        assert len(first["link_dist"]) > 4096  # more than a fresh context reads back with its counters
        alive = [eng.run(s, **kw) for _ in range(12)]
        for r in alive:
            _same_summary(r, first)
            assert np.array_equal(r["link_dist"], first["link_dist"])
        del alive
        again = [eng.run(s, **kw) for _ in range(3)]
        for r in again:
            _same_summary(r, first)
        wide = eng.run(s, **dict(kw, threshold=0.2))  # more links than the blocks were sized for
        want = eng.run(s, **dict(kw, threshold=0.2, flags=T.FLAG_EXACT))
        assert len(wide["link_dist"]) > len(first["link_dist"])
        _same(wide, want)
        s.free()
    finally:
        eng.close()


def test_batch_equals_single(engine):
    """tn93g_run_batch (one graph launch for all jobs) gives every job its single-run result.  The cached graph is
    launched again with eager runs in between, and every launch is checked."""
    specs = [(gen.c0(seed=21, N=120, L=700), T.RESOLVE), (gen.c0(seed=22, N=64, L=2100), T.AVERAGE),
             (gen.c1s(200, seed=23), T.SKIP), (gen.c0(seed=24, N=90, L=333), T.GAPMM),
             (gen.c1s(150, seed=25), T.SUBSET)]
    sets, jobs = [], []
    for codes, mode in specs:
        st = engine.upload(codes)
        sets.append(st)
        jobs.append((st, dict(mode=mode, threshold=0.05, min_overlap=min(300, codes.shape[1] // 2),
                              resolve_fraction=0.05 if mode in (T.RESOLVE, T.SUBSET) else 1.0,
                              subset_mask=(1 << 5) | (1 << 6) | (1 << 15))))
    jobs.append((sets[2], dict(mode=T.RESOLVE, kind=T.CROSS, n1=77, threshold=0.05, min_overlap=300,
                               resolve_fraction=0.05)))
    jobs.append((sets[0], dict(mode=T.RESOLVE, threshold=0.05, min_overlap=300, flags=T.FLAG_EXACT)))
    jobs.append((sets[4], dict(mode=T.AVERAGE, threshold=1.0, min_overlap=300, backend=T.BACKEND_AUTO,
                               flags=T.FLAG_AUTOTUNE)))
    want = [engine.run(st, **kw) for st, kw in jobs]
    for rep in range(3):
        got = engine.run_batch(jobs)
        assert len(got) == len(want)
        for k, (w, g) in enumerate(zip(want, got)):
            print(f"batch launch {rep}, job {k}", flush=True)
            _same_summary(w, g)
        engine.run(sets[1], **jobs[1][1])  # eager work between the launches of the cached graph
    # the same launches issued directly on the streams (TN93G_NO_GRAPH: the profilers' counter passes)
    os.environ["TN93G_NO_GRAPH"] = "1"
    try:
        got = engine.run_batch(jobs)
    finally:
        del os.environ["TN93G_NO_GRAPH"]
    for k, (w, g) in enumerate(zip(want, got)):
        print(f"batch without a graph, job {k}", flush=True)
        _same_summary(w, g)
    for st in sets:
        st.free()


def test_append_runs_match_triangles(engine):
    """A set grown by appends: the APPEND run of a batch gives exactly the pairs a triangle over everything up to the
    batch adds to the triangle before it (histogram difference, links with a column in the batch, bit for bit)."""
    codes = gen.c1s(420, seed=31)
    st = engine.create(1300, 500)
    assert st.append(codes[:300]) == 0
    assert st.append(codes[300:360]) == 300
    assert st.append(codes[360:420]) == 360 and st.size() == 420
    kw = dict(mode=T.RESOLVE, threshold=0.05, min_overlap=500, resolve_fraction=0.05)
    tri = {m: engine.upload(codes[:m]) for m in (300, 360, 420)}
    full = {m: engine.run(s, **kw) for m, s in tri.items()}
    for n1, n2 in ((300, 360), (360, 420), (300, 420)):
        a = engine.run(st, kind=T.APPEND, n1=n1, n2=n2, **kw)
        hi, lo = full[n2], full[n1]
        assert (a["hist"] == hi["hist"] - lo["hist"]).all()
        m = hi["link_col"] >= n1
        for k in ("link_row", "link_col", "link_dist", "link_lo", "link_hi"):
            assert (a[k] == hi[k][m]).all(), (n1, n2, k)
        assert a["found_links"] == hi["found_links"] - lo["found_links"]
        assert a["skipped"] == hi["skipped"] - lo["skipped"]
    for s in tri.values():
        s.free()
    st.free()


def test_serving_alone_vs_mixed():
    """Continuous batching: query batches arriving together are admitted as workers free up and run concurrently;
    every batch's result is the same as when it is computed alone afterwards."""
    from swifttn93_amd import serve
    panel = gen.c1s(1500, seed=41)
    stream = serve.query_stream(panel, seed=42, rate=200.0, seconds=0.06, max_batch=200)
    nq = sum(c.shape[0] for _, c in stream)
    srv = serve.Server(panel, capacity=1500 + nq, workers=3, hbm_budget=1 << 30, run_kw=dict(threshold=0.03))
    try:
        reqs = serve.serve(srv, stream)
        assert len(reqs) >= 4
        assert all(r.error is None for r in reqs), [r.error for r in reqs if r.error]
        overlap = any(a.started < b.done and b.started < a.done for a in reqs for b in reqs if a is not b)
        assert overlap, "no two batches were on the device together"
        for r in reqs:
            print(f"request {r.rid}: rows [{r.n1}, {r.n2})", flush=True)
            alone = srv.run_alone(r)
            for k in ("link_row", "link_col"):
                assert (alone[k] == r.result[k]).all(), (r.rid, k)
            assert ["%g" % x for x in alone["link_dist"]] == ["%g" % x for x in r.result["link_dist"]], r.rid
    finally:
        srv.close()


@pytest.mark.parametrize("mode", [T.RESOLVE, T.AVERAGE, T.SKIP])
def test_dense_n_fast_equals_replay(engine, mode):
    """A set with long N blocks counts N from bit planes in the pair kernel and replays N runs as segments; both
    agree with the site-order replay of every pair and with the host-exact path."""
    codes = gen.c2s(160, seed=51)
    s = engine.upload(codes)
    kw = dict(mode=mode, threshold=1.0, min_overlap=20000, resolve_fraction=0.05 if mode == T.RESOLVE else 1.0)
    fast = engine.run(s, **kw)
    slow = engine.run(s, flags=T.FLAG_ALL_SLOW, **kw)
    exact = engine.run(s, flags=T.FLAG_EXACT, **kw)
    mf = engine.run(s, backend=T.BACKEND_MFMA, **kw)
    mfx = engine.run(s, backend=T.BACKEND_MFMA, flags=T.FLAG_EXACT, **kw)
    _same(fast, slow)
    _same(fast, exact)
    _same(fast, mf)
    _same(exact, mfx)
    assert mf["backend_used"] == T.BACKEND_MFMA
    assert slow["n_replayed"] == 160 * 159 // 2
    assert (fast["link_dist"] == slow["link_dist"]).all()
    assert (fast["link_dist"] == mf["link_dist"]).all()
    s.free()


def _edge_codes(seed, n, L, dense_n):
    """Random bases with 3 % ambiguity codes (B/D/H/V and U included), leading, trailing and inner gap runs (some
    reaching the first or the last site), and with dense_n an N block per sequence."""
    rs = gen.Stream(seed)
    codes = rs.integers(4, n * L).reshape(n, L).astype(np.uint8)
    amb = rs.uniform(n * L).reshape(n, L) < 0.03
    pool = np.array([4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16], dtype=np.uint8)
    pick = pool[rs.integers(len(pool), n * L).reshape(n, L)]
    codes[amb] = pick[amb]
    for i in range(n):
        u = rs.uniform(10)
        if u[0] < 0.4:
            codes[i, :1 + int(u[1] * min(L, 70))] = gen.GAP
        if u[2] < 0.4:
            codes[i, L - 1 - int(u[3] * min(L - 1, 70)):] = gen.GAP
        if u[4] < 0.6:
            a = int(u[5] * L)
            codes[i, a:min(L, a + 1 + int(u[6] * 40))] = gen.GAP
        if dense_n and u[7] < 0.8:
            a = int(u[8] * L)
            codes[i, a:min(L, a + 40 + int(u[9] * 60))] = 15
    return codes


@pytest.mark.parametrize("L,dense", [(31, False), (32, False), (33, False), (64, False), (96, False), (300, False),
                                     (160, True), (256, True)])
def test_marginal_counts_edges(engine, L, dense):
    """The popcount kernel's base sums from per-sequence prefix totals (corrected at events, N runs and gap runs)
    against the site-order replay of every pair and against the MFMA backend's direct 4x4 counts, at lengths on and
    around word boundaries with gap runs touching both ends."""
    codes = _edge_codes(61 + L, 150, L, dense)
    s = engine.upload(codes)
    for mode in (T.RESOLVE, T.AVERAGE, T.GAPMM, T.SUBSET):
        print(f"L={L} dense={dense} mode={mode}", flush=True)
        kw = dict(mode=mode, threshold=1.0, min_overlap=5, resolve_fraction=0.05 if mode == T.RESOLVE else 1.0,
                  subset_mask=(1 << 5) | (1 << 6) | (1 << 15))
        fast = engine.run(s, backend=T.BACKEND_POPC, **kw)
        slow = engine.run(s, flags=T.FLAG_ALL_SLOW, **kw)
        mf = engine.run(s, backend=T.BACKEND_MFMA, **kw)
        _same(fast, slow)
        _same(fast, mf)
        assert (fast["link_dist"] == mf["link_dist"]).all()
    s.free()


def test_event_chunks(engine):
    """Tiles whose events overflow the LDS (15 % ambiguity codes over 2,000 columns: about 300 events per sequence,
    several chunks per tile) and, under -g 0.05, most pairs computed again in AVERAGE over all chunks: the pair
    kernel against the site-order replay of every pair and the MFMA backend."""
    rs = gen.Stream(71)
    n, L = 120, 2000
    codes = rs.integers(4, n * L).reshape(n, L).astype(np.uint8)
    amb = rs.uniform(n * L).reshape(n, L) < 0.15
    pool = np.array([5, 6, 7, 8, 9, 10, 15], dtype=np.uint8)
    codes[amb] = pool[rs.integers(len(pool), n * L).reshape(n, L)][amb]
    codes[::7, :150] = gen.GAP
    s = engine.upload(codes)
    for mode, rf in ((T.RESOLVE, 0.05), (T.RESOLVE, 1.0), (T.AVERAGE, 1.0), (T.GAPMM, 1.0)):
        print(f"mode={mode} rf={rf}", flush=True)
        kw = dict(mode=mode, threshold=1.0, min_overlap=100, resolve_fraction=rf)
        fast = engine.run(s, backend=T.BACKEND_POPC, **kw)
        slow = engine.run(s, flags=T.FLAG_ALL_SLOW, **kw)
        mf = engine.run(s, backend=T.BACKEND_MFMA, **kw)
        _same(fast, slow)
        _same(fast, mf)
    s.free()


@pytest.mark.parametrize("case", ["c1s", "dense_n", "chunks", "edges", "cross_append"])
def test_tile_shapes_agree(engine, case):
    """The pair kernel in tiles of 8, 16 and 32 rows (the autotuner's shapes) gives the same outputs bit for bit, on
    ambiguity-heavy sets (B/D/H/V: replays), dense N, event chunks that overflow the LDS, lengths around word
    boundaries with gap runs at both ends, and -s / appended ranges; the 16-row shape also against exact mode."""
    if case == "c1s":
        codes, L_ov = gen.c1s(300, seed=171), 500
    elif case == "dense_n":
        codes, L_ov = gen.c2s(120, seed=172), 20000
    elif case == "chunks":
        rs = gen.Stream(173)
        n, L = 100, 2000
        codes = rs.integers(4, n * L).reshape(n, L).astype(np.uint8)
        amb = rs.uniform(n * L).reshape(n, L) < 0.15
        pool = np.array([5, 6, 7, 8, 9, 10, 11, 15], dtype=np.uint8)
        codes[amb] = pool[rs.integers(len(pool), n * L).reshape(n, L)][amb]
        L_ov = 100
    else:
        codes, L_ov = _edge_codes(174, 150, 97, False), 5
    s = engine.upload(codes)
    kinds = [dict()] if case != "cross_append" else [dict(kind=T.CROSS, n1=37), dict(kind=T.APPEND, n1=101)]
    for kind in kinds:
        for mode in (T.RESOLVE, T.AVERAGE, T.GAPMM, T.SUBSET):
            kw = dict(mode=mode, threshold=1.0, min_overlap=L_ov, resolve_fraction=0.05 if mode == T.RESOLVE else 1.0,
                      subset_mask=(1 << 5) | (1 << 6) | (1 << 15), backend=T.BACKEND_POPC, **kind)
            print(f"{case} {kind} mode={mode}", flush=True)
            base = engine.run(s, tile_rows=16, **kw)
            for tr in (8, 32):
                other = engine.run(s, tile_rows=tr, **kw)
                _same(base, other)
                assert (base["link_dist"] == other["link_dist"]).all()
                assert base["n_replayed"] == other["n_replayed"]
            exact = engine.run(s, tile_rows=16, flags=T.FLAG_EXACT, **kw)
            _same(base, exact)
    s.free()


def test_autotuned_shapes(engine):
    """With TN93G_FLAG_AUTOTUNE and tile_rows 0 the autotuner times the backends and tile heights on the first job
    of a shape; the tuned job gives the untuned job's outputs."""
    codes = gen.c1s(400, seed=175)
    s = engine.upload(codes)
    kw = dict(mode=T.RESOLVE, threshold=0.05, min_overlap=500, resolve_fraction=0.05)
    want = engine.run(s, tile_rows=16, backend=T.BACKEND_POPC, **kw)
    for _ in range(2):  # the first call tunes, the second takes the tuned choice
        got = engine.run(s, backend=T.BACKEND_AUTO, flags=T.FLAG_AUTOTUNE, **kw)
        _same(want, got)
    s.free()


def test_exact_list_grows(monkeypatch):
    """The host-exact list starts small (TN93G_EXACT_LIST) and a run with more uncertain pairs than it holds is
    repeated with the list it needed: the result equals the fast path's."""
    monkeypatch.setenv("TN93G_EXACT_LIST", "64")
    e = T.Engine(0)  # a context of its own: no list size remembered from earlier runs
    try:
        codes = gen.c1s(300, seed=201)
        s = e.upload(codes)
        kw = dict(mode=T.RESOLVE, threshold=0.05, min_overlap=500, resolve_fraction=0.05)
        exact = e.run(s, flags=T.FLAG_EXACT, **kw)
        assert exact["exact"] and exact["n_host_exact"] > 64
        _same(e.run(s, **kw), exact)
        s.free()
    finally:
        e.close()
