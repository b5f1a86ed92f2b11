"""GPU parity of the paths around a single run, against the oracle and the published fixtures:

* the batch runner: many alignments in one tn93g_run_batch (tn93g --jobs), each case's output byte for byte;
* the process interface as callers use it: tn93g exec'd once per call, with and without -o (links on stdout, the
  JSON on stderr, swifttn93.cpp:2904-2912), quiet or not, from a file or stdin, and error exits;
* alignments of 65,535 columns or more (the reference's slow path, swifttn93.cpp:580-590);
* serving: every served batch's links against the oracle's triangle over everything received, merged or not, and
  alone = mixed on the full 50,000-sequence panel.

Needs an MI355X: run with ``pytest -m gpu``."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO

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


def run_jobs(cases, tmp_path, timeout=600):
    """cases: list of (dir, argv without program name); all run by one `tn93g --jobs` process as batches."""
    bf = tmp_path / "jobs.txt"
    bf.write_text("".join(SEP.join([str(d)] + list(a)) + "\n" for d, a in cases))
    print(f"tn93g --jobs: {len(cases)} cases", flush=True)
    p = subprocess.run([str(_cli()), "--jobs", str(bf)], capture_output=True, timeout=timeout)
    assert p.returncode == 0, p.stderr[-2000:]
    res = []
    for d, _ in cases:
        rc = int((d / "rc.txt").read_text())
        out = (d / "out.txt").read_bytes() if (d / "out.txt").exists() else None
        res.append((rc, out, (d / "stdout.bin").read_bytes()))
    return res


def test_jobs_fuzz_fixtures(fuzz_cases, tmp_path):
    """The 600 reference-built fuzz fixtures as batches of up to 256 jobs in one tn93g_run_batch each."""
    cases = []
    for k, c in enumerate(fuzz_cases):
        d = tmp_path / f"c{k}"
        d.mkdir()
        for key, v in c["files"].items():
            (d / f"{key}.fa").write_bytes(v.encode("latin-1"))
        argv = ["-q"] + [f"{f.strip('{}')}.fa" if f.startswith("{") else f for f in c["flags"]] + ["-o", "out.txt",
                                                                                                   "in.fa"]
        cases.append((d, argv))
    res = run_jobs(cases, tmp_path)
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


def _mode_flags(mode):
    return ["-a", mode] + (["-g", "0.05"] if mode == "resolve" else [])


def test_jobs_c3_vs_oracle(tmp_path):
    """The batched workload's generator (c0 alignments, log-uniform sizes, uniform lengths, -a modes cycling
    resolve/average/skip/gapmm) at sizes the oracle finishes in seconds: 40 alignments of 20-400 sequences x
    300-4,000 columns, plus -s bipartite jobs, an F_ST job and all-pairs-written jobs, all in one batch."""
    jobs = gen.c3(seed=106, count=40, n_range=(20, 400), L_range=(300, 4000))
    cases, want = [], []
    orc = _oracle_cli()

    def add(name, argv, files):
        d = tmp_path / name
        d.mkdir()
        for fn, codes in files.items():
            gen.write_fasta(d / fn, codes, gen.names_for(codes.shape[0], fn[0]))
        p = subprocess.run([str(orc), "-q", *argv, "-o", "want.txt", "in.fa"], cwd=d, capture_output=True,
                           check=True, timeout=600)
        want.append(((d / "want.txt").read_bytes(), p.stdout))
        cases.append((d, ["-q", *argv, "-o", "out.txt", "in.fa"]))

    for k, (codes, mode) in enumerate(jobs):
        L = codes.shape[1]
        t = ["0.015", "0.05", "1"][k % 3]
        add(f"j{k}", ["-t", t, "-l", str(min(500, L // 2)), *_mode_flags(mode)], {"in.fa": codes})
    for k in range(4):  # -s query-by-reference blocks (CROSS), and one -m (F_ST, exact path, run alone)
        codes = gen.c1s(260 + 40 * k, seed=120 + k)
        nq = 40 + 20 * k
        extra = ["-m"] if k == 3 else []
        add(f"s{k}", ["-t", "0.02", "-l", "500", "-s", "ref.fa", *extra, *_mode_flags(gen.C3_MODES[k])],
            {"in.fa": codes[:nq], "ref.fa": codes[nq:]})
    res = run_jobs(cases, tmp_path)
    bad = [(str(d), rc) for (d, _), (rc, out, so), (wo, ws) in zip(cases, res, want)
           if rc != 0 or out != wo or so != ws]
    assert not bad, bad[:5]


@pytest.mark.parametrize("rows", ["8", "32"])
def test_tile_rows_vs_oracle(rows, tmp_path):
    """The CLI with the pair kernel forced to tiles of 8 or 32 rows (TN93G_TILE_ROWS) against the oracle, every -a
    mode, -t 0.015 and every pair written, a single file and -s."""
    codes = gen.c1s(420, seed=181)
    gen.write_fasta(tmp_path / "in.fa", codes)
    gen.write_fasta(tmp_path / "ref.fa", codes[:150], gen.names_for(150, "r"))
    cfgs = [("-a", "resolve", "-g", "0.05"), ("-a", "average"), ("-a", "skip"), ("-a", "gapmm"), ("-a", "RYN")]
    cases, want = [], []
    for ci, cfg in enumerate(cfgs):
        for t in ("0.015", "1"):
            for extra in ([], ["-s", str(tmp_path / "ref.fa")]):
                d = tmp_path / f"k{ci}_{t}_{len(extra)}"
                d.mkdir()
                argv = ["-q", "-t", t, "-l", "500", *cfg, *extra]
                p = subprocess.run([str(_oracle_cli()), *argv, "-o", str(d / "want.txt"), str(tmp_path / "in.fa")],
                                   capture_output=True, check=True, timeout=600)
                want.append(((d / "want.txt").read_bytes(), p.stdout))
                cases.append((d, [*argv, "-o", "out.txt", str(tmp_path / "in.fa")]))
    bf = tmp_path / "jobs.txt"
    bf.write_text("".join(SEP.join([str(d)] + list(a)) + "\n" for d, a in cases))
    p = subprocess.run([str(_cli()), "--batch", str(bf)], capture_output=True, timeout=600,
                       env=dict(os.environ, TN93G_TILE_ROWS=rows))
    assert p.returncode == 0, p.stderr[-2000:]
    for (d, argv), (wo, ws) in zip(cases, want):
        assert (d / "rc.txt").read_text().strip() == "0", argv
        assert (d / "out.txt").read_bytes() == wo, argv
        assert (d / "stdout.bin").read_bytes() == ws, argv


def _exec(argv, cwd, stdin=None, exe=None):
    p = subprocess.run([str(exe or _cli()), *argv], cwd=cwd, stdin=stdin, capture_output=True, timeout=300)
    return p.returncode, p.stdout, p.stderr


PROCESS_CASES = [
    ("o_quiet", ["-q", "-t", "0.02", "-l", "500", "-a", "resolve", "-g", "0.05", "-o", "out.csv", "in.fa"]),
    ("o_verbose", ["-t", "0.02", "-l", "500", "-a", "average", "-o", "out.csv", "in.fa"]),
    ("stdout_quiet", ["-q", "-t", "0.02", "-l", "500", "-a", "resolve", "-g", "0.05", "in.fa"]),
    ("stdout_verbose", ["-t", "0.03", "-l", "500", "-a", "skip", "-f", "csvn", "in.fa"]),
    ("stdout_dash", ["-t", "0.02", "-l", "500", "-o", "-", "-0", "in.fa"]),
    ("stdin", ["-q", "-t", "0.02", "-l", "500", "-a", "gapmm"]),
    ("two_files", ["-t", "0.03", "-l", "500", "-s", "ref.fa", "-f", "csvn", "in.fa"]),
    ("fst_verbose", ["-t", "0.03", "-l", "500", "-s", "ref.fa", "-m", "in.fa"]),
    ("hyphy_stdout", ["-q", "-t", "0.03", "-l", "500", "-f", "hyphy", "small.fa"]),
    ("count", ["-t", "0.02", "-l", "500", "-c", "in.fa"]),
    ("bad_overlap", ["-t", "0.02", "-l", "0", "in.fa"]),
    ("bad_format", ["-q", "-f", "xml", "in.fa"]),
    ("missing_file", ["-q", "nonexistent.fa"]),
    ("unequal_lengths", ["-q", "-t", "0.02", "ragged.fa"]),
    ("version", ["-v"]),
]


def test_process_interface(tmp_path):
    """tn93g exec'd as its own process once per call, as HIV-TRACE and the reference's own scripts call tn93
    (run_network.py:16-17, evaluate.py:75-76): exit code, stdout, stderr and the output file against the oracle
    CLI and, where it was built, the reference solver itself (oracle/_ref), byte for byte."""
    codes = gen.c1s(360, seed=131)
    gen.write_fasta(tmp_path / "in.fa", codes[:240])
    gen.write_fasta(tmp_path / "ref.fa", codes[240:], gen.names_for(120, "r"))
    gen.write_fasta(tmp_path / "small.fa", codes[:30])
    (tmp_path / "ragged.fa").write_bytes(b">a\nACGTACGT\n>b\nACGTAC\n")
    witnesses = [_oracle_cli()]
    ref = REPO / "oracle" / "_ref" / "swifttn93"
    if ref.exists():
        witnesses.append(ref)
    for name, argv in PROCESS_CASES:
        got_dir = tmp_path / f"g_{name}"
        got_dir.mkdir()
        for f in ("in.fa", "ref.fa", "small.fa", "ragged.fa"):
            os.link(tmp_path / f, got_dir / f)
        print(f"process case {name}: {' '.join(argv)}", flush=True)
        stdin = open(tmp_path / "in.fa", "rb") if name == "stdin" else None
        got = _exec(argv, got_dir, stdin=stdin)
        got_file = (got_dir / "out.csv").read_bytes() if (got_dir / "out.csv").exists() else None
        for w in witnesses:
            wd = tmp_path / f"w_{name}_{w.name}"
            wd.mkdir()
            for f in ("in.fa", "ref.fa", "small.fa", "ragged.fa"):
                os.link(tmp_path / f, wd / f)
            stdin = open(tmp_path / "in.fa", "rb") if name == "stdin" else None
            exp = _exec(argv, wd, stdin=stdin, exe=w)
            exp_file = (wd / "out.csv").read_bytes() if (wd / "out.csv").exists() else None
            assert got[0] == exp[0], (name, w.name, got[0], exp[0], got[2][-500:])
            if name in ("bad_overlap", "bad_format", "missing_file"):  # after the usage line: the same message
                assert got[2].endswith(exp[2].split(b"tn93: error:")[-1]), (name, got[2][-300:], exp[2][-300:])
                continue
            assert got[1] == exp[1], (name, w.name, "stdout", got[1][:300], exp[1][:300])
            assert got[2] == exp[2], (name, w.name, "stderr", got[2][:600], exp[2][:600])
            assert got_file == exp_file, (name, w.name, "output file")


def _long_codes(seed, n, L):
    """SARS-CoV-2-shaped sequences of L columns (N blocks, low divergence) with more ambiguity codes (B/D/H/V
    among them: order-dependent pairs) and long gap runs at either end."""
    codes = gen.c2(seed=seed, n=n, L=L)
    rs = gen.Stream(seed + 1)
    gen._ambiguities(rs, codes, 0.002, gen._C1_POOL)
    gen._gap_runs(rs, codes, 0.4, 50, 3000)
    return codes


LONG_CONFIGS = [
    ("-t", "1", "-a", "resolve", "-g", "0.05", "-l", "25000"),
    ("-t", "0.001", "-a", "resolve", "-l", "60000"),
    ("-t", "1", "-a", "average", "-l", "25000"),
    ("-t", "1", "-a", "skip", "-l", "25000", "-f", "csvn"),
    ("-t", "1", "-a", "gapmm", "-l", "25000"),
    ("-t", "1", "-a", "RYN", "-g", "0.5", "-l", "25000"),
]


@pytest.mark.parametrize("L", [65534, 65535, 70000])
def test_long_alignment_over_65535_vs_oracle(L, tmp_path):
    """Alignments at and beyond 65,535 columns, where the reference leaves its packed kernels for the per-pair
    slow path (swifttn93.cpp:580-590): 48 sequences, every -a mode, every pair written, against the oracle."""
    codes = _long_codes(140 + L % 7, 48, L)
    fa = tmp_path / "in.fa"
    gen.write_fasta(fa, codes)
    cases, want = [], []
    for ci, cfg in enumerate(LONG_CONFIGS):
        d = tmp_path / f"k{ci}"
        d.mkdir()
        p = subprocess.run([str(_oracle_cli()), "-q", *cfg, "-o", str(d / "want.txt"), str(fa)],
                           capture_output=True, check=True, timeout=600)
        want.append(((d / "want.txt").read_bytes(), p.stdout))
        cases.append((d, ["-q", *cfg, "-o", "out.txt", str(fa)]))
    for (d, argv), (rc, out, so), (wo, ws) in zip(cases, run_jobs(cases, tmp_path), want):
        assert rc == 0, argv
        assert out == wo, (L, argv, len(out or b""), len(wo))
        assert so == ws, (L, argv, so[:400], ws[:400])


def _oracle_matrix(codes, mode, min_overlap, rf):
    lib = ctypes.CDLL(str(REPO / "oracle" / "_build" / "liboracle.so"))
    lib.oracle_init.argtypes = [ctypes.c_char_p]
    lib.oracle_triangle.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_long,
                                    ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_init(None)
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    n, L = c.shape
    out = np.zeros((n, n))
    lib.oracle_triangle(c.ctypes.data, n, L, mode, min_overlap, rf, out.ctypes.data, None)
    return out


def _check_served(reqs, everything, thr, mind=-1e-10):
    want = _oracle_matrix(everything, 0, 500, 0.05)
    for r in reqs:
        assert r.error is None, r.error
        cols = np.arange(r.n1, r.n2)
        rows_e, cols_e, d_e = [], [], []
        for i in range(r.n2):  # tn93's order: rows ascending, then columns
            js = cols[cols > i]
            d = want[i, js]
            m = (d >= mind) & (d <= thr)
            rows_e += [i] * int(m.sum())
            cols_e += list(js[m])
            d_e += list(d[m])
        res = r.result
        assert list(res["link_row"]) == rows_e and list(res["link_col"]) == cols_e, r.rid
        assert ["%g" % x for x in res["link_dist"]] == ["%g" % x for x in d_e], r.rid


@pytest.mark.parametrize("merge_rows", [0, 1024])
def test_serving_vs_oracle(merge_rows):
    """Served batches (appended to a 1,200-sequence device panel, each scored against the panel and all earlier
    queries) against the oracle's all-pairs triangle over everything received, restricted to each batch's
    columns; with merging, batches waiting together share one APPEND run and are split back by column."""
    from swifttn93_amd import serve
    panel = gen.c1s(1200, seed=91)
    # arrivals within a few milliseconds: batches wait behind the first runs, so a free worker finds several
    stream = serve.query_stream(panel, seed=92 + merge_rows, rate=2000.0, seconds=0.01, max_batch=120)
    nq = sum(c.shape[0] for _, c in stream)
    srv = serve.Server(panel, capacity=1200 + nq, workers=1 if merge_rows else 2, hbm_budget=1 << 30,
                       run_kw=dict(threshold=0.03), merge_rows=merge_rows)
    try:
        reqs = serve.serve(srv, stream)
    finally:
        srv.close()
    print(f"{len(reqs)} batches, {nq} queries, groups {[r.group for r in reqs]}", flush=True)
    assert len(reqs) >= 6
    if merge_rows:
        assert any(r.group > 1 for r in reqs), "no batches were merged"
    everything = np.concatenate([panel] + [c for _, c in stream])
    _check_served(reqs, everything, 0.03)


def test_serving_full_panel_alone_vs_mixed():
    """The serving workload's 50,000 x 1,300 panel: a short Poisson stream with merging and 4 workers; every
    batch's links in the mix = the batch computed alone afterwards (row, column and printed distance)."""
    from swifttn93_amd import serve
    panel = gen.c1(n=50000, clusters=50000 * 200 // 6700)
    stream = serve.query_stream(panel, seed=8, rate=20.0, seconds=1.5)
    nq = sum(c.shape[0] for _, c in stream)
    srv = serve.Server(panel, capacity=50000 + nq, workers=4, hbm_budget=32 << 30, merge_rows=2048)
    try:
        reqs = serve.serve(srv, stream)
        assert all(r.error is None for r in reqs), [r.error for r in reqs if r.error]
        for r in reqs:
            print(f"request {r.rid}: rows [{r.n1}, {r.n2}) group {r.group}", flush=True)
            alone = srv.run_alone(r)
            for k in ("link_row", "link_col"):
                assert (alone[k] == r.result[k]).all(), (r.rid, k)
            assert ["%g" % x for x in alone["link_dist"]] == ["%g" % x for x in r.result["link_dist"]], r.rid
            assert r.result["certified"] or r.result["exact"]
    finally:
        srv.close()


def _haplotypes(L, seed):
    """Four haplotypes of L sites: h1 12 sites from h0, h2 30 other sites from h0, h3 one site from h2."""
    rs = np.random.default_rng(seed)
    h0 = rs.integers(0, 4, L).astype(np.uint8)
    sites = rs.permutation(L)
    h1, h2 = h0.copy(), h0.copy()
    h1[sites[:12]] = (h1[sites[:12]] + 1) % 4
    h2[sites[12:42]] = (h2[sites[12:42]] + 2) % 4
    h3 = h2.copy()
    h3[sites[42]] = (h3[sites[42]] + 1) % 4
    return np.stack([h0, h1, h2, h3])


def test_more_than_2_32_pairs():
    """A triangle of 92,700 sequences (2^32 + 1,597,354 pairs) runs in parts.  The sequences are copies of four
    haplotypes (92,000 + 600 + 50 + 50, shuffled), so the run equals the four haplotypes with those copy numbers:
    the same histogram, link count and weight sum, an overlapping mean enclosure, and as links exactly the h2-h3
    pairs, in row order, at the h2-h3 distance."""
    L, m = 200, [92000, 600, 50, 50]
    H = _haplotypes(L, 31)
    lab = np.random.default_rng(32).permutation(np.repeat(np.arange(4), m))
    n = lab.size
    assert n * (n - 1) // 2 > 2 ** 32
    codes = H[lab]
    kw = dict(mode=T.RESOLVE, threshold=0.015, min_distance=1e-9, min_overlap=10)
    eng = T.Engine(0)
    try:
        small = eng.upload(H, weights=np.array(m))
        want = eng.run(small, **kw)
        print("four haplotypes: links", want["found_links"], "hist nonzero", int((want["hist"] > 0).sum()), flush=True)
        small.free()
        s = eng.upload(codes)
        got = eng.run(s, **kw)
        s.free()
    finally:
        eng.close()
    assert np.array_equal(got["hist"], want["hist"])
    assert got["found_links"] == want["found_links"] == 50 * 50
    assert got["weighted"][0] == want["weighted"][0] == n * (n - 1) / 2
    assert got["skipped"] == want["skipped"] == 0
    assert got["mean_lo"][0] <= want["mean_hi"][0] and want["mean_lo"][0] <= got["mean_hi"][0]
    i_idx, j_idx = [], []
    pos2, pos3 = np.flatnonzero(lab == 2), np.flatnonzero(lab == 3)
    for i in np.sort(np.concatenate([pos2, pos3])):
        other = pos3 if lab[i] == 2 else pos2
        js = other[other > i]
        i_idx += [i] * js.size
        j_idx += list(js)
    assert np.array_equal(got["link_row"], np.array(i_idx)) and np.array_equal(got["link_col"], np.array(j_idx))
    d23 = want["link_dist"][0]
    assert (got["link_dist"] == d23).all() and len(want["link_dist"]) == 1


SPLIT_KINDS = [
    ("triangle", dict(mode=T.RESOLVE, resolve_fraction=0.05)),
    ("triangle", dict(mode=T.AVERAGE)),
    ("triangle", dict(mode=T.GAPMM)),
    ("triangle", dict(mode=T.RESOLVE, resolve_fraction=0.05, flags=T.FLAG_EXACT)),
    ("cross", dict(mode=T.RESOLVE, resolve_fraction=0.05)),
    ("append", dict(mode=T.SKIP)),
]


@pytest.mark.parametrize("k", range(len(SPLIT_KINDS)))
def test_split_equals_one_run(k):
    """Runs cut into parts of at most 20,000 pairs (TN93G_PART_PAIRS) give the whole run's result field for field:
    links, histogram, the mean's enclosure, maxima, counts; with copy numbers and replayed B/D/H/V pairs."""
    kind, kw = SPLIT_KINDS[k]
    codes = gen.c1s(600, seed=33)
    w = np.random.default_rng(34).integers(1, 4, 600)
    kw = dict(kw, threshold=0.03, min_overlap=300)
    if kind == "cross":
        kw.update(kind=T.CROSS, n1=250)
    elif kind == "append":
        kw.update(kind=T.APPEND, n1=420, n2=600)
    eng = T.Engine(0)
    try:
        s = eng.upload(codes, weights=w)
        whole = eng.run(s, **kw)
        os.environ["TN93G_PART_PAIRS"] = "20000"
        try:
            parts = eng.run(s, **kw)
        finally:
            del os.environ["TN93G_PART_PAIRS"]
        s.free()
    finally:
        eng.close()
    print(kind, kw, "links", whole["found_links"], "replayed", whole["n_replayed"], "exact", whole["exact"], flush=True)
    for f in ["hist", "link_row", "link_col", "link_dist", "link_lo", "link_hi"]:
        assert np.array_equal(np.asarray(parts[f]), np.asarray(whole[f])), f
    for f in ["weighted", "max_lo", "max_hi", "found_links", "skipped", "exact", "n_replayed"]:
        assert parts[f] == whole[f], (f, parts[f], whole[f])
    # the fast path's row sums are added by atomics, in no fixed order: the mean's enclosures may differ in their last
    # bits from run to run, and must overlap and agree to far more digits than are printed
    lo, hi, lo1, hi1 = parts["mean_lo"][0], parts["mean_hi"][0], whole["mean_lo"][0], whole["mean_hi"][0]
    assert lo <= hi1 and lo1 <= hi and abs(lo - lo1) <= 1e-12 * abs(lo1) and abs(hi - hi1) <= 1e-12 * abs(hi1)
    if whole["exact"]:
        assert parts["mean_lo"] == whole["mean_lo"] and parts["mean_hi"] == whole["mean_hi"]
    assert len(whole["link_row"]) > 0 and (whole["n_replayed"] > 0 or kw["mode"] == T.SKIP)


def test_split_cli_vs_oracle(tmp_path):
    """The command with its runs cut into parts of at most 5,000 pairs, byte for byte against the oracle: every
    pair written, -s, and copy numbers."""
    codes = gen.c1s(300, seed=35)
    fa, fb = tmp_path / "in.fa", tmp_path / "ref.fa"
    gen.write_fasta(fa, codes[:200], [f"s{i}:{1 + i % 3}" for i in range(200)])
    gen.write_fasta(fb, codes[200:], [f"r{i}" for i in range(100)])
    cases = [["-q", "-t", "1", "-f", "csvn", "-a", "resolve", "-g", "0.05", "-l", "500"],
             ["-q", "-t", "0.03", "-a", "average", "-l", "500", "-s", str(fb)],
             ["-q", "-t", "0.03", "-a", "gapmm", "-l", "500", "-0"]]
    env = dict(os.environ, TN93G_PART_PAIRS="5000")
    for k, argv in enumerate(cases):
        want = subprocess.run([str(_oracle_cli()), *argv, "-o", str(tmp_path / f"w{k}"), str(fa)], capture_output=True,
                              check=True)
        p = subprocess.run([str(_cli()), *argv, "-o", str(tmp_path / f"g{k}"), str(fa)], capture_output=True,
                           env=env, timeout=300)
        assert p.returncode == 0, p.stderr[-500:]
        assert (tmp_path / f"g{k}").read_bytes() == (tmp_path / f"w{k}").read_bytes(), argv
        assert p.stdout == want.stdout, (argv, p.stdout[-300:], want.stdout[-300:])