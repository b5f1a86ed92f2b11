#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference build (oracle/_ref/, see oracle/Makefile).

Run in a container that holds the reference tree (``make -C oracle ref`` first):

    python3 tests/golden/make_fixtures.py

Outputs (all data, no code from the reference):
  fuzz_cases.json.xz   adversarial tiny cases: input FASTA bytes, flags, and the reference's exit code, output file
                       and stdout (the gate of evaluate.py:61-99: -q, -o OUT, FASTA last, byte-compared).
  handler_table.txt.xz the reference ambiguity_handler contribution of every (mode, c1, c2) (ref_harness table).
  pair_vectors.json.xz per-pair inputs (regenerated from a seed by pair_inputs()) and the reference computeTN93
                       outputs (distance bits, histogram index), with and without gap descriptors.
  real/                the reference's real HIV-1 pol sets (data/real_hiv_pol_{dev,holdout}.fa.xz) and the ten
                       published tn93 sha256 pins of application/results/network.jsonl; its real SARS-CoV-2 sets
                       (data/real_sars2_{dev,holdout}.fa.xz, 500 and 459 genomes x 29,903 columns, two thirds of
                       them constant) and sars2_pins.json: the reference's sha256 of file and stdout for SARS2_FLAGS.

``python3 tests/golden/make_fixtures.py sars2`` writes only the SARS-CoV-2 part.

The fuzz distribution mirrors the reference harness's generator (gen_data.py:156-250) in our own code: sizes,
alphabets (IUPAC, lowercase, U, ?, weird bytes), CRLF, wrapping, empty/duplicate/odd/copy-number names, unequal
lengths, leading gaps, near-duplicates, random flags and a second file.
"""
import hashlib
import json
import lzma
import os
import random
import shutil
import struct
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
REF_BIN = REPO / "oracle" / "_ref" / "swifttn93"
HARNESS = REPO / "oracle" / "_ref" / "ref_harness"
REFTREE = Path(os.environ.get("SWIFTTN93_REF_TREE", "/root/reference"))

# subset strings of the -a SUBSET mode, indexed by the subset id of the per-pair records (ref_harness.cpp SUBSETS)
SUBSETS_FOR_IDS = ["RY", "RYSWKM", "N", "BDHV", "RYN?", "ACGT"]

CORE = b"ACGT"
EXTRA = b"acgtRYSWKMBDHVNnrykm-"
EXTRA3 = b"BDHVbdhvRYN"
WEIRD = b"uU?.X*~ "


def fuzz_records(rng, n, L, extra_p, names_mode, equal=True):
    weird_p = extra_p * 0.1 if rng.random() < 0.1 else 0.0
    heavy3 = rng.random() < 0.2  # exercise the 1/3-weight (order-dependent) codes more often
    recs, names = [], []
    for i in range(n):
        if names_mode == "dup" and i and rng.random() < 0.4:
            nm = names[rng.randrange(len(names))]
        elif names_mode == "copies":
            nm = f"s{i}:{rng.randint(1, 4)}" if rng.random() < 0.8 else f"s{i}"
        elif names_mode == "odd" and rng.random() < 0.3:
            nm = rng.choice(["", " ", "a b", "x,y", "q;r", "s\tt", "long" * 20, "c:-3", "c:0", "c:"])
        else:
            nm = f"seq{i}"
        names.append(nm)
        Li = L if equal or rng.random() < 0.7 else max(0, L + rng.randint(-5, 5))
        out = bytearray()
        for _ in range(Li):
            if rng.random() < weird_p:
                out.append(rng.choice(WEIRD))
            elif rng.random() < extra_p:
                out.append(rng.choice(EXTRA3 if heavy3 and rng.random() < 0.5 else EXTRA))
            else:
                out.append(rng.choice(CORE))
        if rng.random() < 0.15:
            a = rng.randint(0, max(0, Li // 3))
            out[:a] = b"-" * a
        if rng.random() < 0.1 and Li > 40:  # an internal gap run long enough for tn93's jump markers
            a = rng.randrange(Li)
            b = min(Li, a + rng.randint(18, 300))
            out[a:b] = b"-" * (b - a)
        recs.append((nm, bytes(out)))
    if rng.random() < 0.5 and n >= 2:
        base = recs[0][1]
        for j in range(1, n):
            if rng.random() < 0.5 and len(recs[j][1]) == len(base):
                s = bytearray(base)
                for _ in range(rng.randint(0, 3)):
                    if s:
                        s[rng.randrange(len(s))] = rng.choice(b"ACGTRN-")
                recs[j] = (recs[j][0], bytes(s))
    return recs


def render(recs, rng):
    wrap = rng.choice([0, 0, 0, 7, 60])
    nl = b"\r\n" if rng.random() < 0.05 else b"\n"
    buf = bytearray()
    for nm, s in recs:
        buf += (b"#" if rng.random() < 0.02 else b">") + nm.encode() + nl
        if rng.random() < 0.05:
            buf += nl
        if wrap and s:
            for i in range(0, len(s), wrap):
                buf += s[i:i + wrap] + nl
        else:
            buf += s + nl
    if rng.random() < 0.05:
        buf = buf.rstrip(b"\r\n")
    return bytes(buf)


def fuzz_case(rng):
    r = rng.random()
    if r < 0.02:
        return {"in": b""}, []
    if r < 0.04:
        return {"in": rng.choice([b">", b">a\n", b">a\nACGT\n>b\n", b"ACGT\n", b">a\nAC\n>b\nACG\n"])}, ["-l", "1"]
    n = rng.choice([1, 2, 2, 3, 4, 5, 6, 8, 12, 20])
    L = rng.choice([1, 2, 5, 17, 40, 64, 65, 100, 127, 128, 129, 200, 300, 150, 250, 320, 400])
    extra_p = rng.choice([0.0, 0.0, 0.02, 0.1, 0.4])
    names_mode = rng.choice(["plain", "plain", "dup", "odd", "copies"])
    equal = rng.random() > 0.03
    files = {"in": render(fuzz_records(rng, n, L, extra_p, names_mode, equal), rng)}
    flags = ["-t", rng.choice(["0", "0.001", "0.015", "0.05", "0.1", "0.3", "0.75", "1", "2"])]
    flags += ["-l", str(rng.choice([1, 1, 1, 2, max(1, L // 4), max(1, L // 2), L, L + 1, 100]))]
    if rng.random() < 0.6:
        flags += ["-a", rng.choice(["resolve", "average", "skip", "gapmm", "RY", "RYSWKM", "N", "bdhv"])]
        if rng.random() < 0.4:
            flags += ["-g", rng.choice(["0", "0.01", "0.05", "0.5", "1"])]
    if rng.random() < 0.15:
        flags += ["-w", rng.choice(["0", "0.001", "0.02"])]
    fmt = rng.choice(["csv", "csv", "csvn", "hyphy"])
    flags += ["-f", fmt]
    for f, p in [("-c", 0.1), ("-n", 0.2), ("-0", 0.2)]:
        if rng.random() < p:
            flags.append(f)
    if rng.random() < 0.15:
        flags += ["-D", rng.choice([";", "|", ":"])]
    if names_mode == "copies" or rng.random() < 0.05:
        flags += ["-d", rng.choice([":", ":", "_"])]
    if rng.random() < 0.2:
        n2 = rng.choice([1, 2, 3, 5, 8])
        L2 = L if rng.random() < 0.95 else L + 1
        files["in2"] = render(fuzz_records(rng, n2, L2, extra_p, rng.choice(["plain", "dup", "copies"])), rng)
        flags += ["-s", "{in2}"]
        if rng.random() < 0.3:
            flags.append("-m")
    if rng.random() < 0.01:
        flags += [rng.choice(["-b", "-H", "-x", "-l"])]
    return files, flags


def run_case(binary, files, flags, td, extra_env=None):
    """Runs `binary -q FLAGS -o out.txt in.fa` in td; returns (rc, out bytes or None, stdout bytes)."""
    names = {k: f"{k}.fa" for k in files}
    for k, data in files.items():
        (td / names[k]).write_bytes(data)
    argv = [str(binary), "-q"] + [names[f.strip("{}")] if f.startswith("{") else f for f in flags]
    argv += ["-o", "out.txt", names["in"]]
    out = td / "out.txt"
    if out.exists():
        out.unlink()
    env = dict(os.environ, OMP_NUM_THREADS="1")
    if extra_env:
        env.update(extra_env)
    p = subprocess.run(argv, cwd=td, capture_output=True, timeout=60, env=env)
    return p.returncode, (out.read_bytes() if out.exists() else None), p.stdout


def make_fuzz(n_cases, seed):
    rng = random.Random(seed)
    cases = []
    with tempfile.TemporaryDirectory() as t:
        td = Path(t)
        for k in range(n_cases):
            files, flags = fuzz_case(rng)
            rc, out, so = run_case(REF_BIN, files, flags, td)
            cases.append({
                "files": {key: v.decode("latin-1") for key, v in files.items()},
                "flags": flags,
                "rc": rc,
                "out": None if out is None else out.decode("latin-1"),
                "stdout": so.decode("latin-1"),
            })
    return cases


# ------------------------------------------------------------------ per-pair vectors

def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def pair_inputs(n, seed):
    """Deterministic per-pair inputs: list of (mode, subset_id, L, min_overlap, rf, s1 bytes, s2 bytes).

    Codes are drawn to stress the order-dependent paths: clean bases, every IUPAC code (incl. B/D/H/V, U, N, ?),
    and gap runs long enough for tn93's jump markers."""
    st = seed
    out = []

    def nxt(m):
        nonlocal st
        st, z = _splitmix(st)
        return z % m

    for _ in range(n):
        mode = nxt(5)
        sid = nxt(6)
        L = 1 + nxt(400)
        amb_p = [0, 2, 10, 40, 150][nxt(5)]  # per mille
        rows = []
        for _r in range(2):
            s = bytearray()
            p = 0
            while p < L:
                u = nxt(1000)
                if u < 15:  # a gap run
                    run = 1 + nxt(80)
                    s += bytes([17]) * min(run, L - p)
                    p += min(run, L - p)
                    continue
                if u < 15 + amb_p:
                    s.append(4 + nxt(13))
                else:
                    s.append(nxt(4))
                p += 1
            rows.append(bytes(s))
        if nxt(3) == 0:  # near-duplicate second row
            s = bytearray(rows[0])
            for _k in range(nxt(4)):
                s[nxt(L)] = [0, 1, 2, 3, 5, 15, 11, 17][nxt(8)]
            rows[1] = bytes(s)
        ov = [1, 1, 1 + nxt(L), L, L + 1][nxt(5)]
        rf = [1.0, 1.0, 0.5, 0.05, 0.01, 0.0][nxt(6)]
        out.append((mode, sid, L, ov, rf, rows[0], rows[1]))
    return out


def make_pairs(n, seed):
    recs = pair_inputs(n, seed)
    blob = bytearray()
    for mode, sid, L, ov, rf, s1, s2 in recs:
        blob += struct.pack("<iiqqd", mode, sid, L, ov, rf) + s1 + s2
    p = subprocess.run([str(HARNESS), "pairs"], input=bytes(blob), capture_output=True, check=True)
    lines = p.stdout.decode().split("\n")[:-1]
    assert len(lines) == n
    res = []
    for ln in lines:
        a, b, c, d = ln.split()
        res.append([a, int(b), c, int(d)])
    return {"seed": seed, "n": n, "results": res}


# option sets of the SARS-CoV-2 pins: the reference's benchmark flags (-t 0.0002 -l 20000 -0), every pair written,
# each -a mode, an overlap bound that skips pairs, and the two sets against each other (-s, and -s -m)
SARS2_FLAGS = [
    ["-t", "0.0002", "-l", "20000", "-0"],
    ["-t", "1", "-a", "resolve", "-g", "0.05", "-l", "20000", "-f", "csvn"],
    ["-t", "0.001", "-a", "average", "-l", "20000"],
    ["-t", "0.001", "-a", "skip", "-l", "5000"],
    ["-t", "0.001", "-a", "gapmm", "-l", "20000"],
    ["-t", "0.0005", "-a", "resolve", "-g", "0.05", "-l", "29800"],
    ["-t", "0.0005", "-a", "resolve", "-g", "0.05", "-l", "20000", "-s", "{second}"],
    ["-t", "0.0005", "-a", "average", "-l", "20000", "-s", "{second}", "-m"],
]


def make_sars2():
    real = HERE / "real"
    real.mkdir(exist_ok=True)
    names = ["real_sars2_dev.fa.xz", "real_sars2_holdout.fa.xz"]
    for f in names:
        shutil.copyfile(REFTREE / "data" / f, real / f)
    files = {"in": lzma.decompress((real / names[0]).read_bytes()),
             "second": lzma.decompress((real / names[1]).read_bytes())}
    pins = []
    with tempfile.TemporaryDirectory() as t:
        for flags in SARS2_FLAGS:
            rc, out, so = run_case(REF_BIN, files, flags, Path(t))
            assert rc == 0 and out is not None, (flags, rc)
            pins.append({"flags": flags, "ref_file_sha256": hashlib.sha256(out).hexdigest(),
                         "ref_stdout_sha256": hashlib.sha256(so).hexdigest(), "file_bytes": len(out)})
            print("sars2", " ".join(flags), len(out))
    (real / "sars2_pins.json").write_text(json.dumps({"files": {"in": names[0], "second": names[1]},
                                                      "pins": pins}, indent=1))


def main():
    if not REF_BIN.exists() or not HARNESS.exists():
        sys.exit("build the reference first: make -C oracle ref")
    if sys.argv[1:] == ["sars2"]:
        make_sars2()
        return
    seed = 20261203
    cases = make_fuzz(int(os.environ.get("N_FUZZ", "600")), seed)
    (HERE / "fuzz_cases.json.xz").write_bytes(lzma.compress(json.dumps({"seed": seed, "cases": cases}).encode()))
    print("fuzz cases:", len(cases), "nonzero rc:", sum(c["rc"] != 0 for c in cases))
    tables = {}
    for sub in ["RY", "RYSWKM", "N", "BDHV", "RYN?", "ACGT"]:
        p = subprocess.run([str(HARNESS), "table", sub], capture_output=True, check=True)
        tables[sub] = p.stdout.decode()
    (HERE / "handler_table.json.xz").write_bytes(lzma.compress(json.dumps(tables).encode()))
    pv = make_pairs(int(os.environ.get("N_PAIRS", "12000")), 777)
    (HERE / "pair_vectors.json.xz").write_bytes(lzma.compress(json.dumps(pv).encode()))
    print("pair vectors:", pv["n"])
    real = HERE / "real"
    real.mkdir(exist_ok=True)
    for f in ["real_hiv_pol_dev.fa.xz", "real_hiv_pol_holdout.fa.xz"]:
        shutil.copyfile(REFTREE / "data" / f, real / f)
    pins = []
    for line in open(REFTREE / "application" / "results" / "network.jsonl"):
        d = json.loads(line)
        pins.append({k: d[k] for k in ["threshold", "t_arg", "ref_file_sha256", "ref_stdout_sha256", "file_bytes"]})
    (real / "network_pins.json").write_text(json.dumps({
        "input_sha256": "70985c02b3ba52585c9ed30db147ca14984a9f5fd03babb3af235432062a0524",
        "flags": ["-q", "-0", "-a", "resolve", "-l", "500", "-g", "0.05", "-f", "csv"],
        "pins": pins}, indent=1))
    make_sars2()


if __name__ == "__main__":
    main()
