"""The command line's multi-threaded FASTA reader (swifttn93_amd/csrc/tn93g_fasta.h) against the oracle's sequential
restatement of tn93's readFASTA (swifttn93.cpp:292-406, length rules 259-273), CPU only.

Random inputs cover what the state machine distinguishes: text before the first record, '>' and '#' record starts,
'>' and '#' inside names and in the middle of sequence lines, \\n, \\r\\n and \\r line ends, lower case, IUPAC codes,
gaps, dropped bytes, copy-number suffixes after several separators, and the error exits (empty name, end of file
inside a name or before any record, an empty first sequence, unequal lengths, within one file or across two).
Every case must give the same exit status, the same message on stderr, and the same names, copy numbers and codes,
at 1, 3 and 8 threads and with scan chunks small enough that records and names straddle chunk boundaries.
"""
import random
import subprocess

import pytest

from conftest import REPO

ALPHA = "ACGTURYSWKMBDHVN?-"


@pytest.fixture(scope="module")
def fasta_check(oracle_bins, tmp_path_factory):
    _, lib = oracle_bins
    exe = tmp_path_factory.mktemp("fasta") / "fasta_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", str(exe),
                    str(REPO / "tests" / "native" / "fasta_check.cpp"), f"-L{lib.parent}", "-loracle",
                    f"-Wl,-rpath,{lib.parent}"], check=True)
    return exe


def _seq(rng, L, dirty):
    s = []
    for _ in range(L):
        r = rng.random()
        c = rng.choice("ACGT") if r < 0.9 else rng.choice(ALPHA)
        if dirty and rng.random() < 0.1:
            c = c.lower()
        s.append(c)
    out = "".join(s)
    if dirty:  # dropped bytes and line breaks inside the sequence
        parts, k = [], 0
        while k < len(out):
            step = rng.randint(1, 80)
            parts.append(out[k:k + step])
            k += step
            if rng.random() < 0.3:
                parts.append(rng.choice([" ", "*", "1", "\t", ".", "\x00", "\xff"]))
            parts.append(rng.choice(["\n", "\r\n", "\r", ""]))
        out = "".join(parts)
    return out


def _name(rng, seps):
    base = "".join(rng.choice("abcXYZ_|.-0123456789") for _ in range(rng.randint(1, 12)))
    r = rng.random()
    if r < 0.3:
        base += rng.choice(seps) + str(rng.randint(-3, 40))
    elif r < 0.4:
        base += rng.choice(seps) + rng.choice(["", "x", "7a", " 5", "0"])
    elif r < 0.5:
        base = rng.choice([">", "#", "a>b", "#c#"]) + base
    return base


def _fasta(rng, n, L, seps, faults):
    eol = rng.choice(["\n", "\r\n", "\r"])
    head = rng.choice(["", "junk before\n", ";comment\n", "\n\n"])
    recs = []
    for k in range(n):
        Lk = L
        if "len" in faults and rng.random() < 0.15:
            Lk = max(0, L + rng.choice([-2, -1, 1, 3]))
        name = _name(rng, seps)
        if "empty" in faults and rng.random() < 0.1:
            name = ""
        start = rng.choice([">", ">", ">", "#"])
        recs.append(start + name + eol + _seq(rng, Lk, rng.random() < 0.5) + eol)
    text = head + "".join(recs)
    if "eof_name" in faults:
        text += ">" + _name(rng, seps)
    if "midline" in faults and len(text) > 10:  # a record start in the middle of a sequence line
        k = rng.randrange(len(text) // 2, len(text))
        text = text[:k] + ">" + _name(rng, seps) + "\n" + _seq(rng, L, False) + text[k:]
    return text.encode("latin-1")


def _cases(tmp_path, seed, count):
    rng = random.Random(seed)
    lines = []
    for c in range(count):
        n = rng.choice([1, 2, 3, 5, 12, 40])
        L = rng.choice([1, 2, 7, 33, 100, 257])
        faults = set(rng.sample(["len", "empty", "eof_name", "midline"], rng.choice([0, 0, 0, 1, 2])))
        seps = rng.choice([":", "|", "_"])
        p1 = tmp_path / f"c{seed}_{c}_1.fa"
        r = rng.random()
        if r < 0.03:
            p1.write_bytes(b"")
        elif r < 0.06:
            p1.write_bytes(b"no records here\nACGT\n")
        elif r < 0.09:
            p1.write_bytes(b">first\n\n>second\nACGT\n")  # an empty first sequence
        else:
            p1.write_bytes(_fasta(rng, n, L, [seps], faults))
        p2 = "-"
        if rng.random() < 0.25:
            q = tmp_path / f"c{seed}_{c}_2.fa"
            L2 = L if rng.random() < 0.7 else L + 1
            q.write_bytes(_fasta(rng, rng.choice([1, 4]), L2, [seps], set()))
            p2 = str(q)
        lines.append(f"{p1}\t{p2}\t{seps}")
    man = tmp_path / f"manifest{seed}.txt"
    man.write_text("\n".join(lines) + "\n")
    return man


@pytest.mark.parametrize("threads,min_chunk", [(1, 1 << 20), (3, 7), (8, 1), (8, 64)])
def test_reader_matches_oracle(fasta_check, tmp_path, threads, min_chunk):
    man = _cases(tmp_path, seed=threads * 1000 + min_chunk, count=400)
    out = subprocess.run([str(fasta_check), str(man), str(threads), str(min_chunk)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-2000:]
    assert "mismatches 0" in out.stdout


def test_reader_real_set(fasta_check, hiv8766, tmp_path):
    """The reference's real 8,766 x 1,065 HIV-1 pol set, whole and split over two files."""
    data = hiv8766.read_bytes()
    cut = data.index(b">", len(data) // 2)
    a, b = tmp_path / "a.fa", tmp_path / "b.fa"
    a.write_bytes(data[:cut])
    b.write_bytes(data[cut:])
    man = tmp_path / "m.txt"
    man.write_text(f"{hiv8766}\t-\t:\n{a}\t{b}\t:\n{a}\t{b}\t|\n")
    for threads, chunk in [(8, 4096), (16, 1 << 20)]:
        out = subprocess.run([str(fasta_check), str(man), str(threads), str(chunk)], capture_output=True, text=True)
        assert out.returncode == 0 and "cases 3 mismatches 0" in out.stdout, out.stdout[-3000:]
