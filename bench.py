#!/usr/bin/env python3
"""Benchmark: byte-exact all-pairs TN93 pairs/s on one MI355X, with the reference CPU path timed beside it.

Workload (one step = one pass of the hot path over the set): the HIV-pol-shaped set of 6,700 sequences x 1,300
columns (swifttn93_amd/gen.py c1), options -t 0.015 -a resolve -g 0.05 -l 500, 22,441,650 pairs.  The packed set is
resident in HBM before the timed region; a step is one tn93g_run: clean/clean counts, ambiguity events,
site-order replays, FP64 tail, histogram, per-row sums in tn93's order and the link list copied back to the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--backend popc|mfma] [--no-cpu] [--config c1|c2|c3|c4|sars2]
                    [--mode resolve|average|skip|gapmm]

--mode runs the same 6,700 x 1,300 set with another -a mode (-g 0.05 only with resolve; the default line is
resolve).

Other workloads (--config): c2, the SARS-CoV-2-shaped long alignment (20,000 x 29,903, N-masked blocks,
-t 0.0005 -a resolve -g 0.05 -l 25000); c3, 64 uneven alignments with mixed -a modes run as one batch (one graph
launch) beside the same jobs run one by one, and the 2,000 x 40,000 -s bipartite block; c4, serving: a 50,000 x
1,300 device-resident panel answering a Poisson stream of query batches (p50/p99 latency, pairs/s); sars2, the
reference's real SARS-CoV-2 dev set (500 x 29,903, two thirds of the columns constant), folded and not.

Prints one JSON line.  Beside `value` (the engine with the set resident) it carries:
  steady_state     the whole tn93 command in a warm process (tn93g's --batch driver, the same invocation repeated):
                   FASTA read (page cache), packing and upload, the run, certification of the printed values,
                   formatting and the distance file written; cold_cli: the same as a fresh process (device init
                   included), which is what the reference's wall time measures;
  parity           sha256 of the GPU command's distance file and stdout against the reference solver's on the same
                   input and flags (every C1 mode, C2, each C3 job and the -s block, C4's served links against the
                   reference's triangle over everything received);
  roofline         for the pair kernel (HIP events on the engine's stream, averaged over the timed steps);
  cpu_baseline     the reference solver (oracle/_ref/swifttn93_avx512, built from the reference sources by
                   oracle/Makefile) on the same input on this host's cores: best of 3 full-command walls with the
                   threads of the process's affinity set (and OMP_NUM_THREADS when set and different), and one run on
                   1 thread; with the host's CPU model, core counts and AVX-512 flags.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))
SEP = "\x1f"


def mode_flags(mode):
    """tn93's flags for the C1 workload in -a mode (-g 0.05 with resolve, as the headline line is quoted)."""
    return ["-t", "0.015", "-a", mode] + (["-g", "0.05"] if mode == "resolve" else []) + ["-l", "500", "-f", "csv"]


VALU_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 39.3: 256 CUs x 64 lanes x 2.4 GHz
MFMA_INT8_TOPS = 5000.0
HBM_BYTES_S = 8.0e12
REF_VARIANT = ("swifttn93 built from the reference sources (oracle/Makefile): g++ -O3 -march=x86-64-v4 "
               "-mavx512vpopcntdq -mavx512vbmi -mavx512vbmi2 -mavx512ifma -mno-avx512dq -fopenmp -static; its "
               "binary-derived vector log (vlog.h, vlog_data.h) is excluded, which also leaves out its AVX512DQ "
               "code paths, so logarithms are libm's; the reference's own -march=native build with that file "
               "measured about 2.3x faster on the 8 cores where the builds were made")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload
    (profiles/c1_popc_hbm_bytes.md: read MB and written MB columns, already scaled by the copy calibration)."""
    f = REPO / "profiles" / "c1_popc_hbm_bytes.md"
    if not f.exists():
        return None
    for row in f.read_text().splitlines():
        cells = [c.strip() for c in row.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:  # the last two columns: read MB, written MB
            return (float(cells[-2]) + float(cells[-1])) * 1e6
    return None


def host_info():
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                info["model"] = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        flags = set()
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("flags"):
                flags = set(ln.split(":", 1)[1].split())
                break
        info["avx512"] = sorted(f for f in flags if f.startswith("avx512"))
    except OSError:
        pass
    return info


def thread_counts():
    """Thread counts the baseline is timed with: the affinity set, and OMP_NUM_THREADS when set and different."""
    aff = len(os.sched_getaffinity(0))
    out = [aff]
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and 0 < int(omp) != aff:
        out.append(int(omp))
    return out


def ref_exe():
    ref = REPO / "oracle" / "_ref" / "swifttn93_avx512"
    if ref.exists():
        return ref, "reference"
    port = REPO / "oracle" / "_build" / "tn93_oracle"
    return (port, "port") if port.exists() else (None, None)


def sha(b):
    return hashlib.sha256(b).hexdigest()


def run_reference(exe, argvs, threads, keep=None):
    """Wall time of the reference solver over a list of argv lists (each with -o OUT substituted), output deleted
    outside the timer; with keep (a list), the sha256 of each run's output file and stdout are appended to it."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    tot = 0.0
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "out.csv"
        for argv in argvs:
            t0 = time.perf_counter()
            p = subprocess.run([str(exe), "-q", "-o", str(out)] + argv, check=True, env=env, capture_output=True,
                               timeout=1200)
            tot += time.perf_counter() - t0
            if keep is not None:
                keep.append((sha(out.read_bytes()), sha(p.stdout)))
            out.unlink()
    return tot


def cpu_baseline(argvs, n_pairs, sample, one_thread=True):
    """The reference on the given runs: best of 3 at each thread count of thread_counts(), and once on 1 thread.
    Returns (baseline dict or None, [(file sha, stdout sha)] of the reference's outputs)."""
    exe, kind = ref_exe()
    if exe is None:
        return None, None
    keep = []
    best = None
    for th in thread_counts():
        for rep in range(3):
            w = run_reference(exe, argvs, th, keep if not keep_done(keep, argvs) else None)
            if best is None or w < best[0]:
                best = (w, th)
    b = {"value": n_pairs / best[0], "unit": "pairs/s", "cores": best[1], "kind": kind,
         "variant": REF_VARIANT if kind == "reference" else "the oracle (scalar CPU restatement)",
         "sample": sample + f"; best of 3 full-command walls (read + compute + write) at {thread_counts()} threads",
         "wall_s": best[0], "host": host_info()}
    if one_thread:
        w1 = run_reference(exe, argvs, 1)
        b["one_thread"] = {"value": n_pairs / w1, "wall_s": w1, "cores": 1}
    return b, keep


def keep_done(keep, argvs):
    return len(keep) >= len(argvs)


def tn93g_batch(cases, td, timeout=1200):
    """Runs `tn93g --batch` (the cases one after the other in one process, each a full tn93 invocation); returns
    per case (rc, output sha, stdout sha, wall seconds)."""
    bf = Path(td) / "cases.txt"
    bf.write_text("".join(SEP.join([str(d)] + list(a)) + "\n" for d, a in cases))
    subprocess.run([str(REPO / "swifttn93_amd" / "bin" / "tn93g"), "--batch", str(bf)], check=True,
                   capture_output=True, timeout=timeout)
    res = []
    for d, _ in cases:
        d = Path(d)
        out = d / "out.csv"
        res.append((int((d / "rc.txt").read_text()), sha(out.read_bytes()) if out.exists() else None,
                    sha((d / "stdout.bin").read_bytes()), float((d / "wall.txt").read_text())))
    return res


def tn93g_jobs(cases, td, timeout=1200):
    """Runs `tn93g --jobs` (all cases prepared, computed as one tn93g_run_batch, then certified and written);
    returns (per case (rc, output sha, stdout sha), wall seconds of the whole command)."""
    bf = Path(td) / "jobs.txt"
    bf.write_text("".join(SEP.join([str(d)] + list(a)) + "\n" for d, a in cases))
    t0 = time.perf_counter()
    subprocess.run([str(REPO / "swifttn93_amd" / "bin" / "tn93g"), "--jobs", str(bf)], check=True,
                   capture_output=True, timeout=timeout)
    wall = time.perf_counter() - t0
    res = []
    for d, _ in cases:
        d = Path(d)
        out = d / "out.csv"
        res.append((int((d / "rc.txt").read_text()), sha(out.read_bytes()) if out.exists() else None,
                    sha((d / "stdout.bin").read_bytes())))
    return res, wall


def cold_cli(argv, td):
    """Wall time of one fresh tn93g process (device init, read, pack, run, certify, format, write)."""
    out = Path(td) / "cold.csv"
    t0 = time.perf_counter()
    subprocess.run([str(REPO / "swifttn93_amd" / "bin" / "tn93g"), "-q", "-o", str(out)] + argv, check=True,
                   capture_output=True, timeout=1200)
    w = time.perf_counter() - t0
    out.unlink()
    return w


def roofline_for(n_pairs, L, pair_ms):
    words = (L + 31) // 32
    ops = 22 * words
    achieved = ops * n_pairs / (pair_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": VALU_INT32_TOPS, "unit": "TOPS",
            "frac": achieved / VALU_INT32_TOPS, "traffic": None, "kernel": "k_pairs<1, true, true, true>",
            "kernel_ms": pair_ms, "ops_per_pair": ops}


def bench_c2(a, T, gen):
    import numpy as np
    codes = gen.c2()
    n, L = codes.shape
    n_pairs = n * (n - 1) // 2
    eng = T.Engine(0)
    s = eng.upload(codes)
    kw = dict(mode=T.RESOLVE, threshold=0.0005, min_overlap=25000, resolve_fraction=0.05, backend=T.BACKEND_POPC)
    for _ in range(a.warmup):
        r = eng.run(s, **kw)
    t0 = time.perf_counter()
    pk = []
    for _ in range(a.steps):
        r = eng.run(s, **kw)
        pk.append(r["pair_kernel_ms"])
    dt = time.perf_counter() - t0
    value = n_pairs * a.steps / dt
    line = {"metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": dt * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32+f64", "data": "synthetic (seeded SARS-CoV-2-shaped generator, swifttn93_amd/gen.py c2)",
            "config": {"workload": "SARS-CoV-2-shaped long alignment, 20,000 x 29,903, N-masked blocks, "
                                   "-t 0.0005 -a resolve -g 0.05 -l 25000", "pairs": n_pairs, "parallelism": "1 GPU"},
            "roofline": roofline_for(n_pairs, L, float(np.mean(pk))),
            "links": int(r["found_links"]), "replayed_pairs": int(r["n_replayed"])}
    s.free()
    eng.close()
    flags = ["-t", "0.0005", "-a", "resolve", "-g", "0.05", "-l", "25000", "-f", "csv"]
    with tempfile.TemporaryDirectory() as td:
        fa = Path(td) / "c2.fa"
        gen.write_fasta(fa, codes)
        cases = []
        for k in range(3):  # the first case also pays the device's initialisation
            d = Path(td) / f"e{k}"
            d.mkdir()
            cases.append((d, ["-q", *flags, "-o", "out.csv", str(fa)]))
        res = tn93g_batch(cases, td)
        warm = float(np.median([w for _, _, _, w in res[1:]]))
        line["steady_state"] = {"value": n_pairs / warm, "wall_s": warm, "cold_cli_wall_s": cold_cli([*flags, str(fa)],
                                                                                                     td)}
        if not a.no_cpu:
            exe, kind = ref_exe()
            if exe:
                keep = []
                th = thread_counts()[0]
                w = run_reference(exe, [[*flags, str(fa)]], th, keep)
                line["cpu_baseline"] = {"value": n_pairs / w, "unit": "pairs/s", "cores": th, "kind": kind,
                                        "variant": REF_VARIANT, "host": host_info(),
                                        "sample": "whole workload, one full command run (read + compute + write); "
                                                  "no 1-thread run (about 16 x longer)", "wall_s": w}
                line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
                line["steady_state_over_cpu"] = line["steady_state"]["value"] / line["cpu_baseline"]["value"]
                ok = all(rc == 0 and (o, so) == keep[0] for rc, o, so, _ in res)
                line["parity"] = {"sha256_ok": ok, "output_sha256": res[0][1], "checked": "distance file + stdout of "
                                  "every GPU command vs the reference's"}
    print(json.dumps(line))


def fasta_codes(path):
    """(names, codes) of an aligned FASTA file, codes as ValidChars indices (upper or lower case)."""
    import numpy as np
    lut = np.full(256, 255, dtype=np.uint8)
    for k, ch in enumerate("ACGTURYSWKMBDHVN?-"):
        lut[ord(ch)] = lut[ord(ch.lower())] = k
    names, rows = [], []
    for block in Path(path).read_text().split(">")[1:]:
        lines = block.splitlines()
        names.append(lines[0])
        rows.append(lut[np.frombuffer("".join(lines[1:]).encode(), dtype=np.uint8)])
    return names, np.stack(rows)


def bench_sars2(a, T, gen):
    """The reference's real SARS-CoV-2 dev set (data/real_sars2_dev.fa.xz, kept under tests/golden/real/): 500
    genomes x 29,903 columns, of which 20,673 hold one clean base in every genome and are folded out of the device
    arrays, with the reference's benchmark flags -t 0.0002 -l 20000 -0; the engine resident folded and with every
    column kept, the whole command warm and cold, and the reference beside it."""
    import lzma
    import numpy as np
    src = REPO / "tests" / "golden" / "real" / "real_sars2_dev.fa.xz"
    flags = ["-t", "0.0002", "-l", "20000", "-0"]
    with tempfile.TemporaryDirectory() as td:
        fa = Path(td) / "sars2_dev.fa"
        fa.write_bytes(lzma.decompress(src.read_bytes()))
        _, codes = fasta_codes(fa)
        n, L = codes.shape
        n_pairs = n * (n - 1) // 2
        eng = T.Engine(0)
        kw = dict(mode=T.RESOLVE, threshold=0.0002, min_overlap=20000, resolve_fraction=1.0)
        runs = {}
        for name, nofold in [("folded", "0"), ("all_columns", "1")]:
            os.environ["TN93G_NOFOLD"] = nofold
            try:
                s = eng.upload(codes)
            finally:
                del os.environ["TN93G_NOFOLD"]
            for _ in range(max(a.warmup, 1)):
                r = eng.run(s, **kw)
            steps = max(a.steps, 20)
            pk, t0 = [], time.perf_counter()
            for _ in range(steps):
                r = eng.run(s, **kw)
                pk.append(r["pair_kernel_ms"])
            dt = (time.perf_counter() - t0) / steps
            runs[name] = {"ms_per_step": dt * 1e3, "pairs_per_s": n_pairs / dt, "pair_kernel_ms": float(np.mean(pk)),
                          "folded_columns": int(r["folded_columns"]), "links": int(r["found_links"])}
            s.free()
        eng.close()
        value = runs["folded"]["pairs_per_s"]
        line = {"metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": max(a.steps, 20),
                "warmup": a.warmup, "ms_per_step": runs["folded"]["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "int32+f64",
                "data": "real (the reference's SARS-CoV-2 dev set, 500 GenBank genomes aligned to MN908947)",
                "config": {"workload": "real SARS-CoV-2 dev set, 500 x 29,903, -t 0.0002 -l 20000 -0",
                           "pairs": n_pairs, "parallelism": "1 GPU"},
                "folding": runs}
        cases = []
        for k in range(6):  # the first case also pays the device's initialisation
            d = Path(td) / f"e{k}"
            d.mkdir()
            cases.append((d, ["-q", *flags, "-o", "out.csv", str(fa)]))
        res = tn93g_batch(cases, td)
        warm = float(np.median([w for _, _, _, w in res[1:]]))
        line["steady_state"] = {"value": n_pairs / warm, "wall_s": warm,
                                "cold_cli_wall_s": min(cold_cli([*flags, str(fa)], td) for _ in range(3))}
        if not a.no_cpu:
            base, keep = cpu_baseline([[*flags, str(fa)]], n_pairs, "the whole workload (one command)")
            if base:
                line["cpu_baseline"] = base
                line["gpu_over_cpu"] = value / base["value"]
                line["steady_state_over_cpu"] = line["steady_state"]["value"] / base["value"]
                line["cold_cli_over_cpu"] = base["wall_s"] / line["steady_state"]["cold_cli_wall_s"]
                ok = all(rc == 0 and (o, so) == keep[0] for rc, o, so, _ in res)
                line["parity"] = {"sha256_ok": ok, "output_sha256": res[0][1],
                                  "checked": "distance file + stdout of every GPU command vs the reference's"}
    print(json.dumps(line))


def bench_c3(a, T, gen):
    import numpy as np
    jobs_codes = gen.c3()
    eng = T.Engine(0)
    modes = {"resolve": T.RESOLVE, "average": T.AVERAGE, "skip": T.SKIP, "gapmm": T.GAPMM}
    sets, jobs = [], []
    for codes, mode in jobs_codes:
        st = eng.upload(codes)
        sets.append(st)
        jobs.append((st, dict(mode=modes[mode], threshold=0.015, min_overlap=min(500, codes.shape[1] // 2),
                              resolve_fraction=0.05 if mode == "resolve" else 1.0, backend=T.BACKEND_AUTO,
                              flags=T.FLAG_AUTOTUNE)))
    n_pairs = sum(c.shape[0] * (c.shape[0] - 1) // 2 for c, _ in jobs_codes)
    # the -s bipartite block, 2,000 queries x 40,000 references, as one more job of the same batch
    bc = gen.c3_bipartite()
    bs = eng.upload(bc)
    bkw = dict(mode=T.RESOLVE, kind=T.CROSS, n1=2000, threshold=0.015, min_overlap=500, resolve_fraction=0.05,
               backend=T.BACKEND_AUTO, flags=T.FLAG_AUTOTUNE)
    bip_pairs = 2000 * 40000
    jobs.append((bs, bkw))
    for _ in range(max(a.warmup, 1)):  # the first batch also tunes every job shape
        eng.run_batch(jobs)
    # plain path: the same jobs (the -s block included) one tn93g_run each
    t0 = time.perf_counter()
    pk = 0.0
    for _ in range(a.steps):
        for st, kw in jobs:
            pk += eng.run(st, **kw)["pair_kernel_ms"]
    t_plain = (time.perf_counter() - t0) / a.steps
    # batched: one tn93g_run_batch (one graph launch) per step
    t0 = time.perf_counter()
    bpk = 0.0
    for _ in range(a.steps):
        res = eng.run_batch(jobs)
        bpk += max(r["pair_kernel_ms"] for r in res)  # the merged pair launches (device clock)
    t_batch = (time.perf_counter() - t0) / a.steps
    bpk /= a.steps
    merged = [r["pair_kernel_ms"] > 0 for r in res]
    value = (n_pairs + bip_pairs) / t_batch
    rb = res[-1]
    # the -s block alone, for reference
    eng.run(bs, **bkw)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        eng.run(bs, **bkw)
    t_bip = (time.perf_counter() - t0) / a.steps
    import collections
    shapes = collections.Counter(("mfma" if r["backend_used"] == T.BACKEND_MFMA else "popc") + f"x{r['tile_rows_used']}"
                                 for r in res)
    # roofline of the batch's dominant kernel, the pair launches shared by the jobs of one kernel variant and tile
    # height (device clock around all of them): 22 int32 ops per word pair (count_word) over the merged jobs' pairs
    Lw = [c.shape[0] * (c.shape[0] - 1) // 2 * ((c.shape[1] + 31) // 32) for c, _ in jobs_codes]
    Lw.append(bip_pairs * ((bc.shape[1] + 31) // 32))
    Lw_m = sum(w for w, m in zip(Lw, merged) if m)
    achieved = 22 * Lw_m / (bpk * 1e-3) / 1e12 if bpk > 0 else 0.0
    line = {"metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32+f64", "data": "synthetic (seeded generators, swifttn93_amd/gen.py c3, c3_bipartite)",
            "config": {"workload": "64 alignments, 50-3,000 sequences x 300-10,000 columns, -a resolve/average/skip/"
                                   "gapmm cycling, -t 0.015, and the -s block of 2,000 x 40,000 at 1,300 columns, "
                                   "all in one batch (one graph launch; backends and tile heights autotuned per job "
                                   "shape)", "pairs": n_pairs + bip_pairs, "parallelism": "1 GPU"},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_INT32_TOPS, "unit": "TOPS",
                         "frac": achieved / VALU_INT32_TOPS, "traffic": None,
                         "kernel": f"k_pairs_multi (launches over the tiles of {sum(merged)} of the "
                                   f"{len(merged)} jobs, one per kernel variant and tile height)", "kernel_ms": bpk,
                         "ops_per_word_pair": 22},
            "tuned_shapes": dict(shapes),
            "plain_pair_kernel_ms": pk / a.steps,
            "plain_pairs_per_s": (n_pairs + bip_pairs) / t_plain, "plain_ms_per_step": t_plain * 1e3,
            "batch_over_plain": t_plain / t_batch,
            "bipartite": {"pairs": bip_pairs, "pairs_per_s": bip_pairs / t_bip, "ms": t_bip * 1e3,
                          "links": int(rb["found_links"])}}
    for st in sets:
        st.free()
    bs.free()
    eng.close()
    n_pairs_jobs = n_pairs
    with tempfile.TemporaryDirectory() as td:
        argvs, cases = [], []
        for k, (codes, mode) in enumerate(jobs_codes):
            fa = Path(td) / f"j{k}.fa"
            gen.write_fasta(fa, codes)
            argv = ["-t", "0.015", "-a", mode, "-l", str(min(500, codes.shape[1] // 2))] + \
                (["-g", "0.05"] if mode == "resolve" else []) + [str(fa)]
            argvs.append(argv)
            d = Path(td) / f"g{k}"
            d.mkdir()
            cases.append((d, ["-q", *argv[:-1], "-o", "out.csv", argv[-1]]))
        qa, ra = Path(td) / "q.fa", Path(td) / "r.fa"
        gen.write_fasta(qa, bc[:2000], gen.names_for(2000, "q"))
        gen.write_fasta(ra, bc[2000:], gen.names_for(40000, "r"))
        bip_argv = ["-t", "0.015", "-a", "resolve", "-g", "0.05", "-l", "500", "-s", str(ra), str(qa)]
        d = Path(td) / "gbip"
        d.mkdir()
        cases.append((d, ["-q", *bip_argv[:-1], "-o", "out.csv", bip_argv[-1]]))
        # the whole batch as tn93 commands: every job read, packed, computed in one tn93g_run_batch, certified and
        # written by one tn93g --jobs process (the first of two runs pays the device's initialisation)
        tn93g_jobs(cases, td)
        got, jobs_wall = tn93g_jobs(cases, td)
        line["steady_state"] = {"value": (n_pairs_jobs + bip_pairs) / jobs_wall, "wall_s": jobs_wall,
                                "what": "tn93g --jobs over the 64 alignments and the -s block: read, pack, one "
                                        "tn93g_run_batch, certify, format, write (one process, warm device)"}
        if not a.no_cpu:
            exe, kind = ref_exe()
            if exe:
                th = thread_counts()[0]
                keep = []
                w = run_reference(exe, argvs, th, keep)
                wb = run_reference(exe, [bip_argv], th, keep)
                line["cpu_baseline"] = {"value": (n_pairs + bip_pairs) / (w + wb), "unit": "pairs/s", "cores": th,
                                        "kind": kind, "variant": REF_VARIANT, "host": host_info(),
                                        "sample": "the 64 alignments and the -s block, one full command run each "
                                                  "(read + compute + write)", "wall_s": w + wb,
                                        "jobs": {"value": n_pairs / w, "wall_s": w},
                                        "bipartite": {"value": bip_pairs / wb, "wall_s": wb}}
                line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
                line["steady_state_over_cpu"] = line["steady_state"]["value"] / line["cpu_baseline"]["value"]
                bad = [k for k, ((rc, o, so), want) in enumerate(zip(got, keep)) if rc != 0 or (o, so) != want]
                line["parity"] = {"sha256_ok": not bad, "mismatched_jobs": bad,
                                  "checked": "distance file + stdout of each of the 64 jobs and the -s block (tn93g "
                                             "--jobs) vs the reference's"}
    print(json.dumps(line))


def bench_c4(a, T, gen):
    import numpy as np
    from swifttn93_amd import serve
    n_panel = 50000
    panel = gen.c1(n=n_panel, clusters=n_panel * 200 // 6700)
    print(f"c4: generating a {a.serve_seconds:g} s query stream", file=sys.stderr, flush=True)
    stream = serve.query_stream(panel, seed=8, rate=20.0, seconds=a.serve_seconds)
    nq = sum(c.shape[0] for _, c in stream)
    srv = serve.Server(panel, capacity=n_panel + nq, workers=4, hbm_budget=32 << 30, merge_rows=a.merge_rows)
    # warm the worker contexts (their run buffers) with the last panel rows as if they had just arrived
    for e in srv.free:
        e.run(srv.set, kind=T.APPEND, n1=n_panel - 512, n2=n_panel, **srv.run_kw)
    print(f"c4: serving {len(stream)} batches, {nq} queries", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    reqs = serve.serve(srv, stream)
    wall = time.perf_counter() - t0
    errors = [r.error for r in reqs if r.error]
    lat = np.array([r.latency for r in reqs]) * 1e3
    pairs = sum((r.n2 * (r.n2 - 1) - r.n1 * (r.n1 - 1)) // 2 for r in reqs)
    runs = {}
    for r in reqs:
        runs.setdefault((r.started, r.done), []).append(r)
    busy = sum(d - s for s, d in runs)
    srv.close()
    line = {"metric": METRIC, "value": pairs / wall, "unit": "pairs/s", "n_gpus": 1, "steps": len(reqs), "warmup": 4,
            "ms_per_step": float(np.mean(lat)), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32+f64", "data": "synthetic (seeded c1 generator panel, Poisson query stream, "
                                          "swifttn93_amd/serve.py)",
            "config": {"workload": f"serving: 50,000 x 1,300 device-resident panel, query batches of 1-512 at "
                                   f"20/s for {a.serve_seconds:g} s, each against the panel and all earlier queries, "
                                   "-t 0.015 -a resolve -g 0.05 -l 500, 4 workers, 32 GiB budget, waiting batches "
                                   f"merged up to {a.merge_rows} queries per run", "pairs": pairs,
                       "parallelism": "1 GPU"},
            "roofline": None,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "batches": len(reqs), "queries": nq, "errors": len(errors), "runs": len(runs),
            "merged_batches": sum(1 for r in reqs if r.group > 1),
            "pairs_per_busy_s": pairs / busy, "offered_pairs_per_s": pairs / a.serve_seconds}
    if not a.no_cpu:
        exe, kind = ref_exe()
        if exe:
            line["parity"], line["cpu_baseline"] = c4_parity(exe, kind, panel, stream, reqs, n_panel)
    print(json.dumps(line))


def c4_parity(exe, kind, panel, stream, reqs, n_panel, checks=8):
    """Served links against the reference on `checks` batches spread over the stream: for batch k, the reference's
    -s run of everything received before it (panel and earlier batches, in arrival order, as the first file) against
    the batch (second file), plus its triangle over the batch itself, gives exactly the pairs (i, j), i < j, j in the
    batch, each with the earlier sequence first as in the served APPEND run.  Names carry the global index."""
    import numpy as np
    th = thread_counts()[0]
    env = dict(os.environ, OMP_NUM_THREADS=str(th))
    flags = ["-q", "-t", "0.015", "-a", "resolve", "-g", "0.05", "-l", "500", "-f", "csv"]
    picks = sorted(set(int(x) for x in np.linspace(0, len(reqs) - 1, min(checks, len(reqs)))))
    bad, nlinks, wall, pairs = [], 0, 0.0, 0
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        from swifttn93_amd import gen
        gen.write_fasta(td / "panel.fa", panel, [f"s{i}" for i in range(n_panel)])
        files = [td / "panel.fa"]
        for k, (t, codes) in enumerate(stream):
            r = reqs[k]
            f = td / f"b{k}.fa"
            gen.write_fasta(f, codes, [f"s{i}" for i in range(r.n1, r.n2)])
            files.append(f)
        for k in picks:
            r = reqs[k]
            earlier = td / "earlier.fa"
            with open(earlier, "wb") as out:
                for f in files[:k + 1]:  # the panel, then batches 0 .. k-1
                    out.write(f.read_bytes())
            want = set()
            t0 = time.perf_counter()
            for argv in (["-s", str(files[k + 1]), str(earlier)], [str(files[k + 1])]):
                subprocess.run([str(exe), *flags, "-o", str(td / "ref.csv"), *argv], check=True, capture_output=True,
                               timeout=1200, env=env)
                with open(td / "ref.csv") as fh:
                    next(fh)
                    for ln in fh:
                        x, y, d = ln.rstrip("\n").split(",")
                        want.add((int(x[1:]), int(y[1:]), d))
            wall += time.perf_counter() - t0
            q = r.n2 - r.n1
            pairs += r.n1 * q + q * (q - 1) // 2
            got = set()
            if r.result is not None:
                for i, j, d in zip(r.result["link_row"], r.result["link_col"], r.result["link_dist"]):
                    got.add((int(i), int(j), "%g" % d))
            nlinks += len(want)
            if got != want:
                bad.append(k)
    parity = {"sha256_ok": None, "links_identical": not bad, "batches_checked": len(picks), "mismatched": bad,
              "reference_links": nlinks,
              "checked": "served links (row, column, %g distance) of batches spread over the stream vs the reference "
                         "solver: -s (everything received before the batch) x (the batch), plus the batch's own "
                         "triangle"}
    baseline = {"value": pairs / wall if wall else None, "unit": "pairs/s", "cores": th, "kind": kind,
                "variant": REF_VARIANT, "host": host_info(), "wall_s": wall,
                "sample": f"the reference's runs of the parity check ({len(picks)} batches, {pairs} pairs: read + "
                          "compute + write)"}
    return parity, baseline


METRIC = "byte-exact pairs/sec at 1 GPU vs host cores; roofline %; serving p50/p99"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--backend", default="auto", choices=["auto", "popc", "mfma"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (torch), the reference for the HBM byte counters")
    ap.add_argument("--config", default="c1", choices=["c1", "c2", "c3", "c4", "sars2"])
    ap.add_argument("--mode", default="resolve", choices=["resolve", "average", "skip", "gapmm"],
                    help="c1: tn93's -a mode (resolve runs with -g 0.05)")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="c4: length of the query stream")
    ap.add_argument("--tile-rows", type=int, default=0, choices=[0, 8, 16, 32],
                    help="c1: rows of the pair kernel's tiles (0: the default, or the autotuner's with --autotune)")
    ap.add_argument("--autotune", action="store_true", help="c1: let the autotuner pick the backend and tile rows")
    ap.add_argument("--merge-rows", type=int, default=2048,
                    help="c4: waiting batches merged into one run up to this many queries (0: no merging)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {a.gpus} is not supported (sharding across GPUs is "
              f"a design only in this build)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import swifttn93_amd as T
    from swifttn93_amd import gen
    if a.copy_calibration:
        import torch
        x = torch.empty(a.copy_calibration << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        del x, y
        print(f"copy calibration: {a.copy_calibration} MiB read + {a.copy_calibration} MiB written per copy",
              file=sys.stderr)
    if a.config == "c2":
        return bench_c2(a, T, gen)
    if a.config == "c3":
        return bench_c3(a, T, gen)
    if a.config == "c4":
        return bench_c4(a, T, gen)
    if a.config == "sars2":
        return bench_sars2(a, T, gen)

    codes = gen.c1()
    n, L = codes.shape
    n_pairs = n * (n - 1) // 2
    backend = {"auto": T.BACKEND_AUTO, "popc": T.BACKEND_POPC, "mfma": T.BACKEND_MFMA}[a.backend]
    eng = T.Engine(0)
    s = eng.upload(codes)
    tmode = {"resolve": T.RESOLVE, "average": T.AVERAGE, "skip": T.SKIP, "gapmm": T.GAPMM}[a.mode]
    kw = dict(mode=tmode, threshold=0.015, min_overlap=500, resolve_fraction=0.05 if a.mode == "resolve" else 1.0,
              backend=backend, tile_rows=a.tile_rows, flags=T.FLAG_AUTOTUNE if a.autotune else 0)
    for _ in range(a.warmup):
        r = eng.run(s, **kw)
    t0 = time.perf_counter()
    pk, gm = [], []
    for _ in range(a.steps):
        r = eng.run(s, **kw)
        pk.append(r["pair_kernel_ms"])
        gm.append(r["gpu_ms"])
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / a.steps
    value = n_pairs * a.steps / dt
    pair_ms = float(np.mean(pk))
    words = (L + 31) // 32
    popc = r["backend_used"] == T.BACKEND_POPC
    rows = r["tile_rows_used"]
    # popcount backend: 12 logic + 10 popcount-accumulate int32 operations per 32-bit word, against the VALU peak;
    # MFMA backend: the 4 x 4 one-hot product, 16 int8 multiply-adds = 32 operations per site, against the int8 peak
    ops_per_pair = 22 * words if popc else 32 * L
    peak = VALU_INT32_TOPS if popc else MFMA_INT8_TOPS
    achieved = ops_per_pair * n_pairs / (pair_ms * 1e-3) / 1e12
    # template arguments: backend (1 popcount, 2 MFMA), fast path, dense-N planes, marginal form (not in SKIP mode,
    # whose ambiguity sites are not visited)
    kname = ("k_pairs<1, true, false, %s, %d>" % ("false" if a.mode == "skip" else "true", rows // 4)) if popc \
        else "k_pairs<2, true, false, false, 4>"
    traffic = profiled_traffic(kname) if a.mode == "resolve" and popc else None  # profiled: the resolve workload
    roofline = {"bound": "valu" if popc else "mfma", "achieved": achieved, "peak": peak, "unit": "TOPS",
                "frac": achieved / peak, "traffic": traffic,
                "kernel": f"{kname} ({'popcount' if popc else 'int8 MFMA'} counts + events + routing, tiles of "
                          f"{rows} x 64)",
                "kernel_ms": pair_ms, "ops_per_pair": ops_per_pair,
                "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/c1_popc_hbm_bytes.md",
                "hbm_frac": None if traffic is None else traffic / (pair_ms * 1e-3) / HBM_BYTES_S}
    line = {
        "metric": METRIC,
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32+f64",
        "data": "synthetic (seeded HIV-pol-shaped generator, swifttn93_amd/gen.py)",
        "config": {"workload": "HIV-pol-shaped single set, 6,700 x 1,300, " + " ".join(mode_flags(a.mode)[:-2]),
                   "pairs": n_pairs, "parallelism": "1 GPU"},
        "value_is": "the engine with the packed set resident in HBM: tn93g_run (counts, events, replays, tail, "
                    "histogram, row sums, links copied to the host); steady_state adds the reader, packing, "
                    "certification, formatting and the file",
        "roofline": roofline,
        "gpu_ms_per_step": float(np.mean(gm)),
        "links": int(r["found_links"]), "replayed_pairs": int(r["n_replayed"]),
        "enclosed_pairs": int(r["n_uncertain"]),  # distances carried as an enclosure and certified on output
    }
    s.free()
    eng.close()
    modes = ["resolve", "average", "skip"] if a.mode == "resolve" else [a.mode]
    with tempfile.TemporaryDirectory() as td:
        fa = Path(td) / "c1.fa"
        gen.write_fasta(fa, codes)
        # the whole command, warm: the measured mode 6 times in one tn93g process (the first also pays the
        # device's initialisation and is left out), then each parity mode once
        cases = []
        for k in range(6):
            d = Path(td) / f"w{k}"
            d.mkdir()
            cases.append((d, ["-q", *mode_flags(a.mode), "-o", "out.csv", str(fa)]))
        for m in modes:
            d = Path(td) / f"p_{m}"
            d.mkdir()
            cases.append((d, ["-q", *mode_flags(m), "-o", "out.csv", str(fa)]))
        res = tn93g_batch(cases, td)
        warm = float(np.median([w for _, _, _, w in res[1:6]]))
        cold = cold_cli([*mode_flags(a.mode), str(fa)], td)
        line["steady_state"] = {"value": n_pairs / warm, "wall_s": warm, "cold_cli_wall_s": cold,
                                "cold_cli_value": n_pairs / cold,
                                "what": "tn93 command in a warm tn93g process: read FASTA, pack + upload, run, "
                                        "certify, format, write the distance file (median of 5)"}
        if not a.no_cpu:
            cb, keep = cpu_baseline([[*mode_flags(a.mode), str(fa)]], n_pairs,
                                    f"whole workload (6,700 x 1,300, {n_pairs} pairs, -a {a.mode})")
            line["cpu_baseline"] = cb
            if cb:
                line["gpu_over_cpu"] = value / cb["value"]
                line["steady_state_over_cpu"] = line["steady_state"]["value"] / cb["value"]
                line["cold_cli_over_cpu"] = line["steady_state"]["cold_cli_value"] / cb["value"]
                # parity: every mode's GPU command output against the reference's, at full size
                exe, _ = ref_exe()
                refs = {a.mode: keep[0]}
                for m in modes:
                    if m not in refs:
                        kk = []
                        run_reference(exe, [[*mode_flags(m), str(fa)]], cb["cores"], kk)
                        refs[m] = kk[0]
                checks = {m: res[6 + q][0] == 0 and (res[6 + q][1], res[6 + q][2]) == refs[m]
                          for q, m in enumerate(modes)}
                checks["timed_runs"] = all(rc == 0 and (o, so) == refs[a.mode] for rc, o, so, _ in res[:6])
                line["parity"] = {"sha256_ok": all(checks.values()), "by_mode": checks,
                                  "output_sha256": res[6][1],
                                  "checked": "distance file + stdout of the GPU command vs the reference solver's, "
                                             "same input and flags"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
