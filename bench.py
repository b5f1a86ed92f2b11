#!/usr/bin/env python3
"""Benchmark: byte-exact all-pairs TN93 pairs/s on one MI355X, with the reference CPU path timed beside it.

Workload (one step = one pass of the hot path over the set): the HIV-pol-shaped set of 6,700 sequences x 1,300
columns (swifttn93_amd/gen.py c1), options -t 0.015 -a resolve -g 0.05 -l 500, 22,441,650 pairs.  The packed set is
resident in HBM before the timed region; a step is one tn93g_run: clean/clean counts, ambiguity events,
site-order replays, FP64 tail, histogram, per-row sums in tn93's order and the link list copied back to the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--backend popc|mfma] [--no-cpu] [--config c1|c2|c3|c4]
                    [--mode resolve|average|skip|gapmm]

--mode runs the same 6,700 x 1,300 set with another -a mode (-g 0.05 only with resolve; the default line is
resolve).

Other workloads (--config): c2, the SARS-CoV-2-shaped long alignment (20,000 x 29,903, N-masked blocks,
-t 0.0005 -a resolve -g 0.05 -l 25000); c3, 64 uneven alignments with mixed -a modes run as one batch (one graph
launch) beside the same jobs run one by one, and the 2,000 x 40,000 -s bipartite block; c4, serving: a 50,000 x
1,300 device-resident panel answering a Poisson stream of query batches (p50/p99 latency, pairs/s).

Prints one JSON line.  `roofline` is for the pair kernel (HIP events on the engine's stream, averaged over the
timed steps); `cpu_baseline` times the reference solver (oracle/_ref/swifttn93_avx512, built from the reference
sources by oracle/Makefile) on the same input on this host's cores, or the oracle when that build is absent.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

def mode_flags(mode):
    """tn93's flags for the C1 workload in -a mode (-g 0.05 with resolve, as the headline line is quoted)."""
    return ["-t", "0.015", "-a", mode] + (["-g", "0.05"] if mode == "resolve" else []) + ["-l", "500", "-f", "csv"]
VALU_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 39.3: 256 CUs x 64 lanes x 2.4 GHz
MFMA_INT8_TOPS = 5000.0
HBM_BYTES_S = 8.0e12


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


def cpu_baseline(fa, n_pairs, mode):
    ref = REPO / "oracle" / "_ref" / "swifttn93_avx512"
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    if ref.exists():
        kind, cmd = "reference", [str(ref), "-q"]
    else:
        kind, cmd = "port", [str(REPO / "oracle" / "_build" / "tn93_oracle"), "-q"]
        if not Path(cmd[0]).exists():
            return None
    best = None
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "out.csv"
        for _ in range(3):
            t0 = time.perf_counter()
            subprocess.run(cmd + ["-o", str(out)] + mode_flags(mode) + [str(fa)], check=True, env=env,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
            out.unlink()
    return {"value": n_pairs / best, "unit": "pairs/s", "cores": threads, "kind": kind,
            "sample": f"whole workload (6,700 x 1,300, {n_pairs} pairs), full command wall incl. read and write, "
                      f"best of 3, OMP_NUM_THREADS={threads}", "wall_s": best}


def run_reference(exe, args_list, threads):
    """Wall time of the reference solver over a list of (argv, fasta) runs, output deleted outside the timer."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    tot = 0.0
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "out.csv"
        for argv in args_list:
            t0 = time.perf_counter()
            subprocess.run([str(exe), "-q", "-o", str(out)] + argv, check=True, env=env, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, timeout=1200)
            tot += time.perf_counter() - t0
            out.unlink()
    return tot


def ref_exe():
    ref = REPO / "oracle" / "_ref" / "swifttn93_avx512"
    if ref.exists():
        return ref, "reference"
    port = REPO / "oracle" / "_build" / "tn93_oracle"
    return (port, "port") if port.exists() else (None, None)


def roofline_for(n_pairs, L, pair_ms):
    words = (L + 31) // 32
    ops = 22 * words
    achieved = ops * n_pairs / (pair_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": VALU_INT32_TOPS, "unit": "TOPS",
            "frac": achieved / VALU_INT32_TOPS, "traffic": None, "kernel": "k_pairs<1, true, true, true>", "kernel_ms": pair_ms,
            "ops_per_pair": ops}


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
    if not a.no_cpu:
        exe, kind = ref_exe()
        if exe:
            threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
            with tempfile.TemporaryDirectory() as td:
                fa = Path(td) / "c2.fa"
                gen.write_fasta(fa, codes)
                w = run_reference(exe, [["-t", "0.0005", "-a", "resolve", "-g", "0.05", "-l", "25000", str(fa)]],
                                  threads)
            line["cpu_baseline"] = {"value": n_pairs / w, "unit": "pairs/s", "cores": threads, "kind": kind,
                                    "sample": "whole workload, one full command run (read + compute + write)",
                                    "wall_s": w}
            line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
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
    for _ in range(max(a.warmup, 1)):
        eng.run_batch(jobs)
    # plain path: the same jobs one tn93g_run each
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
    value = n_pairs / t_batch
    # the -s bipartite block
    bc = gen.c3_bipartite()
    bs = eng.upload(bc)
    bkw = dict(mode=T.RESOLVE, kind=T.CROSS, n1=2000, threshold=0.015, min_overlap=500, resolve_fraction=0.05,
               backend=T.BACKEND_AUTO, flags=T.FLAG_AUTOTUNE)
    eng.run(bs, **bkw)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        rb = eng.run(bs, **bkw)
    t_bip = (time.perf_counter() - t0) / a.steps
    bip_pairs = 2000 * 40000
    # roofline of the batch's dominant kernel, the pair launch shared by the jobs of one kernel variant: 22 int32 ops
    # per word pair (count_word) over the merged jobs' pairs
    Lw = [c.shape[0] * (c.shape[0] - 1) // 2 * ((c.shape[1] + 31) // 32) for c, _ in jobs_codes]
    Lw_m = sum(w for w, m in zip(Lw, merged) if m)
    achieved = 22 * Lw_m / (bpk * 1e-3) / 1e12 if bpk > 0 else 0.0
    line = {"metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32+f64", "data": "synthetic (seeded generators, swifttn93_amd/gen.py c3, c3_bipartite)",
            "config": {"workload": "64 alignments, 50-3,000 sequences x 300-10,000 columns, -a resolve/average/skip/"
                                   "gapmm cycling, -t 0.015, one batch (one graph launch)", "pairs": n_pairs,
                       "parallelism": "1 GPU"},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_INT32_TOPS, "unit": "TOPS",
                         "frac": achieved / VALU_INT32_TOPS, "traffic": None,
                         "kernel": f"k_pairs_multi (one launch over the tiles of {sum(merged)} of the "
                                   f"{len(merged)} jobs)", "kernel_ms": bpk,
                         "ops_per_word_pair": 22},
            "plain_pair_kernel_ms": pk / a.steps,
            "plain_pairs_per_s": n_pairs / t_plain, "plain_ms_per_step": t_plain * 1e3,
            "batch_over_plain": t_plain / t_batch,
            "bipartite": {"pairs": bip_pairs, "pairs_per_s": bip_pairs / t_bip, "ms": t_bip * 1e3,
                          "links": int(rb["found_links"])}}
    for st in sets:
        st.free()
    bs.free()
    eng.close()
    if not a.no_cpu:
        exe, kind = ref_exe()
        if exe:
            threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
            with tempfile.TemporaryDirectory() as td:
                argvs = []
                for k, (codes, mode) in enumerate(jobs_codes):
                    fa = Path(td) / f"j{k}.fa"
                    gen.write_fasta(fa, codes)
                    argvs.append(["-t", "0.015", "-a", mode, "-l", str(min(500, codes.shape[1] // 2))]
                                 + (["-g", "0.05"] if mode == "resolve" else []) + [str(fa)])
                w = run_reference(exe, argvs, threads)
            line["cpu_baseline"] = {"value": n_pairs / w, "unit": "pairs/s", "cores": threads, "kind": kind,
                                    "sample": "the 64 alignments, one full command run each (read + compute + write)",
                                    "wall_s": w}
            line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line))


def bench_c4(a, T, gen):
    import numpy as np
    from swifttn93_amd import serve
    n_panel = 50000
    panel = gen.c1(n=n_panel, clusters=n_panel * 200 // 6700)
    stream = serve.query_stream(panel, seed=8, rate=20.0, seconds=a.serve_seconds)
    nq = sum(c.shape[0] for _, c in stream)
    srv = serve.Server(panel, capacity=n_panel + nq, workers=4, hbm_budget=32 << 30)
    # warm the worker contexts (their run buffers) with the last panel rows as if they had just arrived
    for e in srv.free:
        e.run(srv.set, kind=T.APPEND, n1=n_panel - 512, n2=n_panel, **srv.run_kw)
    t0 = time.perf_counter()
    reqs = serve.serve(srv, stream)
    wall = time.perf_counter() - t0
    errors = [r.error for r in reqs if r.error]
    lat = np.array([r.latency for r in reqs]) * 1e3
    pairs = sum((r.n2 * (r.n2 - 1) - r.n1 * (r.n1 - 1)) // 2 for r in reqs)
    busy = sum(r.done - r.started for r in reqs)
    srv.close()
    line = {"metric": METRIC, "value": pairs / wall, "unit": "pairs/s", "n_gpus": 1, "steps": len(reqs), "warmup": 4,
            "ms_per_step": float(np.mean(lat)), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32+f64", "data": "synthetic (seeded c1 generator panel, Poisson query stream, "
                                          "swifttn93_amd/serve.py)",
            "config": {"workload": f"serving: 50,000 x 1,300 device-resident panel, query batches of 1-512 at "
                                   f"20/s for {a.serve_seconds:g} s, each against the panel and all earlier queries, "
                                   "-t 0.015 -a resolve -g 0.05 -l 500, 4 workers, 32 GiB budget",
                       "pairs": pairs, "parallelism": "1 GPU"},
            "roofline": None,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "batches": len(reqs), "queries": nq, "errors": len(errors),
            "pairs_per_busy_s": pairs / busy, "offered_pairs_per_s": pairs / a.serve_seconds}
    print(json.dumps(line))


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
    ap.add_argument("--config", default="c1", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--mode", default="resolve", choices=["resolve", "average", "skip", "gapmm"],
                    help="c1: tn93's -a mode (resolve runs with -g 0.05)")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="c4: length of the query stream")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {a.gpus} is not supported (sharding across GPUs is "
              f"a design only in this build)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import swifttn93_amd as T
    from swifttn93_amd import gen
    if a.config == "c2":
        return bench_c2(a, T, gen)
    if a.config == "c3":
        return bench_c3(a, T, gen)
    if a.config == "c4":
        return bench_c4(a, T, gen)

    if a.copy_calibration:
        import torch
        x = torch.empty(a.copy_calibration << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(f"copy calibration: {a.copy_calibration} MiB read + {a.copy_calibration} MiB written per copy",
              file=sys.stderr)
    codes = gen.c1()
    n, L = codes.shape
    n_pairs = n * (n - 1) // 2
    backend = {"auto": T.BACKEND_AUTO, "popc": T.BACKEND_POPC, "mfma": T.BACKEND_MFMA}[a.backend]
    eng = T.Engine(0)
    s = eng.upload(codes)
    tmode = {"resolve": T.RESOLVE, "average": T.AVERAGE, "skip": T.SKIP, "gapmm": T.GAPMM}[a.mode]
    kw = dict(mode=tmode, threshold=0.015, min_overlap=500, resolve_fraction=0.05 if a.mode == "resolve" else 1.0,
              backend=backend)
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
    popc = backend in (T.BACKEND_AUTO, T.BACKEND_POPC)
    # popcount backend: 12 logic + 10 popcount-accumulate int32 operations per 32-bit word, against the VALU peak;
    # MFMA backend: the 4 x 4 one-hot product, 16 int8 multiply-adds = 32 operations per site, against the int8 peak
    ops_per_pair = 22 * words if popc else 32 * L
    peak = VALU_INT32_TOPS if popc else MFMA_INT8_TOPS
    achieved = ops_per_pair * n_pairs / (pair_ms * 1e-3) / 1e12
    # template arguments: backend (1 popcount, 2 MFMA), fast path, dense-N planes, marginal form (not in SKIP mode,
    # whose ambiguity sites are not visited)
    kname = ("k_pairs<1, true, false, %s>" % ("false" if a.mode == "skip" else "true")) if popc \
        else "k_pairs<2, true, false, false>"
    traffic = profiled_traffic(kname) if a.mode == "resolve" else None  # the profile is of the resolve workload
    roofline = {"bound": "valu" if popc else "mfma", "achieved": achieved, "peak": peak, "unit": "TOPS",
                "frac": achieved / peak, "traffic": traffic,
                "kernel": f"{kname} ({'popcount' if popc else 'int8 MFMA'} counts + events + routing)",
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
        "roofline": roofline,
        "gpu_ms_per_step": float(np.mean(gm)),
        "links": int(r["found_links"]), "replayed_pairs": int(r["n_replayed"]),
        "enclosed_pairs": int(r["n_uncertain"]),  # distances carried as an enclosure and certified on output
    }
    s.free()
    eng.close()
    if not a.no_cpu:
        with tempfile.TemporaryDirectory() as td:
            fa = Path(td) / "c1.fa"
            gen.write_fasta(fa, codes)
            cb = cpu_baseline(fa, n_pairs, a.mode)
        line["cpu_baseline"] = cb
        if cb:
            line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
