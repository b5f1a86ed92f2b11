#!/usr/bin/env python3
"""Benchmark: byte-exact all-pairs TN93 pairs/s on one MI355X, with the reference CPU path timed beside it.

Workload (one step = one pass of the hot path over the set): the HIV-pol-shaped set of 6,700 sequences x 1,300
columns (swifttn93_amd/gen.py c1), options -t 0.015 -a resolve -g 0.05 -l 500, 22,441,650 pairs.  The packed set is
resident in HBM before the timed region; a step is one tn93g_run: clean/clean counts, ambiguity events,
site-order replays, FP64 tail, histogram, per-row sums in tn93's order and the link list copied back to the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--backend popc|mfma] [--no-cpu]

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

MODE_FLAGS = ["-t", "0.015", "-a", "resolve", "-g", "0.05", "-l", "500", "-f", "csv"]
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
        if len(cells) == 5 and cells[0] == kernel:
            return (float(cells[3]) + float(cells[4])) * 1e6
    return None


def cpu_baseline(fa, n_pairs):
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
            subprocess.run(cmd + ["-o", str(out)] + MODE_FLAGS + [str(fa)], check=True, env=env,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
            out.unlink()
    return {"value": n_pairs / best, "unit": "pairs/s", "cores": threads, "kind": kind,
            "sample": f"whole workload (6,700 x 1,300, {n_pairs} pairs), full command wall incl. read and write, "
                      f"best of 3, OMP_NUM_THREADS={threads}", "wall_s": best}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--backend", default="auto", choices=["auto", "popc", "mfma"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (torch), the reference for the HBM byte counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {a.gpus} is not supported (the multi-GPU sharding is "
              f"described in DESIGN.md)", file=sys.stderr)
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
        print(f"copy calibration: {a.copy_calibration} MiB read + {a.copy_calibration} MiB written per copy",
              file=sys.stderr)
    codes = gen.c1()
    n, L = codes.shape
    n_pairs = n * (n - 1) // 2
    backend = {"auto": T.BACKEND_AUTO, "popc": T.BACKEND_POPC, "mfma": T.BACKEND_MFMA}[a.backend]
    eng = T.Engine(0)
    s = eng.upload(codes)
    kw = dict(mode=T.RESOLVE, threshold=0.015, min_overlap=500, resolve_fraction=0.05, backend=backend)
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
    ops_per_pair = 22 * words  # popcount backend: 12 logic + 10 popcount-accumulate per 32-bit word
    achieved = ops_per_pair * n_pairs / (pair_ms * 1e-3) / 1e12
    kname = "k_pairs<1, true>" if backend in (T.BACKEND_AUTO, T.BACKEND_POPC) else "k_pairs<2, true>"
    traffic = profiled_traffic(kname)
    roofline = {"bound": "valu", "achieved": achieved, "peak": VALU_INT32_TOPS, "unit": "TOPS",
                "frac": achieved / VALU_INT32_TOPS, "traffic": traffic,
                "kernel": f"{kname} (popcount counts + events + routing)", "kernel_ms": pair_ms,
                "ops_per_pair": ops_per_pair,
                "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/c1_popc_hbm_bytes.md",
                "hbm_frac": None if traffic is None else traffic / (pair_ms * 1e-3) / HBM_BYTES_S}
    line = {
        "metric": "byte-exact pairs/sec at 1 GPU vs host cores; roofline %; serving p50/p99",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32+f64",
        "data": "synthetic (seeded HIV-pol-shaped generator, swifttn93_amd/gen.py)",
        "config": {"workload": "HIV-pol-shaped single set, 6,700 x 1,300, -t 0.015 -a resolve -g 0.05 -l 500",
                   "pairs": n_pairs, "parallelism": "1 GPU"},
        "roofline": roofline,
        "gpu_ms_per_step": float(np.mean(gm)),
        "links": int(r["found_links"]), "replayed_pairs": int(r["n_replayed"]),
        "uncertain_pairs": int(r["n_uncertain"]),
    }
    s.free()
    eng.close()
    if not a.no_cpu:
        with tempfile.TemporaryDirectory() as td:
            fa = Path(td) / "c1.fa"
            gen.write_fasta(fa, codes)
            cb = cpu_baseline(fa, n_pairs)
        line["cpu_baseline"] = cb
        if cb:
            line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
