#!/usr/bin/env python3
"""Headline benchmark: LDE + Blake2s Merkle commit of M31 columns on one MI355X.

Workload (one "step"): 256 M31 columns of log size 20, uniform random, resident in
HBM as evaluations on CanonicCoset(20) (bit-reversed order):
  interpolate (circle IFFT) -> coefficients,
  evaluate on CanonicCoset(21) (blowup 2) with precomputed twiddles -> LDE,
  Blake2s Merkle tree over the 256 LDE columns (2^21 leaves, 22 layers).
Metric: committed LDE elements per second (256 * 2^21 per step).

Prints one JSON line.  Per-kernel durations come from HIP events recorded on the
library's stream around every launch (a separate, non-timed pass).  The CPU
baseline is the oracle (the project's C restatement, OpenMP) on a bounded sample
of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "stwo-gpu_amd"))

N_COLS = 256
LOG_N = 20
LOG_BLOWUP = 1
SEED = 0x5EED0002
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(C, n):
    """Per-launch compulsory bytes (read once + write once) of each kernel of a step."""
    m = 2 * n  # LDE length
    return {
        "fft_inv_contig": 8 * C * n,           # read evaluations, write partial IFFT
        "fft_lde_strided": 4 * C * n + 4 * C * n + 4 * C * m,  # read partial IFFT, write coefficients + partial LDE
        "fft_fwd_contig": 8 * C * m,           # read, write LDE
        "merkle_leaf": 4 * C * m + 32 * m,     # read every LDE value, write 2^21 leaf hashes
    }


def pipeline_bytes(C, n):
    """Whole-step algorithmic bytes: 28 C n + 256 n (interp 8Cn, eval 12Cn, leaves 8Cn + 64n, nodes)."""
    m = 2 * n
    return 8 * C * n + 12 * C * n + 4 * C * m + 32 * m + 32 * (2 * m - 2) + 32 * (m - 1)


def cpu_baseline(sample_cols):
    """Oracle LDE + Merkle on `sample_cols` columns of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import stwo_hip as S

    n = 1 << LOG_N
    evals = [O.fill_uniform_m31(n, SEED, c * n) for c in range(sample_cols)]
    root = S.canonic_half_coset(LOG_N + LOG_BLOWUP)
    tw, itw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    t0 = time.perf_counter()
    coeffs, ldes = O.lde_columns(evals, LOG_BLOWUP, tw, itw)
    layers = O.merkle_commit(ldes)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return sample_cols * (n << LOG_BLOWUP) / dt, dt, threads, ldes[0], layers


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sample-cols", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device copy (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this backend runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import stwo_hip as S

    S.check(S.lib().stwo_hip_init(0))
    C, n = N_COLS, 1 << LOG_N
    m = n << LOG_BLOWUP
    evals = [S.BaseFieldVec.uninitialized(n) for _ in range(C)]
    for c, col in enumerate(evals):
        S.fill_uniform(col, SEED, c * n)
    coeffs = [S.BaseFieldVec.uninitialized(n) for _ in range(C)]
    ldes = [S.BaseFieldVec.uninitialized(m) for _ in range(C)]
    root = S.canonic_half_coset(LOG_N + LOG_BLOWUP)
    tw, itw = S.cached_twiddles(root)
    tree_len = 1 << root.log_size
    tree = S.MerkleTree(ldes, [LOG_N + LOG_BLOWUP] * C)
    S.synchronize()
    if args.calibrate_copy:
        # copy the first 256 evaluation columns' worth of bytes: 1 GiB read + 1 GiB written
        a, b = S.BaseFieldVec.uninitialized(1 << 28), S.BaseFieldVec.uninitialized(1 << 28)
        S.device_copy(b, a)
        S.synchronize()
        del a, b

    def step():
        S.lde_columns_async(evals, coeffs, ldes, [LOG_N] * C, LOG_BLOWUP, tw, itw, tree_len)
        tree.commit_async()

    for _ in range(args.warmup):
        step()
    S.synchronize()
    S.timer_begin()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ms = S.timer_end()
    wall = time.perf_counter() - t0
    S.synchronize()
    ms_per_step = ms / args.steps
    elems = C * m
    value = elems * args.steps / (ms / 1e3)

    # per-kernel durations (separate pass with events around every launch)
    S.kernel_stats_reset()
    S.kernel_timing(True)
    reps = 3
    for _ in range(reps):
        step()
    S.synchronize()
    S.kernel_timing(False)
    names = ["fft_inv_contig", "fft_lde_strided", "fft_fwd_contig", "merkle_leaf", "merkle_subtree"]
    stats = {k: S.kernel_stats(k) for k in names}
    alg = algorithmic_bytes(C, n)
    kernels = {}
    for k, (cnt, tot) in stats.items():
        if cnt:
            avg = tot / cnt
            kernels[k] = {"launches_per_step": cnt / reps, "avg_ms": round(avg, 4)}
            if k in alg:
                kernels[k]["GBps"] = round(alg[k] / (avg * 1e-3) / 1e9, 1)
    dom = max((k for k in alg if k in kernels), key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches_per_step"])
    d_avg = kernels[dom]["avg_ms"]
    achieved = alg[dom] / (d_avg * 1e-3) / 1e9
    pipe = pipeline_bytes(C, n)
    root_hash = tree.root()
    traffic = None
    tf = os.path.join(ROOT, "profiles", "lde_merkle_256x2e20_blowup2_traffic.json")
    if os.path.exists(tf):  # measured HBM bytes per dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled)
        rec = json.load(open(tf))["kernels"].get(dom)
        if rec:
            traffic = rec["read_bytes"] + rec["write_bytes"]

    out = {
        "metric": "LDE+Merkle M31 elems/s",
        "value": value,
        "unit": "elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (M31)",
        "data": "synthetic (uniform M31, seeded counter-based stream)",
        "config": {"workload": "cfg2 LDE+commit: 256 M31 columns log 20 -> LDE log 21 (blowup 2) + Blake2s Merkle",
                   "columns": C, "log_size": LOG_N, "log_blowup": LOG_BLOWUP, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg[dom], "avg_launch_ms": d_avg},
        "pipeline": {"algorithmic_bytes_per_step": pipe,
                     "achieved_GBps": round(pipe / (ms_per_step * 1e-3) / 1e9, 1),
                     "frac_of_hbm_peak": round(pipe / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "root": "".join(f"{w:08x}" for w in root_hash),
        "wall_s_timed_region": round(wall, 4),
    }
    if not args.no_cpu_baseline:
        cv, dt, threads, lde0, layers = cpu_baseline(args.cpu_sample_cols)
        parity = bool((ldes[0].to_numpy() == lde0).all())
        out["cpu_baseline"] = {"value": cv, "unit": "elems/s", "cores": threads, "kind": "port",
                               "sample": f"{args.cpu_sample_cols} of the 256 columns (same seeds), full LDE + "
                                         f"Merkle tree over them; {dt:.2f} s",
                               "lde_column0_matches_gpu": parity}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
