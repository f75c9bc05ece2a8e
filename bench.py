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
        "fft_inv_strided": 8 * C * n,          # read partial IFFT, write coefficients
        "fft_lde_strided": 4 * C * n + 4 * C * n + 4 * C * m,  # read partial IFFT, write coefficients + partial LDE
        "fft_fwd_strided": 4 * C * n + 4 * C * m,  # read coefficients (upper half implicit zeros), write partial LDE
        "fft_fwd_contig": 8 * C * m,           # read, write LDE
        "merkle_leaf": 4 * C * m + 32 * m,     # read every LDE value, write 2^21 leaf hashes
    }


# VALU roofline: one 32-bit VALU instruction per lane per cycle on 64 lanes of each of 256
# CUs at 2.4 GHz.  Algorithmic instruction counts: a butterfly is 12 lane-instructions
# (64-bit product by a doubled twiddle, shift, add, and two min-based canonicalisations
# for the sum and the difference); a Blake2s compression is 960 (10 rounds x 8 G x 12:
# add3, xor, alignbit per quarter step).
VALU_PEAK_OPS = 256 * 64 * 2.4e9
OPS_PER_BUTTERFLY = 12
OPS_PER_COMPRESSION = 960


def algorithmic_valu_ops(C, n):
    """Per-launch algorithmic VALU lane-instructions of each kernel of a step (blowup 2)."""
    m = 2 * n
    kc, ks_inv, ks_fwd = 12, (n.bit_length() - 1) - 12, (m.bit_length() - 1) - 13  # forward skips the top layer
    return {
        "fft_inv_contig": OPS_PER_BUTTERFLY * C * (n // 2) * kc,
        "fft_lde_strided": OPS_PER_BUTTERFLY * C * ((n // 2) * ks_inv + 2 * (n // 2) * ks_fwd),
        "fft_fwd_contig": OPS_PER_BUTTERFLY * C * (m // 2) * kc,
        "merkle_leaf": OPS_PER_COMPRESSION * m * ((C + 15) // 16),
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


def bench_lde_commit(args):
    import numpy as np
    import stwo_hip as S

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
    # one plan: pointer tables uploaded once, FFT pass schedule autotuned on these buffers,
    # every step replayed as one hipGraph (kernel-timing passes run it unrecorded)
    plan = S.CommitPlan(evals, coeffs, ldes, [LOG_N] * C, LOG_BLOWUP, tw, itw, tree_len, autotune=True)
    tree = plan
    schedule = S.lde_schedule(LOG_N, LOG_BLOWUP, C)
    S.synchronize()

    def step():
        plan.run(True)

    def step_eager():
        plan.run(False)

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
        step_eager()
    S.synchronize()
    S.kernel_timing(False)
    names = ["fft_inv_contig", "fft_inv_strided", "fft_lde_strided", "fft_fwd_strided", "fft_fwd_contig", "merkle_leaf",
             "merkle_layer", "merkle_subtree"]
    stats = {k: S.kernel_stats(k) for k in names}
    alg = algorithmic_bytes(C, n)
    vops = algorithmic_valu_ops(C, n)
    kernels = {}
    for k, (cnt, tot) in stats.items():
        if cnt:
            avg = tot / cnt
            kernels[k] = {"launches_per_step": cnt / reps, "avg_ms": round(avg, 4)}
            if k in alg:
                kernels[k]["GBps"] = round(alg[k] / (avg * 1e-3) / 1e9, 1)
            if k in vops:
                kernels[k]["valu_frac"] = round(vops[k] / (avg * 1e-3) / VALU_PEAK_OPS, 4)
    dom = max((k for k in alg if k in kernels), key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches_per_step"])
    d_avg = kernels[dom]["avg_ms"]
    achieved = alg[dom] / (d_avg * 1e-3) / 1e9
    pipe = pipeline_bytes(C, n)
    root_hash = tree.root()
    traffic = _traffic("lde_merkle_256x2e20_blowup2", dom)

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
        "config": {"workload": "LDE+commit: 256 M31 columns log 20 -> LDE log 21 (blowup 2) + Blake2s Merkle",
                   "columns": C, "log_size": LOG_N, "log_blowup": LOG_BLOWUP, "parallelism": "single GPU",
                   "fft_schedule": {"contiguous_layers": schedule[0], "fused_strided": schedule[1]},
                   "launch": ("hipGraph replay of the commit plan" if plan.uses_graph() else
                              "stream launches of the commit plan (measured faster than its graph)")},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg[dom], "avg_launch_ms": d_avg},
        # the same kernel against the VALU roofline (the step is VALU-bound: exact M31 arithmetic)
        "valu_roofline": {"bound": "valu", "kernel": dom,
                          "achieved": round(vops[dom] / (d_avg * 1e-3) / 1e12, 2) if dom in vops else None,
                          "peak": round(VALU_PEAK_OPS / 1e12, 2), "unit": "T int32 lane-ops/s",
                          "frac": kernels[dom].get("valu_frac"), "ops_per_launch": vops.get(dom),
                          "ops_per_butterfly": OPS_PER_BUTTERFLY, "ops_per_compression": OPS_PER_COMPRESSION},
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
        if args.cpu_sample_cols == C:  # the whole workload: the roots must agree
            out["cpu_baseline"]["root_matches_gpu"] = bool((layers[0].reshape(-1) == root_hash).all())
    return out


def _timed_steps(S, step, steps, warmup):
    for _ in range(warmup):
        step()
    S.synchronize()
    S.timer_begin()
    for _ in range(steps):
        step()
    ms = S.timer_end()
    return ms / steps


def _kernel_avg(S, step, names, reps=3):
    S.kernel_stats_reset()
    S.kernel_timing(True)
    for _ in range(reps):
        step()
    S.synchronize()
    S.kernel_timing(False)
    res = {}
    for k in names:
        cnt, tot = S.kernel_stats(k)
        if cnt:
            res[k] = {"launches_per_step": cnt / reps, "avg_ms": round(tot / cnt, 4)}
    return res


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def _cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _line(metric, value, unit, ms, args, config, roofline=None, extra=None, higher=True):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32 (M31/QM31)", "data": "synthetic (seeded uniform)", "config": config}
    if roofline:
        out["roofline"] = roofline
    if extra:
        out.update(extra)
    return out


def _traffic(profile, kernel):
    """Measured HBM bytes per dispatch of `kernel` (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration) from profiles/<profile>_traffic.json, or None."""
    tf = os.path.join(ROOT, "profiles", profile + "_traffic.json")
    if not os.path.exists(tf):
        return None
    rec = json.load(open(tf))["kernels"].get(kernel)
    return rec["read_bytes"] + rec["write_bytes"] if rec else None


def _roof(bytes_, avg_ms, kernel, profile=None):
    gbs = bytes_ / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": _traffic(profile, kernel) if profile else None,
            "algorithmic_bytes_per_launch": bytes_, "avg_launch_ms": avg_ms}


def bench_small(args):
    """16 columns of log 16: interpolate, evaluate on log 17, Merkle commit; root checked vs the oracle."""
    import numpy as np
    import stwo_hip as S

    C, L = 16, 16
    n, m = 1 << L, 1 << (L + 1)
    evals = [S.BaseFieldVec.uninitialized(n) for _ in range(C)]
    for c, col in enumerate(evals):
        S.fill_uniform(col, 0x5EED0001, c * n)
    coeffs = [S.BaseFieldVec.uninitialized(n) for _ in range(C)]
    ldes = [S.BaseFieldVec.uninitialized(m) for _ in range(C)]
    root = S.canonic_half_coset(L + 1)
    tw, itw = S.cached_twiddles(root)
    tree = S.MerkleTree(ldes, [L + 1] * C)

    def step_calls():  # one asynchronous call per operation
        S.lde_columns_async(evals, coeffs, ldes, [L] * C, 1, tw, itw, 1 << root.log_size)
        tree.commit_async()
    ms_calls = _timed_steps(S, step_calls, args.steps, args.warmup)
    plan = S.CommitPlan(evals, coeffs, ldes, [L] * C, 1, tw, itw, 1 << root.log_size)
    ms = _timed_steps(S, lambda: plan.run(True), args.steps, args.warmup)
    assert (plan.root() == tree.root()).all()
    extra = {"root": "".join(f"{w:08x}" for w in tree.root()), "launch": "hipGraph replay of the commit plan",
             "ms_per_step_async_calls": round(ms_calls, 4)}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        host = [O.fill_uniform_m31(n, 0x5EED0001, c * n) for c in range(C)]
        otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
        t0 = time.perf_counter()
        _, olde = O.lde_columns(host, 1, otw, oitw)
        oroot = O.merkle_commit(olde)[0]
        dt = time.perf_counter() - t0
        extra["parity_root_matches_oracle"] = bool((oroot == tree.root()).all())
        extra["cpu_baseline"] = {"value": C * m / dt, "unit": "elems/s", "kind": "port", "sample": "whole config",
                                 "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    return _line("LDE+Merkle M31 elems/s (small plain path)", C * m / (ms * 1e-3), "elems/s", ms, args,
                 {"workload": "small LDE+commit: 16 M31 columns log 16 -> LDE log 17 + Blake2s Merkle"}, extra=extra)


def bench_batch_inverse(args):
    """batch_inverse of 2^24 uniform QM31 elements (AoS)."""
    import stwo_hip as S

    n = 1 << 24
    src, dst = S.SecureFieldVec.uninitialized(n), S.SecureFieldVec.uninitialized(n)
    S.check(S.lib().stwo_hip_fill_uniform_m31(src.ptr, 4 * n, 0x5EED0002, 1 << 40))

    def step():
        S.lib().batch_inverse_secure_field(src.ptr, dst.ptr, n)
    ms = _timed_steps(S, step, args.steps, args.warmup)
    ks = _kernel_avg(S, step, ["batch_inverse_qm31"])
    avg = ks["batch_inverse_qm31"]["avg_ms"]
    extra = {"kernels": ks, "note": "ms_per_step is the synchronous ABI call (launch + sync)"}
    if not args.no_cpu_baseline:
        O = _oracle()
        host = O.fill_uniform_m31(4 * n, 0x5EED0002, 1 << 40)
        t0 = time.perf_counter()
        want = O.batch_inverse_qm31(host)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": n / dt, "unit": "elems/s", "kind": "port", "sample": f"whole input; {dt:.2f} s",
                                 "cores": _cores()}
        extra["parity_matches_oracle"] = bool((dst.to_numpy().reshape(-1) == want.reshape(-1)).all())
    return _line("batch_inverse QM31 elems/s", n / (ms * 1e-3), "elems/s", ms, args,
                 {"workload": "batch_inverse: 2^24 QM31 (AoS)"},
                 _roof(32 * n, avg, "batch_inverse_qm31", "batch_inverse_qm31_2e24"), extra)


def bench_fri(args):
    """fold_circle_into_line of a QM31 circle evaluation at log 22, then fold_line down to log 4,
    every folding coefficient seeded up front: the fused chain (each block folds a 2^13-value
    range 13 times in LDS; two launches), and beside it the per-layer path (18 launches)."""
    import numpy as np
    import stwo_hip as S

    L = 22
    rng = np.random.default_rng(0x5EED0002)
    src = S.SecureColumn.uninitialized(1 << L)
    for k in range(4):
        S.fill_uniform(src.columns[k], 0x5EED0002, (2 << 40) + k * (1 << L))
    layers = [S.SecureColumn.zeros(1 << l) for l in range(L - 1, 3, -1)]
    root = S.canonic_half_coset(L)  # the circle domain's half coset: its tree serves every line layer
    tw, itw = S.cached_twiddles(root)
    tree_len = 1 << root.log_size
    alphas = [S.random_qm31(rng) for _ in range(len(layers))]

    def step():
        S.fri_fold_chain_async(src, layers, alphas, itw, tree_len)

    def step_per_layer():
        S.fold_circle_into_line_async(layers[0], src, alphas[0], itw, tree_len - len(layers[0]))
        for i in range(1, len(layers)):
            S.fold_line_async(layers[i - 1], layers[i], alphas[i], itw, tree_len - len(layers[i - 1]))
    ms = _timed_steps(S, step, args.steps, args.warmup)
    ms_layers = _timed_steps(S, step_per_layer, args.steps, args.warmup)
    ks = _kernel_avg(S, step, ["fri_fold_chain", "fri_fold_chain_tail"])
    ks.update(_kernel_avg(S, step_per_layer, ["fold_circle_into_line", "fold_line"]))
    # algorithmic bytes: every layer read once and written once (the circle fold also reads the
    # layer it accumulates into), plus the twiddle words each fold reads
    circle_bytes = 16 * (1 << L) + 2 * 16 * (1 << (L - 1)) + 4 * (1 << (L - 2))
    line = {l: 16 * (1 << l) + 16 * (1 << (l - 1)) + 4 * (1 << (l - 1)) for l in range(L - 1, 4, -1)}
    total = circle_bytes + sum(line.values())
    # the wide launch (2^13-input blocks): the circle fold and the 12 line folds below it (inputs of
    # log 21 .. 10); its later folds read the block's previous outputs from LDS, so count only their
    # writes and twiddles
    wide = circle_bytes + sum(16 * (1 << (l - 1)) + 4 * (1 << (l - 1)) for l in range(L - 1, L - 13, -1))
    extra = {"kernels": ks, "algorithmic_bytes_per_step": total, "ms_per_step_per_layer_launches": round(ms_layers, 4),
             "launches_per_step": {"fused_chain": 2, "per_layer": len(layers)}}
    if not args.no_cpu_baseline:
        # one fresh pass on the device (the timed steps accumulate into layer 0) against the oracle
        layers[0] = S.SecureColumn.zeros(1 << (L - 1))
        step()
        S.synchronize()
        O = _oracle()
        host = np.stack([O.fill_uniform_m31(1 << L, 0x5EED0002, (2 << 40) + k * (1 << L)) for k in range(4)])
        t0 = time.perf_counter()
        cur = O.fold_circle_into_line(np.zeros((4, 1 << (L - 1)), np.uint32), host, root.initial_index,
                                      root.step_size, alphas[0])
        d = S.Coset(root.initial_index, root.step_size, root.log_size)
        outs = [cur]
        for i in range(1, len(layers)):
            cur = O.fold_line(cur, d.initial_index, d.step_size, alphas[i])
            d = d.double()
            outs.append(cur)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": total / dt / 1e9, "unit": "GB/s", "kind": "port",
                                 "sample": f"whole fold chain; {dt:.3f} s", "cores": _cores()}
        extra["parity_matches_oracle"] = bool(all((layers[i].to_numpy() == outs[i]).all() for i in range(len(layers))))
    return _line("FRI fold chain GB/s (log 22 circle -> line -> log 4)", total / (ms * 1e-3) / 1e9, "GB/s", ms, args,
                 {"workload": "FRI: fold_circle_into_line log 22 -> 21, then fold_line to log 4, seeded alphas "
                              "(fused chain: two launches)"},
                 _roof(wide, ks["fri_fold_chain"]["avg_ms"], "fri_fold_chain", "fri_fold_2e22_to_2e4"), extra)


def bench_quotients(args):
    """accumulate_quotients: 200 M31 columns at log 21, 3 random QM31 points, 1-3 samples per column."""
    import numpy as np
    import stwo_hip as S

    L, C = 21, 200
    n = 1 << L
    rng = np.random.default_rng(0x5EED0003)
    cols = [S.BaseFieldVec.uninitialized(n) for _ in range(C)]
    for c, col in enumerate(cols):
        S.fill_uniform(col, 0x5EED0003, c * n)
    points = [S.random_circle_point_qm31(rng) for _ in range(3)]
    per_point = [[] for _ in range(3)]
    for c in range(C):
        k = int(rng.integers(1, 4))
        for pi in sorted(rng.choice(3, size=k, replace=False)):
            per_point[pi].append((c, S.random_qm31(rng)))
    batches = [(points[i], per_point[i]) for i in range(3) if per_point[i]]
    samples = S.QuotientSamples(batches)
    rc = S.random_qm31(rng)
    out = S.SecureColumn.uninitialized(n)
    dom = S.canonic_half_coset(L)

    ptrs = S.column_pointers(cols)  # host pointer table built once, as a prover holds it

    def step():
        S.accumulate_quotients_async(dom, cols, rc, samples, out, col_ptrs=ptrs)
    ms = _timed_steps(S, step, args.steps, args.warmup)
    ks = _kernel_avg(S, step, ["accumulate_quotients"])
    alg = 4 * C * n + 16 * n
    n_samples = sum(len(cv) for _, cv in batches)
    extra = {"kernels": ks}
    if not args.no_cpu_baseline:
        O = _oracle()
        host = [O.fill_uniform_m31(n, 0x5EED0003, c * n) for c in range(C)]
        t0 = time.perf_counter()
        want = O.accumulate_quotients(dom.initial_index, dom.step_size, L, host, rc, batches)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": n / dt, "unit": "rows/s", "kind": "port", "sample": f"whole workload; {dt:.2f} s",
                                 "cores": _cores()}
        extra["parity_matches_oracle"] = bool((out.to_numpy() == want).all())
    return _line("accumulate_quotients rows/s", n / (ms * 1e-3), "rows/s", ms, args,
                 {"workload": f"quotients: {C} M31 columns log {L}, 3 QM31 points, {n_samples} samples"},
                 _roof(alg, ks["accumulate_quotients"]["avg_ms"], "accumulate_quotients", "quotients_200x2e21_3pts"),
                 extra)


def mixed_log_sizes(n_cols=1200, seed=0x5EED0004):
    """log sizes drawn from {10..22} with weights proportional to 1/size."""
    import numpy as np
    ks = np.arange(10, 23)
    p = 2.0 ** -ks
    return [int(k) for k in np.random.default_rng(seed).choice(ks, n_cols, p=p / p.sum())]


def bench_batched(args):
    """Mixed-size commit: 1200 columns of log 10..22, LDE at blowup 2 and one Merkle tree with
    mixed-degree injection.  Batched = one plan (size classes side by side on auxiliary
    streams, whole sequence replayed as a hipGraph, autotuned pass schedules).  Plain = the
    per-column trait path through the reference ABI: copy, interpolate, extend (zeroed
    buffer + copy), evaluate per column, then commit_on_layer per tree layer."""
    import numpy as np
    import stwo_hip as S

    logs = mixed_log_sizes()
    C = len(logs)
    L = S.lib()
    evals = []
    for c, l in enumerate(logs):
        v = S.BaseFieldVec.uninitialized(1 << l)
        S.fill_uniform(v, 0x5EED0004, c << 22)
        evals.append(v)
    coeffs = [S.BaseFieldVec.uninitialized(1 << l) for l in logs]
    ldes = [S.BaseFieldVec.uninitialized(2 << l) for l in logs]
    max_log = max(logs) + 1
    root = S.canonic_half_coset(max_log)
    tw, itw = S.cached_twiddles(root)
    tree_len = 1 << root.log_size
    elems = sum(2 << l for l in logs)
    t0 = time.perf_counter()
    plan = S.CommitPlan(evals, coeffs, ldes, logs, 1, tw, itw, tree_len, autotune=True)
    plan_s = time.perf_counter() - t0

    ms_graph = _timed_steps(S, lambda: plan.run(True), args.steps, args.warmup)  # the autotuned launch mode
    ms_stream = _timed_steps(S, lambda: plan.run(False), args.steps, args.warmup)
    launch = "hipGraph replay" if plan.uses_graph() else "stream launches (measured faster than the graph)"
    names = ["fft_inv_tile", "fft_fwd_tile", "fft_inv_contig", "fft_fwd_contig", "fft_inv_strided", "fft_fwd_strided",
             "fft_lde_strided", "fft_tiny_inv", "fft_tiny_fwd", "merkle_leaf", "merkle_layer", "merkle_subtree"]
    ks = _kernel_avg(S, lambda: plan.run(False), names)
    for v in ks.values():
        v["ms_per_step"] = round(v["avg_ms"] * v["launches_per_step"], 4)
    root_batched = plan.root()
    layer0 = plan.layers[0].to_numpy().reshape(-1)

    # plain per-column path (synchronous reference ABI calls, as the trait implementation issues them)
    p_coeffs = [S.BaseFieldVec.uninitialized(1 << l) for l in logs]
    by_layer = {}
    for c, l in enumerate(logs):
        by_layer.setdefault(l + 1, []).append(c)

    def plain_step():
        exts = []
        for c, l in enumerate(logs):
            n = 1 << l
            L.copy_uint32_t_vec_from_device_to_device(evals[c].ptr, p_coeffs[c].ptr, n)
            L.interpolate(n // 2, p_coeffs[c].ptr, itw, tree_len, n)
            ext = L.cuda_alloc_zeroes_uint32_t(2 * n)
            L.copy_uint32_t_vec_from_device_to_device(p_coeffs[c].ptr, ext, n)
            L.evaluate(n, ext, tw, tree_len, 2 * n)
            exts.append(ext)
        prev = None
        for k in range(max_log, -1, -1):
            cols = [S.BaseFieldVec.view(exts[c], 2 << logs[c]) for c in by_layer.get(k, [])]
            nxt = S.commit_on_layer(k, prev, cols)
            prev = nxt
        for e in exts:
            L.cuda_free_memory(e)
        S.check()
        return prev
    plain_steps = max(1, min(args.steps, 3))
    for _ in range(1):
        plain_step()
    S.synchronize()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        top = plain_step()
    S.synchronize()
    ms_plain = (time.perf_counter() - t0) * 1e3 / plain_steps
    plain_root = top.to_numpy().reshape(-1)
    pipe = sum(28 * (1 << l) for l in logs) + sum(96 * (1 << k) for k in range(max_log + 1))
    gbs = pipe / (ms_graph * 1e-3) / 1e9
    extra = {
        "batched_plan_ms": round(ms_graph, 4), "batched_stream_ms": round(ms_stream, 4), "launch": launch,
        "plain_per_column_ms": round(ms_plain, 3), "speedup_vs_plain": round(ms_plain / ms_graph, 1),
        "plan_create_and_autotune_s": round(plan_s, 3), "columns": C,
        "kernels_stream_mode": ks,
        "columns_per_log": {str(k): logs.count(k) for k in range(10, 23)},
        "root": "".join(f"{w:08x}" for w in root_batched),
        "plain_root_matches_batched": bool((plain_root == layer0).all()),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
        host = [O.fill_uniform_m31(1 << l, 0x5EED0004, c << 22) for c, l in enumerate(logs)]
        t0 = time.perf_counter()
        olde = [None] * C
        for k in sorted(set(logs)):
            idx = [c for c in range(C) if logs[c] == k]
            _, ls = O.lde_columns([host[c] for c in idx], 1, otw, oitw)
            for c, v in zip(idx, ls):
                olde[c] = v
        oroot = O.merkle_commit(olde)[0]
        dt = time.perf_counter() - t0
        extra["parity_root_matches_oracle"] = bool((oroot == root_batched).all())
        extra["cpu_baseline"] = {"value": elems / dt, "unit": "elems/s", "kind": "port", "sample": "whole config",
                                 "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    plan.close()
    return _line("LDE+Merkle M31 elems/s (batched mixed sizes)", elems / (ms_graph * 1e-3), "elems/s",
                 ms_graph, args,
                 {"workload": f"mixed commit: {C} M31 columns, log sizes 10..22 (weights 1/size), LDE blowup 2, one "
                              f"Blake2s tree with mixed-degree injection; {launch}"},
                 {"bound": "hbm", "kernel": "whole step (all kernels)", "achieved": round(gbs, 1),
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                  "algorithmic_bytes_per_launch": pipe, "avg_launch_ms": round(ms_graph, 4)}, extra)


def serving_jobs(n_jobs, rate, seed=0x5EED0005, logs=(12, 21), cols=(32, 64, 128, 256, 512)):
    """Poisson arrivals (seeded): (arrival_s, n_cols, log_n, trace seed)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    c = rng.choice(list(cols), n_jobs)
    lg = rng.integers(logs[0], logs[1], n_jobs)
    return [(float(t[i]), int(c[i]), int(lg[i]), 0xA000 + i) for i in range(n_jobs)]


def _serve(S, jobs, budget, slots, pow_bits, last_log, work_limit=0, max_batch=1, max_batch_work=0):
    """Submit each job at its arrival time; returns per-job ((submit, start, end), proof words)."""
    srv = S.ProofServer(budget, slots, work_limit, max_batch, max_batch_work)
    t0 = time.perf_counter()
    done, ids = {}, {}
    i = 0
    while len(done) < len(jobs):
        now = time.perf_counter() - t0
        while i < len(jobs) and jobs[i][0] <= now:
            _, C, ln, seed = jobs[i]
            ids[srv.submit(S.wide_fib_params(C, ln, 3, pow_bits, last_log), seed)] = i
            i += 1
        wait_us = 200 if i < len(jobs) else 100_000
        if i < len(jobs):
            wait_us = max(20, min(wait_us, int((jobs[i][0] - now) * 1e6)))
        r = srv.wait(wait_us)
        while r is not None:
            jid, st, proof, times = r
            if st != 0:
                raise RuntimeError(f"job {ids[jid]} failed with status {st}")
            done[ids[jid]] = (times, proof.size)
            r = srv.wait(0)
    srv.close()
    return done


def _serve_plain(jobs, pow_bits, last_log):
    """The reference-shaped path: jobs one after another, each through one synchronous call per
    backend operation (wide_fib_plain); arrival, start and end times per job."""
    import wide_fib_plain
    t0 = time.perf_counter()
    done = {}
    for i, (arrival, C, ln, seed) in enumerate(jobs):
        now = time.perf_counter() - t0
        if now < arrival:
            time.sleep(arrival - now)
        start = time.perf_counter() - t0
        proof = wide_fib_plain.prove(C, ln, 3, pow_bits, last_log, seed)
        end = time.perf_counter() - t0
        done[i] = ((arrival, start, end, end - start), proof.size)
    return done


def bench_serving(args):
    """Wide-Fibonacci proof serving: a seeded Poisson stream of prove jobs (n_cols in
    {32..512}, log_n in {12..20}, blowup 2, 3 queries) through three modes on one GPU:
      * batched: the continuous-batching server (8 slot streams, HBM admission, a work limit,
        same-phase batching of same-shape small jobs);
      * device prover one job at a time (the server with one slot, no batching);
      * plain: the reference-shaped path (one synchronous part-1 ABI call per backend operation,
        host channel, one gather per tree for the decommitments).
    Capacity: the burst stream (offered load above what one GPU proves).  Equal p99: every mode
    runs the same job stream at rising Poisson rates; the latency target is 1.5x the plain
    path's own p99 at the lowest rate, and each mode is credited with the proofs/s of the
    highest rate whose p99 stays within it.  A burst of small same-shape jobs (log_n 12..14)
    shows what the same-phase batching adds over overlapping streams."""
    import numpy as np
    import stwo_hip as S

    n_jobs, pow_bits, last_log = args.serve_jobs, 10, 2
    budget = args.serve_budget_gib << 30
    wl, mb, mbw = args.serve_work_limit, args.serve_max_batch, args.serve_batch_work
    # warm the twiddle cache and the memory pool with one pass of every shape, off the clock
    shapes = sorted({(C, ln) for _, C, ln, _ in serving_jobs(n_jobs, 1000.0)})
    warm = S.ProofServer(budget, args.serve_slots)
    for C, ln in shapes:
        warm.submit(S.wide_fib_params(C, ln, 3, pow_bits, last_log), 1)
    for _ in shapes:
        r = warm.wait(60_000_000)
        if r is None or r[1] != 0:
            raise RuntimeError(f"warm-up job failed: {r and r[1]}: {S.last_error()}")
    warm.close()

    def summarize(done, n):
        lat = np.array([done[i][0][2] - done[i][0][0] for i in range(n)])
        first = min(done[i][0][0] for i in range(n))
        last = max(done[i][0][2] for i in range(n))
        return {"proofs_per_s": round(n / (last - first), 1), "p50_ms": round(float(np.percentile(lat, 50) * 1e3), 2),
                "p99_ms": round(float(np.percentile(lat, 99) * 1e3), 2), "makespan_s": round(last - first, 4),
                "host_ms_per_job": round(float(np.mean([done[i][0][3] for i in range(n)])) * 1e3, 3)}

    modes = {
        "batched": lambda js: _serve(S, js, budget, args.serve_slots, pow_bits, last_log, wl, mb, mbw),
        "device_prover_one_at_a_time": lambda js: _serve(S, js, budget, 1, pow_bits, last_log),
        "plain_reference_shaped": lambda js: _serve_plain(js, pow_bits, last_log),
    }
    rate = args.serve_rate
    capacity = {m: summarize(f(serving_jobs(n_jobs, rate)), n_jobs) for m, f in modes.items()}

    # equal p99: rising rates, each mode until its p99 leaves the target
    rates = sorted(args.serve_sweep)
    slo_ms, sweep, best = None, {m: [] for m in modes}, {}
    for m in ["plain_reference_shaped", "device_prover_one_at_a_time", "batched"]:
        for r in rates:
            res = summarize(modes[m](serving_jobs(n_jobs, r)), n_jobs)
            res["rate"] = r
            sweep[m].append(res)
            if slo_ms is None:
                slo_ms = round(1.5 * res["p99_ms"], 2)
            if res["p99_ms"] > slo_ms:
                break
            best[m] = res
    at_slo = {m: best.get(m) for m in modes}

    def ratio(a, b):
        return round(at_slo[a]["proofs_per_s"] / at_slo[b]["proofs_per_s"], 2) if at_slo[a] and at_slo[b] else None

    # small same-shape jobs in a burst (all arrive within ~1 ms, so arrivals do not pace them):
    # streams only vs streams + same-phase batching
    small = serving_jobs(n_jobs, 1e5, seed=0x5EED0006, logs=(12, 15), cols=(32, 64))
    small_res = {
        "batched_same_phase": summarize(_serve(S, small, budget, args.serve_slots, pow_bits, last_log, wl, mb, mbw),
                                        n_jobs),
        "streams_only": summarize(_serve(S, small, budget, args.serve_slots, pow_bits, last_log, wl), n_jobs),
        "device_prover_one_at_a_time": summarize(_serve(S, small, budget, 1, pow_bits, last_log), n_jobs),
    }
    elems = sum(C << ln for _, C, ln, _ in serving_jobs(n_jobs, rate))
    extra = {"capacity": capacity,
             "throughput_ratio_vs_plain": round(capacity["batched"]["proofs_per_s"] /
                                                capacity["plain_reference_shaped"]["proofs_per_s"], 2),
             "throughput_ratio_vs_one_at_a_time": round(capacity["batched"]["proofs_per_s"] /
                                                        capacity["device_prover_one_at_a_time"]["proofs_per_s"], 2),
             "equal_p99": {"p99_target_ms": slo_ms, "at_target": at_slo,
                           "ratio_batched_vs_plain": ratio("batched", "plain_reference_shaped"),
                           "ratio_batched_vs_one_at_a_time": ratio("batched", "device_prover_one_at_a_time"),
                           "sweep": sweep},
             "small_same_shape_burst": {"workload": f"{n_jobs} jobs arriving at 1e5/s, n_cols in {{32, 64}}, "
                                                    "log_n in 12..14",
                                        **small_res,
                                        "ratio_batching_vs_streams": round(
                                            small_res["batched_same_phase"]["proofs_per_s"] /
                                            small_res["streams_only"]["proofs_per_s"], 2)},
             "jobs": n_jobs, "arrival_rate_per_s": rate, "slots": args.serve_slots, "work_limit": wl,
             "max_batch": mb, "max_batch_work": mbw, "hbm_budget_gib": args.serve_budget_gib,
             "trace_elements_total": elems, "pow_bits": pow_bits, "last_layer_log": last_log}
    return _line("served wide-Fibonacci proofs/s (continuous batching)", capacity["batched"]["proofs_per_s"],
                 "proofs/s", 1e3 / capacity["batched"]["proofs_per_s"], args,
                 {"workload": f"serving: {n_jobs} Poisson(rate {rate}/s) wide-Fibonacci prove jobs, n_cols in "
                              "{32,64,128,256,512}, log_n in 12..20, blowup 2, 3 queries, Blake2s channel"},
                 extra=extra)


class _Slab:
    """One device allocation of n_cols columns of 2^log_n values back to back, with non-owning
    column views (the reference benches' 2^16 small columns without one allocation each)."""

    def __init__(self, S, n_cols, log_n):
        self.buf = S.BaseFieldVec.uninitialized(n_cols << log_n)
        self.cols = [S.BaseFieldVec.view(self.buf.ptr + (c << (log_n + 2)), 1 << log_n) for c in range(n_cols)]


def bench_ref_interpolate(args):
    """The reference's benches/interpolate_columns.rs shape: interpolate_columns of 2^16 columns of
    2^10 values (in place, through the reference ABI with its host pointer array)."""
    import stwo_hip as S

    log_n, n_cols = 10, 1 << 16
    slab = _Slab(S, n_cols, log_n)
    S.check(S.lib().stwo_hip_fill_uniform_m31(slab.buf.ptr, n_cols << log_n, 0x5EED0008, 0))
    coset = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(coset)
    tw, itw = S.cached_twiddles(coset)
    # the host pointer arrays are built once (a Rust caller passes its Vec of column pointers)
    import numpy as np
    ptrs = S._ptr_array([c.ptr for c in slab.cols])
    logs = np.full(n_cols, log_n, np.uint32)
    L = S.lib()

    def step_ref():  # reference ABI: synchronous, uploads the pointer array itself
        L.interpolate_columns(1 << (log_n - 1), ptrs, tree.itwiddles.ptr, tree.itwiddles.size, n_cols, 1 << log_n)

    def step():
        S.check(L.stwo_hip_interpolate_columns_async(ptrs, logs.ctypes.data_as(S.u32p), n_cols, itw, 1 << coset.log_size))
    ms = _timed_steps(S, step_ref, args.steps, args.warmup)
    ms_async = _timed_steps(S, step, args.steps, args.warmup)
    ks = _kernel_avg(S, step, ["fft_inv_tile"])
    elems = n_cols << log_n
    extra = {"kernels": ks, "ms_per_step_async_batched": round(ms_async, 4),
             "note": "ms_per_step is the synchronous reference-ABI call (pointer-array upload + launch + sync)"}
    if not args.no_cpu_baseline:
        O = _oracle()
        sample = 4096
        host = O.fill_uniform_m31(sample << log_n, 1, 0).reshape(sample, -1)
        otw, oitw = O.precompute_twiddles(coset.initial_index, coset.step_size, coset.log_size)
        t0 = time.perf_counter()
        O.lde_columns(list(host), 0, otw, oitw)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": (sample << log_n) / dt, "unit": "elems/s", "kind": "port", "cores": _cores(),
                                 "sample": f"{sample} of the 2^16 columns; {dt:.3f} s"}
    return _line("interpolate_columns M31 elems/s (2^16 columns x 2^10)", elems / (ms * 1e-3), "elems/s", ms, args,
                 {"workload": "reference bench interpolate_columns: 2^16 columns of 2^10, in place"},
                 _roof(8 * elems, ks["fft_inv_tile"]["avg_ms"], "fft_inv_tile", "interpolate_columns_2e16x2e10"), extra)


def bench_ref_evaluate(args):
    """The reference's benches/evaluate_columns.rs shape: evaluate_polynomials of 2^16 polynomials
    of 2^10 coefficients at blowup 4 (domain 2^12).  value: the batched path from the coefficients
    (the zero padding is implicit: no extend copy)."""
    import stwo_hip as S

    log_n, blow, n_cols = 10, 2, 1 << 16
    m = log_n + blow
    coef = _Slab(S, n_cols, log_n)
    S.check(S.lib().stwo_hip_fill_uniform_m31(coef.buf.ptr, n_cols << log_n, 0x5EED0009, 0))
    out = _Slab(S, n_cols, m)
    big = S.canonic_half_coset(m)
    tw, itw = S.cached_twiddles(big)
    tl = 1 << big.log_size

    # the host pointer arrays are built once (a Rust caller passes its Vec of column pointers)
    import numpy as np
    src, dst = S._ptr_array([c.ptr for c in coef.cols]), S._ptr_array([c.ptr for c in out.cols])
    cl, el = np.full(n_cols, log_n, np.uint32), np.full(n_cols, m, np.uint32)
    L = S.lib()

    def step():
        S.check(L.stwo_hip_evaluate_columns_async(src, cl.ctypes.data_as(S.u32p), dst, el.ctypes.data_as(S.u32p),
                                                  n_cols, tw, tl))

    ms = _timed_steps(S, step, args.steps, args.warmup)
    ks = _kernel_avg(S, step, ["fft_fwd_tile"])
    extra = {"kernels": ks}
    if not args.no_cpu_baseline:
        O = _oracle()
        import numpy as np
        sample = 512
        host = O.fill_uniform_m31(sample << log_n, 2, 0).reshape(sample, -1)
        otw, _ = O.precompute_twiddles(big.initial_index, big.step_size, big.log_size)
        t0 = time.perf_counter()
        for c in range(sample):
            ext = np.zeros(1 << m, np.uint32)
            ext[:1 << log_n] = host[c]
            O.evaluate(ext, big.initial_index, otw)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": (sample << m) / dt, "unit": "elems/s", "kind": "port", "cores": 1,
                                 "sample": f"{sample} of the 2^16 polynomials, one call per column; {dt:.3f} s"}
    elems = n_cols << m
    return _line("evaluate_polynomials M31 elems/s (2^16 x 2^10, blowup 4)", elems / (ms * 1e-3), "elems/s", ms, args,
                 {"workload": "reference bench evaluate_columns: 2^16 polynomials of 2^10 coefficients, blowup 4"},
                 _roof(4 * (n_cols << log_n) + 4 * elems, ks["fft_fwd_tile"]["avg_ms"], "fft_fwd_tile",
                       "evaluate_polynomials_2e16x2e10_blowup4"), extra)


def bench_ref_bit_reverse(args):
    """The reference's benches/bit_reverse.rs shape: bit_reverse_column of 2^28 M31 (value) and
    2^28 QM31 (AoS, in the extra fields), in place."""
    import stwo_hip as S

    n = 1 << 28
    res = {}
    for secure in (False, True):
        V = S.SecureFieldVec if secure else S.BaseFieldVec
        col = V.uninitialized(n)
        S.check(S.lib().stwo_hip_fill_uniform_m31(col.ptr, 4 * n if secure else n, 0x5EED0007, 0))
        ms = _timed_steps(S, lambda: S.bit_reverse_column(col), args.steps, args.warmup)
        ks = _kernel_avg(S, lambda: S.bit_reverse_column(col), ["bit_reverse"])
        res["qm31" if secure else "m31"] = (ms, ks, (32 if secure else 8) * n)
        del col
    ms, ks, b = res["m31"]
    extra = {"kernels": ks, "qm31": {"ms_per_step": round(res["qm31"][0], 4), "kernels": res["qm31"][1],
                                     "roofline": _roof(res["qm31"][2], res["qm31"][1]["bit_reverse"]["avg_ms"],
                                                       "bit_reverse_qm31", "bit_reverse_2e28")},
             "note": "ms_per_step is the synchronous reference-ABI call"}
    if not args.no_cpu_baseline:
        O = _oracle()
        host = O.fill_uniform_m31(1 << 26, 3, 0)
        t0 = time.perf_counter()
        O.bit_reverse(host)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": (1 << 26) / dt, "unit": "elems/s", "kind": "port", "cores": _cores(),
                                 "sample": f"2^26 M31; {dt:.3f} s"}
    return _line("bit_reverse M31 elems/s (2^28)", n / (ms * 1e-3), "elems/s", ms, args,
                 {"workload": "reference bench bit_reverse: 2^28 M31 in place (and 2^28 QM31)"},
                 _roof(b, ks["bit_reverse"]["avg_ms"], "bit_reverse", "bit_reverse_2e28"), extra)


def bench_ref_batch_inverse(args):
    """The reference's benches/batch_inverse.rs shape: batch_inverse of 2^28 M31 (value) and 2^28
    QM31 (in the extra fields)."""
    import stwo_hip as S

    n = 1 << 28
    res = {}
    for secure in (False, True):
        V = S.SecureFieldVec if secure else S.BaseFieldVec
        src, dst = V.uninitialized(n), V.uninitialized(n)
        S.check(S.lib().stwo_hip_fill_uniform_m31(src.ptr, 4 * n if secure else n, 0x5EED0006, 0))
        name = "batch_inverse_qm31" if secure else "batch_inverse_m31"
        ms = _timed_steps(S, lambda: S.batch_inverse(src, dst), args.steps, args.warmup)
        ks = _kernel_avg(S, lambda: S.batch_inverse(src, dst), [name])
        res[name] = (ms, ks, (32 if secure else 8) * n)
        del src, dst
    ms, ks, b = res["batch_inverse_m31"]
    q = res["batch_inverse_qm31"]
    extra = {"kernels": ks, "qm31": {"ms_per_step": round(q[0], 4), "kernels": q[1],
                                     "roofline": _roof(q[2], q[1]["batch_inverse_qm31"]["avg_ms"], "batch_inverse_qm31",
                                                       "batch_inverse_2e28")},
             "note": "ms_per_step is the synchronous reference-ABI call"}
    if not args.no_cpu_baseline:
        O = _oracle()
        host = O.fill_uniform_m31(1 << 25, 4, 0)
        host[host == 0] = 1
        t0 = time.perf_counter()
        O.batch_inverse_m31(host)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": (1 << 25) / dt, "unit": "elems/s", "kind": "port", "cores": _cores(),
                                 "sample": f"2^25 M31; {dt:.3f} s"}
    return _line("batch_inverse M31 elems/s (2^28)", n / (ms * 1e-3), "elems/s", ms, args,
                 {"workload": "reference bench batch_inverse: 2^28 M31 (and 2^28 QM31)"},
                 _roof(b, ks["batch_inverse_m31"]["avg_ms"], "batch_inverse_m31", "batch_inverse_2e28"), extra)


CONFIGS = {"lde-commit": bench_lde_commit, "small": bench_small, "batch-inverse": bench_batch_inverse,
           "fri": bench_fri, "quotients": bench_quotients, "mixed-commit": bench_batched, "serving": bench_serving,
           "ref-interpolate": bench_ref_interpolate, "ref-evaluate": bench_ref_evaluate,
           "ref-bit-reverse": bench_ref_bit_reverse, "ref-batch-inverse": bench_ref_batch_inverse}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="lde-commit", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-sample-cols", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=120)
    ap.add_argument("--serve-rate", type=float, default=4000.0,
                    help="Poisson arrival rate (jobs/s); the default offers more than one GPU proves, so "
                         "proofs/s is the server's capacity")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-work-limit", type=int, default=1 << 31,
                    help="trace elements (n_cols x 2^log_n) in flight at most (0: no limit); 2^31, four of "
                         "the largest jobs, measured best with 8 slots")
    ap.add_argument("--serve-budget-gib", type=int, default=96)
    ap.add_argument("--serve-sweep", type=float, nargs="*", default=[25.0, 50.0, 100.0, 200.0, 400.0, 800.0, 1600.0],
                    help="arrival rates (jobs/s) of the equal-p99 comparison")
    ap.add_argument("--serve-max-batch", type=int, default=8,
                    help="same-shape jobs proved as one batch by the server (1: no same-phase batching)")
    ap.add_argument("--serve-batch-work", type=int, default=1 << 22,
                    help="only jobs of at most this many trace elements (n_cols x 2^log_n) batch")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device copy (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this backend runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    import stwo_hip as S

    S.check(S.lib().stwo_hip_init(0))
    if args.calibrate_copy:
        # 1 GiB read + 1 GiB written by the library's plain copy kernel
        a, b = S.BaseFieldVec.uninitialized(1 << 28), S.BaseFieldVec.uninitialized(1 << 28)
        S.device_copy(b, a)
        S.synchronize()
        del a, b
    print(json.dumps(CONFIGS[args.config](args)))


if __name__ == "__main__":
    main()
