#!/usr/bin/env python3
"""Headline benchmark: pair interactions/s and steps/s of the ion-flow step at
2^20 superions (single dust, all-pairs), FP32-compensated pair arithmetic,
with the plain FP64 GPU path and the FP64 CPU oracle timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line on stdout; progress goes to stderr.

A step is one full reference step (opendust.py:1059-1384: collisions, KDK,
flags, absorbed counts and momenta, reinjection, orbit sums, charge update,
dust forces) over all N ions, with the ion state resident on the GPU.
pair interactions/s = N (N - 1) x steps/s.

Extra modes (used for the committed profiles):
    --copy-calibration   plain device copies of known size (rocprof PMC scale)
    --pair-only K        K launches of the main pair kernel only
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 (vector with packed ops; also matrix), TFLOP/s
FLOP_PER_PAIR = 24         # algorithmic FLOPs of one pair term
# HBM bytes per k_pair32_sym launch at 2^20 ions for a given kernel variant and
# number of source blocks per work-group, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# scaled by a 1 GiB copy's readings (x2.000, x1.000), with the XCD-aware tile launch
# order of the engine: profiles/k_pair32_sym_2e20_pmc.md
# (variants 2 and 4 spill registers; their scratch traffic is included)
SYM_TRAFFIC_2E20 = {  # (kernel variant, source blocks per work-group): bytes per launch
    (1, 1): 20.625e9, (1, 2): 13.926e9, (1, 4): 10.656e9, (1, 8): 9.014e9, (1, 16): 8.275e9, (1, 32): 7.890e9,
    (2, 1): 40.530e9, (2, 2): 25.591e9, (2, 4): 18.874e9, (2, 8): 14.837e9, (2, 16): 13.301e9, (2, 32): 12.711e9,
    (3, 1): 19.686e9, (3, 2): 13.698e9, (3, 4): 10.542e9, (3, 8): 9.070e9, (3, 16): 8.286e9, (3, 32): 7.694e9,
    (4, 1): 26.310e9, (4, 2): 16.920e9, (4, 4): 12.250e9, (4, 8): 10.366e9, (4, 16): 9.452e9, (4, 32): 9.830e9,
    (6, 1): 20.625e9, (6, 2): 14.032e9, (6, 4): 10.691e9, (6, 8): 9.057e9, (6, 16): 8.325e9, (6, 32): 8.095e9,
    (7, 1): 20.721e9, (7, 2): 14.036e9, (7, 4): 10.715e9, (7, 8): 9.090e9, (7, 16): 8.332e9, (7, 32): 8.110e9,
    (9, 4): 7.225e9, (9, 8): 5.726e9, (9, 16): 4.930e9,
    (10, 4): 6.524e9, (10, 8): 4.319e9, (10, 16): 3.241e9, (10, 32): 2.803e9,
    (11, 1): 21.417e9, (11, 2): 14.630e9, (11, 4): 10.970e9, (11, 8): 9.214e9, (11, 16): 8.386e9, (11, 32): 8.136e9,
    (12, 1): 18.450e9, (12, 2): 11.139e9, (12, 4): 6.519e9, (12, 8): 4.318e9, (12, 16): 3.247e9, (12, 32): 2.825e9,
    (13, 1): 21.992e9, (13, 2): 11.048e9, (13, 4): 7.779e9, (13, 8): 5.937e9, (13, 16): 5.052e9, (13, 32): 4.665e9,
}
# variants whose registers spill (scratch traffic of their own); for the others
# the schedule fixes the traffic: the partial-sum writes are identical and the
# position reads (0.4-1.5 GB) differ, so variants 1, 3, 6 and 7 agree within
# 5 % at every profiled block count
SYM_SPILLING_VARIANTS = (2, 4)


def pair_evaluations(n, sym, pairs_launch, n_flagged, n_absorbed_rows):
    """Pair-term evaluations of one step (see the bench line's
    kernel_pair_evaluations_per_step)."""
    main = pairs_launch / 2.0 if sym else float(pairs_launch)
    in_block = float(n) * (n - 1) - pairs_launch if sym else 0.0
    corrections = 2.0 * (n - n_flagged) * n_flagged + float(n) * (n_flagged + n_absorbed_rows)
    return {"main_launch": main, "in_block": in_block, "flagged_corrections": corrections,
            "total": main + in_block + corrections, "per_interaction": (main + in_block + corrections) / (n * (n - 1.0))}


def sym_traffic_2e20(tile, chunk, rows=1):
    """(bytes per launch, source) of k_pair32_sym at 2^20 ions: the PMC
    reading of this (variant, block count), else that of variant 1 with the
    same block count when neither spills and the variant does not group
    target blocks (rows > 1 writes fewer partials), else (None, None)."""
    if (tile, chunk) in SYM_TRAFFIC_2E20:
        return SYM_TRAFFIC_2E20[(tile, chunk)], "pmc"
    if rows == 1 and tile not in SYM_SPILLING_VARIANTS and (1, chunk) in SYM_TRAFFIC_2E20:
        return SYM_TRAFFIC_2E20[(1, chunk)], "pmc of variant 1 with the same schedule (neither spills)"
    return None, None
METRIC = "pair interactions/s + steps/s at 2^20 ions (FP32-comp vs FP64); sweep sims/hour"
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
# Algorithmic HBM bytes per ion of the per-ion stages of a step (FP32c): what
# the stage must read and write once, flagged-ion lists excluded.
STAGE_BYTES_PER_ION = {
    "kick_drift": 156,  # read X, V, F (72); write X, V (48), flags (4), FP32 positions pi, pj (32)
    "list": 4,          # read the flags
    "kick2": 117,       # read pi (16), flags (4), X, V, F (72); write V (24), face (1)
    "inflow": 68,       # read pj (16), flags (4); read and write F (48)
}


# cutoff mode: k_kick2 also reads the flagged sources' share G (24); k_inflow
# only reads the flags and writes the reinjected ions' rows (the kept ions'
# share of the reinjected sources is added by the cell correction kernel)
STAGE_BYTES_PER_ION_CUTOFF = dict(STAGE_BYTES_PER_ION, kick2=141, inflow=4)


def cell_stage_bytes_per_ion(n_cells):
    """Binning (read X 24 + flags 4, write key and value 8), one 8-bit radix
    pass per digit of the cell ids (read key 4 for the histogram, read and
    write key and value 16), gather (read key, value 8 and the position 16,
    write the sorted position 16)."""
    passes = 1
    while passes < 4 and (n_cells >> (8 * passes)) != 0:
        passes += 1
    return 36 + 20 * passes + 40, passes


def hbm_stages(stage_ms, n_ions, bytes_per_ion):
    """GB/s of the per-ion stages against the HBM peak."""
    out = {}
    for name, b in bytes_per_ion.items():
        ms = stage_ms.get(name, 0.0)
        if ms <= 0:
            continue
        gbs = b * n_ions / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "bytes_per_ion": b, "GB_per_s": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_calibration(nbytes=1 << 30, reps=8):
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def cpu_baseline(cfg, x, target_s=12.0, small_steps=10):
    """The FP64 oracle on the host CPU, rebuilt there (-O3 -march=native), all
    host threads: (a) ion-ion force rows of a sample of the benchmark's 2^20
    ions against all of them (the workload of `value`), (b) full oracle steps
    of the 16384-ion single-dust workload (two O(N^2) force evaluations per
    step, as the reference's step)."""
    from oracle import oracle
    from opendust_amd import workloads

    build = oracle.use_native_build()
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or oracle.threads()
    n = x.shape[0]
    rows = 8 * nthreads
    t1 = 0.0
    while t1 < 1.0 and rows < n:
        rows = min(n, rows * 4)
        t0 = time.perf_counter()
        oracle.pair_forces(cfg, x, rows=(0, rows), nthreads=nthreads)
        t1 = time.perf_counter() - t0
    rows = int(min(n, max(rows, rows * target_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    oracle.pair_forces(cfg, x, rows=(0, rows), nthreads=nthreads)
    dt = time.perf_counter() - t0
    pairs = rows * (n - 1)
    import contextlib
    import io

    with contextlib.redirect_stdout(io.StringIO()):
        od0 = workloads.single_dust(N=16384, n=small_steps, precision="fp64", seed=12345)
        cfg0, x0, v0 = workloads.prepare(od0, alpha=workloads.synthetic_alpha(od0), np_seed=7, precision="fp64")
    q0 = np.array([cfg0.dust_q0[0]])
    t0 = time.perf_counter()
    oracle.run(cfg0, x0, v0, q0, 0, small_steps, nthreads=nthreads)
    dt0 = time.perf_counter() - t0
    n0 = cfg0.n_ions
    return {
        "value": pairs / dt,
        "unit": "pair interactions/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "FP64 oracle ion-ion force rows 0..%d of %d ions against all %d ions (%.1f s); "
                  "the oracle step evaluates this O(N^2) force twice per step" % (rows, n, n, dt),
        "cpu_model": oracle.cpu_model(),
        "os_cpu_count": os.cpu_count(),
        "threads": nthreads,
        "build": build,
        "small_workload_full_steps": {
            "workload": "single dust, 16384 superions, R=5 r_D_e, H=10 r_D_e, FP64, full oracle steps",
            "steps": small_steps,
            "s_per_step": dt0 / small_steps,
            "steps_per_s": small_steps / dt0,
            "pair_interactions_per_s": float(n0) * (n0 - 1) * small_steps / dt0,
        },
    }


def cutoff_leg(steps):
    """2^22 superions, cylinder R = 10, H = 20 r_D_e, ion-ion force truncated
    at 2 r_D_e through the per-step cell list; in-range pair interactions/s."""
    from opendust_amd import engine, workloads

    od = workloads.cutoff_large(N=2 ** 22, n=steps + 1, seed=99, trapSeed=7)
    cfg, x0, v0 = workloads.prepare(od, alpha=None, np_seed=11, precision="fp32c")
    e = engine.Engine(cfg)
    e.set_ions(x0, v0)
    _, variant = e.tune(2)
    e.run(1, engine.RUN_PROFILE_PAIR)
    e.pair_timing(reset=True)
    t0 = time.perf_counter()
    e.run(steps, engine.RUN_PROFILE_PAIR)
    wall = time.perf_counter() - t0
    ms, nl = e.pair_timing()
    in_range, cand = e.cutoff_pair_count()
    _, _, dims = e.cells()
    e.run(2, engine.RUN_PROFILE_STAGES)
    st, _ = e.stage_timing()
    e.close()
    n_cells = dims[0] * dims[1] * dims[2]
    cb, passes = cell_stage_bytes_per_ion(n_cells)
    return {
        "workload": "single dust, 2^22 superions, R=10 r_D_e, H=20 r_D_e, hard cutoff 2 r_D_e, cell list rebuilt "
                    "every step",
        "steps": steps,
        "ms_per_step": wall / steps * 1e3,
        "in_range_pairs_per_step": in_range,
        "candidate_pairs_per_step": cand,
        "pair_interactions_per_s": in_range * steps / wall,
        "cell_kernel_ms": ms / max(nl, 1),
        "cell_kernel_tflops_alg": in_range * FLOP_PER_PAIR / (ms / max(nl, 1) * 1e-3) / 1e12,
        "cell_kernel_frac_of_fp32_peak": in_range * FLOP_PER_PAIR / (ms / max(nl, 1) * 1e-3) / 1e12
                                         / FP32_PEAK_TFLOPS,
        "cell_kernel_variant": {"sym_tile": variant, "targets_per_lane": {1: 1, 2: 2, 3: 4, 4: 1, 5: 2, 6: 4}.get(variant),
                                "sources_from_lds": variant in (4, 5, 6)},
        "grid": {"dims": list(dims), "cells": n_cells, "radix_passes": passes},
        "stage_ms": st,
        "hbm_stages": hbm_stages(st, 2 ** 22, dict(STAGE_BYTES_PER_ION_CUTOFF, cells=cb)),
    }


def sweep_leg(steps, seq_steps, n_sims=64, sweep_len=5000):
    """Parameter sweep: 64 independent simulations (N log-uniform in [2^12,
    2^17], M in [0.5, 2], 1-4 dusts), FP32c, graph-replayed steps.  Batched
    (one od_batch) against the same simulations run one after another through
    single contexts; the step loop of a sweep_len-step sweep is extrapolated
    from `steps` (batched) and `seq_steps` (per simulation) timed steps.  Setup
    (initial states, trap fit) is not timed on either side."""
    import contextlib
    import io

    from opendust_amd import engine, sweep, workloads

    with contextlib.redirect_stdout(io.StringIO()):
        ods = workloads.sweep(n_sims=n_sims, n=sweep_len, precision="fp32c", seed=7, trapSeed=5,
                              trapSamples=4_000_000)
    np.random.seed(21)
    cfgs, states = sweep.prepare_batch(ods, device=0)
    pairs = float(sum(c.n_ions * (c.n_ions - 1.0) for c in cfgs))
    ions = int(sum(c.n_ions for c in cfgs))
    b = engine.Batch(cfgs)
    for s_, (x0, v0) in enumerate(states):
        b.set_ions(s_, x0, v0)
    chunk, tile = b.tune(3)
    log("sweep: autotuner chose %d source blocks per work-group, kernel variant %d %s"
        % (chunk, tile, engine.sym_variants()[tile - 1]))
    b.run(3, engine.RUN_PROFILE_PAIR, records=False)
    b.pair_timing(reset=True)
    b.run(3, engine.RUN_PROFILE_PAIR, records=False)
    kms, nl, ppl = b.pair_timing()
    b.run(16, engine.RUN_GRAPH, records=False)
    t0 = time.perf_counter()
    b.run(steps, engine.RUN_GRAPH, records=False)
    t_batch = (time.perf_counter() - t0) / steps
    b.close()
    log("sweep batched: %.2f ms/step" % (t_batch * 1e3))
    t_seq = 0.0
    for s_, (cfg, (x0, v0)) in enumerate(zip(cfgs, states)):
        e = engine.Engine(cfg)
        e.set_ions(x0, v0)
        e.run(seq_steps, engine.RUN_GRAPH)  # captures the graph the timed run replays
        t0 = time.perf_counter()
        e.run(seq_steps, engine.RUN_GRAPH)
        t_seq += (time.perf_counter() - t0) / seq_steps
        e.close()
    log("sweep sequential: %.2f ms per step of all simulations" % (t_seq * 1e3))
    kernel_ms = kms / max(nl, 1)
    return {
        "workload": "%d simulations, N log-uniform in [2^12, 2^17] (sum %d ions), M uniform in [0.5, 2], 1-4 dusts "
                    "on the axis, FP32c, %d steps each" % (n_sims, ions, sweep_len),
        "pairs_per_step": pairs,
        "batched_ms_per_step": t_batch * 1e3,
        "sequential_ms_per_step": t_seq * 1e3,
        "sims_per_hour_batched": n_sims / (sweep_len * t_batch / 3600.0),
        "sims_per_hour_sequential": n_sims / (sweep_len * t_seq / 3600.0),
        "batched_over_sequential": t_seq / t_batch,
        "pair_interactions_per_s_batched": pairs / t_batch,
        "batch_pair_kernel_ms": kernel_ms,
        "batch_pair_kernel_tflops_alg": ppl * FLOP_PER_PAIR / (kernel_ms * 1e-3) / 1e12,
        "timed_steps": {"batched": steps, "sequential_per_sim": seq_steps},
        "tuned_sym_chunk": chunk,
        "tuned_sym_tile": tile,
    }


def issue_probe():
    """Build and run opendust_amd/csrc/od_issue_probe.hip: the VALU issue
    rate of full-rate FMAs, of the transcendental ops alone and after FMAs,
    and of the pair term with and without its three transcendental ops (the
    measured ceiling of the pair kernels)."""
    import subprocess
    import tempfile

    from opendust_amd import build

    exe = os.path.join(tempfile.mkdtemp(prefix="od_probe_"), "od_issue_probe")
    csrc = os.path.join(ROOT, "opendust_amd", "csrc")
    cmd = [build.hipcc(), "--offload-arch=" + build.ARCH, "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang",
           "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", csrc, "-o", exe,
           os.path.join(csrc, "od_issue_probe.hip")]
    subprocess.run(cmd, check=True, capture_output=True)
    res = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    fr, tr = res["law_fullrate"], res["law_trans"]
    # 67 VALU instructions per loop iteration of 4 pair terms in both laws
    res["law_trans_slots_per_transcendental"] = 1.0 + (tr["ms"] / fr["ms"] - 1.0) * 67.0 / 12.0
    print(json.dumps(res))
    return res


def sweep_full(n_sims=64, sweep_len=5000, seq_steps=200, out_path=None):
    """The whole parameter sweep end to end, setup included: batched, every
    simulation's sweep_len steps (simulateBatch: trap fits, initial states,
    tables, batch, graph-replayed steps, records); sequential, the same
    simulations one after another (simulate-like: trap fit, initial state,
    context, graph-replayed steps, records), each timed over seq_steps steps
    and its step loop extrapolated to sweep_len (its setup measured)."""
    import contextlib
    import io

    from opendust_amd import engine, sweep, workloads

    def make():
        with contextlib.redirect_stdout(io.StringIO()):
            return workloads.sweep(n_sims=n_sims, n=sweep_len, precision="fp32c", seed=7, trapSeed=5,
                                   trapSamples=4_000_000)

    ods = make()
    np.random.seed(21)
    log("sweep end to end: batched, %d steps" % sweep_len)
    t0 = time.perf_counter()
    last = [t0]

    def progress(done, total):
        if time.perf_counter() - last[0] > 30.0 or done == total:
            last[0] = time.perf_counter()
            log("sweep end to end: batched, %d of %d steps, %.0f s" % (done, total, last[0] - t0))

    with contextlib.redirect_stdout(io.StringIO()):
        sweep.simulateBatch(ods, "0", useGraph=True, progress=progress)
    t_batch = time.perf_counter() - t0
    log("sweep end to end: batched %.1f s" % t_batch)
    ods = make()
    np.random.seed(21)
    t_seq, t_setup = 0.0, 0.0
    for i, od in enumerate(ods):
        t0 = time.perf_counter()
        cfg, x0, v0 = workloads.prepare(od, alpha=None, np_seed=None, precision="fp32c")
        e = engine.Engine(cfg)
        e.set_ions(x0, v0)
        t1 = time.perf_counter()
        e.run(seq_steps, engine.RUN_GRAPH)
        t2 = time.perf_counter()
        e.close()
        t_setup += t1 - t0
        t_seq += (t1 - t0) + (t2 - t1) * sweep_len / seq_steps
        if i % 16 == 15:
            log("sweep end to end: sequential, %d of %d simulations" % (i + 1, n_sims))
    res = {
        "workload": "%d simulations of the sweep generator (seed 7), %d steps each, FP32c" % (n_sims, sweep_len),
        "batched_wall_s": t_batch,
        "sequential_wall_s_extrapolated": t_seq,
        "sequential_setup_s": t_setup,
        "sequential_steps_timed_per_sim": seq_steps,
        "sims_per_hour_batched": n_sims / t_batch * 3600.0,
        "sims_per_hour_sequential": n_sims / t_seq * 3600.0,
        "batched_over_sequential": t_seq / t_batch,
    }
    print(json.dumps(res))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-ions", type=int, default=2 ** 20)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--pair-only", type=int, default=0)
    ap.add_argument("--pair-mode", choices=["sym", "rows", "auto"], default="sym")
    ap.add_argument("--cutoff-steps", type=int, default=3,
                    help="also time the cutoff workload (2^22 ions, R=10, H=20 r_D_e, r_c=2 r_D_e)")
    ap.add_argument("--sweep-steps", type=int, default=10,
                    help="also time the 64-simulation sweep, batched (this many steps) and sequential")
    ap.add_argument("--sweep-seq-steps", type=int, default=5)
    ap.add_argument("--sym-chunk", type=int, default=0,
                    help="source blocks per work-group of the symmetric kernel (0: autotune, or the default "
                         "with --no-tune)")
    ap.add_argument("--sym-tile", type=int, default=0,
                    help="symmetric kernel variant (1-based; tile shape and waves per SIMD); 0: autotune, or "
                         "the size-dependent default with --no-tune or --sym-chunk")
    ap.add_argument("--no-tune", action="store_true", help="skip the symmetric kernel's autotuner")
    ap.add_argument("--issue-probe", action="store_true",
                    help="only: the VALU issue-rate probe (opendust_amd/csrc/od_issue_probe.hip)")
    ap.add_argument("--sweep-full", action="store_true",
                    help="only: the whole 64-simulation, 5000-step sweep end to end, batched and sequential")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only: the 64-simulation sweep leg (--sweep-steps, --sweep-seq-steps), its JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this engine runs one simulation per GPU and this build measures one GPU only; "
              "--gpus %d is not supported (use --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.sweep_full:
        sweep_full()
        return
    if args.issue_probe:
        issue_probe()
        return
    if args.sweep_only:
        print(json.dumps(sweep_leg(args.sweep_steps, args.sweep_seq_steps)))
        return

    from opendust_amd import engine, workloads

    N = args.n_ions
    log("setup: single dust, N =", N)
    t0 = time.time()
    od = workloads.single_dust(N=N, n=args.steps + args.warmup, precision="fp32c", seed=12345,
                               trapSeed=2024)
    cfg, x0, v0 = workloads.prepare(od, alpha=None, np_seed=7, precision="fp32c")
    cfg.pair_mode = {"rows": engine.PAIR_ROWS, "sym": engine.PAIR_SYM, "auto": engine.PAIR_AUTO}[args.pair_mode]
    cfg.sym_chunk = args.sym_chunk
    cfg.sym_tile = args.sym_tile
    log("setup done in %.1f s (p_t=%.4f p_s=%.4f)" % (time.time() - t0, od.p_t, od.p_s))

    eng = engine.Engine(cfg)
    eng.set_ions(x0, v0)
    mode, pairs_launch = eng.pair_info()
    if not args.no_tune and not args.pair_only and not args.sym_chunk and not args.sym_tile:
        chunk, tile = eng.tune(2)
        log("autotuner chose %d source blocks per work-group, kernel variant %d %s"
            % (chunk, tile, engine.sym_variants()[tile - 1]))
    chunk, tile, bands = eng.sym_schedule()

    if args.pair_only:
        for _ in range(args.pair_only):
            eng.pair_forces(x0)
        log("pair-only launches done")
        return

    log("warmup", args.warmup)
    if args.warmup:
        eng.run(args.warmup, engine.RUN_PROFILE_PAIR)
    eng.pair_timing(reset=True)
    log("timed", args.steps, "steps")
    t0 = time.perf_counter()
    rec = eng.run(args.steps, engine.RUN_PROFILE_PAIR)  # returns after a stream synchronize
    wall = time.perf_counter() - t0
    pair_ms, launches = eng.pair_timing()
    x_mid, v_mid = eng.get_ions()
    q_mid = eng.get_dust_charges()
    # per-stage HIP-event times of a few more steps (outside the timed region)
    eng.run(3, engine.RUN_PROFILE_STAGES)
    stage_ms, _ = eng.stage_timing()
    flags_last, _, n_flagged = eng.step_info()
    n_absorbed_rows = int(np.count_nonzero(flags_last & ((1 << engine.MAX_DUST) - 1)))
    eng.close()
    log("stages (ms/step):", {k: round(v, 4) for k, v in stage_ms.items()})

    ms_per_step = wall / args.steps * 1e3
    pairs_step = float(N) * (N - 1)
    value = pairs_step * args.steps / wall
    kernel_ms = pair_ms / max(launches, 1)
    achieved = pairs_launch * FLOP_PER_PAIR / (kernel_ms * 1e-3) / 1e12
    traffic, traffic_src = (sym_traffic_2e20(tile, chunk, engine.sym_variants()[tile - 1][3])
                            if (mode == engine.PAIR_SYM and N == 2 ** 20 and tile) else (None, None))
    log("fp32c: %.2f ms/step, pair kernel %.2f ms, %.3e pairs/s" % (ms_per_step, kernel_ms, value))

    # FP64 GPU path and FP32c-vs-FP64 dust force error from the same state
    fp64 = None
    if args.fp64_steps > 0:
        step_idx = args.warmup + args.steps
        cfg64 = od.make_config(od.reinjection_tables(), od.alphaTrap, precision="fp64", seed=od.seed)
        res = {}
        for prec, c in (("fp64", cfg64), ("fp32c", cfg)):
            e = engine.Engine(c)
            e.set_ions(x_mid, v_mid)
            e.set_dust_charges(q_mid)
            e.set_step(step_idx)
            t0 = time.perf_counter()
            r = e.run(args.fp64_steps)
            res[prec] = (time.perf_counter() - t0, r)
            e.close()
        t64, r64 = res["fp64"]
        r32 = res["fp32c"][1]
        fo64 = r64["force_orbit"][0, 0]
        fo32 = r32["force_orbit"][0, 0]
        fp64 = {
            "steps": args.fp64_steps,
            "ms_per_step": t64 / args.fp64_steps * 1e3,
            "pair_interactions_per_s": pairs_step * args.fp64_steps / t64,
            "fp32c_speedup": value / (pairs_step * args.fp64_steps / t64),
            "dust_force_rel_err_first_step": float(np.linalg.norm(fo32 - fo64) / np.linalg.norm(fo64)),
        }
        log("fp64:", fp64)

    cut = None
    if args.cutoff_steps > 0:
        log("cutoff workload")
        cut = cutoff_leg(args.cutoff_steps)
        log("cutoff:", cut)

    swp = None
    if args.sweep_steps > 0:
        log("sweep workload")
        swp = sweep_leg(args.sweep_steps, args.sweep_seq_steps)
        log("sweep:", swp)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(cfg, x_mid)
        log("cpu:", cpu)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "steps_per_s": 1e3 / ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "single dust sphere, 2^20 superions, all-pairs ion-ion force, Maxwellian Ar+ flow M=1, "
                        "cylinder R=5 r_D_e H=10 r_D_e, consistent charge, dt=0.01/w_pi",
            "n_ions": N,
            "n_dust": 1,
            "pair_arithmetic": "f32 with compensated per-ion sums",
            "state_and_reductions": "f64",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            # profiled per (variant, schedule); the untuned default at 2^20 is variant 10 (row groups of
            # 4 target blocks) with 8 source blocks per work-group
            "traffic": traffic,
            "traffic_source": traffic_src,
            "traffic_unit": "bytes per launch",
            "kernel": "k_pair32_sym" if mode == engine.PAIR_SYM else "k_pair32<2>",
            "kernel_ms": kernel_ms,
            "flop_per_pair": FLOP_PER_PAIR,
            "pairs_per_launch": pairs_launch,
            "sym_chunk": chunk,
            "sym_tile": tile,
            "sym_bands": bands,
            # operand_form: 0 DPP/SGPR operands, 1 VGPR operands, 2 sources read from LDS
            "sym_tile_shape": (dict(zip(("targets_per_lane", "waves_per_simd", "operand_form",
                                         "target_blocks_per_work_group"),
                                        engine.sym_variants()[tile - 1]))
                               if tile else None),
        },
        "stage_ms": stage_ms,
        "hbm_stages": hbm_stages(stage_ms, N, STAGE_BYTES_PER_ION),
        "flagged_ions_last_step": n_flagged,
        # pair-term evaluations the kernels perform in a step (the metric counts
        # N(N-1) interactions once per step): the main launch (one evaluation per
        # unordered pair between blocks, symmetric kernel; one per ordered pair
        # otherwise), the in-block pairs, and the flagged ions' corrections (their
        # share as sources at the drifted and at the reinjected positions, and the
        # full rows of the absorbed and of the reinjected ions), last step's counts
        "kernel_pair_evaluations_per_step": pair_evaluations(N, mode == engine.PAIR_SYM, pairs_launch, n_flagged,
                                                             n_absorbed_rows),
        "cpu_baseline": cpu,
        "fp64_gpu": fp64,
        "cutoff_2e22": cut,
        "sweep_64": swp,
        "absorbed_last_step": int(rec["absorbed"][-1, 0]),
    }
    # targets of the workload (the dust-force accuracy is a correctness
    # requirement: the run fails when it is missed)
    dust_err = fp64["dust_force_rel_err_first_step"] if fp64 else None
    out["targets"] = {
        "fp32c_pair_kernel_frac_ge_0.60": out["roofline"]["frac"] >= 0.60,
        "fp32c_ge_5x_fp64": (fp64["fp32c_speedup"] >= 5.0) if fp64 else None,
        "dust_force_rel_err_le_1e-5": (dust_err <= 1e-5) if dust_err is not None else None,
        "sweep_ge_3x_sequential": (swp["batched_over_sequential"] >= 3.0) if swp else None,
    }
    print(json.dumps(out))
    if dust_err is not None and not dust_err <= 1e-5:
        log("dust force error %.3g against the FP64 path exceeds the 1e-5 target" % dust_err)
        sys.exit(1)


if __name__ == "__main__":
    main()
