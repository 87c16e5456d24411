#!/usr/bin/env python3
"""Benchmark of the PLT hot path on one MI355X.

Workload (one "step"): one eval + pdf pass of the HIP BSDF kernel over a resident batch of
2^20 synthetic PLT BSDF queries (seed 0; 50/50 rough dielectric / rough Drude conductor,
K-correlation / Gaussian PSD roughness, coherence-aware beams, FP32). Prints one JSON line.

The CPU baseline is the oracle restatement (oracle/liboracle.so, OpenMP over the host cores)
timed on a bounded sample of the same batch.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
METRIC = "path samples/s @1080p depth 8; BSDF evals/s; surrogate vs exact at ≤1% rel err"
BYTES_PER_QUERY = 84  # 16 planes x 4 B in + 5 planes x 4 B out
HBM_PEAK_GBS = 8000.0
N_QUERIES = 1 << 20


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the timed region, run 5 plain device copies of MIB MiB (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters to bytes)")
    return ap.parse_args()


def traffic_from_profile():
    """HBM bytes per launch of the eval kernel from the committed rocprofv3 PMC summary, or None."""
    p = ROOT / "profiles" / "bsdf_eval_2p20_pmc.md"
    if not p.exists():
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(queries_np, budget_s: float):
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # the oracle is the CPU baseline (test infrastructure)
    from pltfalcor_amd.spectra import drude_profiles
    n_prof, k_prof = drude_profiles()
    prof = bytes(n_prof) + bytes(k_prof)
    threads = max(1, min(16, os.cpu_count() or 1))
    sample = queries_np.copy()
    n = sample.shape[1]
    O.c0_eval(sample, prof, 2, threads=threads)  # warm-up
    best, total, passes = float("inf"), 0.0, 0
    while total < budget_s or passes < 2:
        t0 = time.perf_counter()
        O.c0_eval(sample, prof, 2, threads=threads)
        dt = time.perf_counter() - t0
        best = min(best, dt)
        total += dt
        passes += 1
        if passes >= 20000:
            break
    return {"value": n / best, "unit": "BSDF evals/s", "cores": threads, "kind": "port",
            "sample": f"the same {n}-query batch, FP32 oracle, best of {passes} passes "
                      f"({total:.1f} s wall on {threads} threads)"}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus N>1 is not supported "
                          "(the path shards by pixels/queries, see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    from pltfalcor_amd.bsdf import BSDFBatch, default_context

    torch.cuda.set_device(0)
    ctx = default_context(0)
    batch = BSDFBatch(ctx, N_QUERIES).generate(seed=0)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        batch.eval(stream)
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        batch.eval(stream)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in ev]))

    value = N_QUERIES * args.steps / wall
    achieved = BYTES_PER_QUERY * N_QUERIES / (kern_ms * 1e-3) / 1e9
    traffic = traffic_from_profile()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "BSDF evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "PLT BSDF eval+pdf batch, 2^20 queries from seed 0, FP32 "
                               "(50/50 rough dielectric IOR U[1.3,1.8] / rough Drude conductor 64-bin n,k; "
                               "q U[0.01,0.3] um, l U[1,20] um, coherence area U[1,100] um^2, "
                               "cosine wi/wo, lambda U[380,780] nm, random Stokes)",
                   "queries": N_QUERIES, "kernel": "eval_queries_kernel"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_ms": kern_ms, "algorithmic_bytes_per_launch": BYTES_PER_QUERY * N_QUERIES},
        "cpu_baseline": None,
    }
    if args.copy_calibration > 0:
        nel = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = nel * 4
    if not args.no_cpu_baseline:
        q = batch.queries.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(q, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
