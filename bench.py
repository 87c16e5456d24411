#!/usr/bin/env python3
"""Benchmark of the PLT path-tracing hot path on one MI355X. Prints one JSON line.

Default workload (one "step"): the PLTPT pass renders 4 frames (4 spp) of a 1920x1080 image at
max depth 8 of the procedural scene with 256 subdivision-5 icospheres (5,242,884 triangles) in 8
material kinds on a 20 x 20 ground quad under one 3 x 3 quad area light (6500 K blackbody), spectral
MIS over 4 hero wavelengths, NEE + MIS, Russian roulette. Scene, BVH, materials and bounce records
are resident in HBM before timing starts. `value` is path samples per second (pixels x spp / s).

The roofline figure is for the kernel that takes the most time (sample or solve), from HIP
events recorded around each kernel launch on the pass's stream during the timed region, with
algorithmic bytes counted from the bounce records of the timed frames:
  sample kernel: 84 B per traced path segment (ray 32 + hit 16 + 36 B bounce record written)
                 + 32 B per pixel (albedo and normal written);
  solve kernel:  36 B per path segment (its bounce record read once) + 16 B per pixel (colour).
Shadow rays, BVH nodes and triangles are not algorithmic bytes; rocprofv3's FETCH_SIZE /
WRITE_SIZE give the actual traffic (profiles/).

Each frame also runs the passes that follow PLTPT in the reference's render graph: the colour is
accumulated in FP64 and albedo and normal in FP32 (AccumulatePass), and the accumulated colour is
tone-mapped with auto exposure (ToneMapper); `secondary.graph_tail` gives their per-launch times.

`secondary.lights` renders the same workload with an importance-sampled environment map and rect,
disc, sphere and distant analytic lights added (every light type of the reference's light
sampling at once).

`secondary.bsdf_eval_batch` is the 2^20-query BSDF eval + pdf batch (BSDF evals/s), the second
part of the metric. The CPU baseline is the oracle restatement (C++, OpenMP over the host cores)
rendering a bounded sample of the same frame.

`--workload bsdf` makes the BSDF batch the primary line instead. `--workload surrogate` measures
the learned BSDF surrogate (an MLP on the matrix cores): it is trained on 2^24 exact samples of
the scattering lobes of the coated and grated materials and evaluated on 2^22 held-out queries;
`value` is surrogate evaluations per second (network + decode + exact pass for out-of-distribution
queries), the roofline is the fused inference kernel's BF16 matrix-core rate, and the line carries
the held-out error against the exact lobes and the exact lobes' own throughput (the A/B). The
default line carries a smaller surrogate run (2^22 training samples) as `secondary.surrogate`.
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
HBM_PEAK_GBS = 8000.0
BSDF_BYTES_PER_QUERY = 84  # 16 planes x 4 B in + 5 planes x 4 B out
N_QUERIES = 1 << 20
WIDTH, HEIGHT, SPP, MAX_BOUNCES = 1920, 1080, 4, 8
SAMPLE_BYTES_PER_SEGMENT, SAMPLE_BYTES_PER_PIXEL = 84, 32
SOLVE_BYTES_PER_SEGMENT, SOLVE_BYTES_PER_PIXEL = 36, 16
# split solve pass: a light event is 20 words as written for an NEE sample (meta, 4 radiances, 4
# wavelengths, position, direction, size, distance, type and light pdfs, walk pdf), a stored walk
# vertex 7 words (+ 7 words of walk origin / direction / base per event that has vertices, folded
# in here), and a region record 16 B per 64 paths
EVENT_BYTES, WALK_VERTEX_BYTES, REGION_BYTES = 80, 28, 16
# the staged unwind's sum kernel reads each event's meta word and stored RGB contribution
SUM_BYTES_PER_EVENT = 16
# a deferred manifold walk: its queue entry, the 12 event words it starts from (meta, vertex base,
# origin, direction, position, wavelength) and the 3 words of its last direction; its vertices are
# read once and written back once (2 x WALK_VERTEX_BYTES each)
WALK_BYTES = 4 + 48 + 12


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("path", "bsdf", "surrogate", "c4", "bvh"), default="path")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the BSDF-batch secondary measurement")
    ap.add_argument("--solve", choices=("megakernel", "split", "autotune"), default="autotune",
                    help="solve-pass schedule; autotune (default) times the candidate launch configurations on "
                         "the workload before the warm-up and keeps the fastest")
    ap.add_argument("--train-epochs", type=int, default=6,
                    help="training epochs of the surrogate (--workload surrogate / c4); 0 keeps the untrained network "
                         "with its statistics set (timing and counters of the inference kernel without the training)")
    ap.add_argument("--bvh", choices=("host", "device"), default="host",
                    help="the scene's BVH: the host builder's (binned SAH, made at scene creation) or rebuilt on "
                         "the GPU (PLOC)")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="override a PLTPT option of the path workload (repeatable), e.g. --option doMNEE=0 to see "
                         "what the manifold walks cost in the per-kernel times; the line then names the overrides")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the timed region, run 5 plain device copies of MIB MiB (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters to bytes)")
    return ap.parse_args()


def traffic_from_profile(name: str):
    """HBM bytes per launch of a kernel from its committed rocprofv3 PMC summary, or None."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def copy_calibration(mib: int):
    import torch
    nel = mib * (1 << 20) // 4
    src = torch.ones(nel, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    return nel * 4


# ------------------------------------------------------------------------------------ BSDF batch


def bsdf_batch(steps: int, warmup: int):
    import numpy as np
    import torch
    from pltfalcor_amd.bsdf import BSDFBatch, default_context
    batch = BSDFBatch(default_context(0), N_QUERIES).generate(seed=0)
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        batch.eval(stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        batch.eval(stream)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in ev]))
    achieved = BSDF_BYTES_PER_QUERY * N_QUERIES / (kern_ms * 1e-3) / 1e9
    return batch, {
        "value": N_QUERIES * steps / wall, "unit": "BSDF evals/s", "ms_per_step": wall * 1e3 / steps,
        "workload": "PLT BSDF eval+pdf batch, 2^20 queries from seed 0, FP32 (50/50 rough dielectric IOR "
                    "U[1.3,1.8] / rough Drude conductor 64-bin n,k; q U[0.01,0.3] um, l U[1,20] um, coherence "
                    "area U[1,100] um^2, cosine wi/wo, lambda U[380,780] nm, random Stokes)",
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile("bsdf_eval_2p20_pmc.md"),
                     "kernel": "eval_queries_kernel", "kernel_ms": kern_ms,
                     "algorithmic_bytes_per_launch": BSDF_BYTES_PER_QUERY * N_QUERIES}}


def bsdf_cpu_baseline(queries_np, budget_s: float):
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
    return {"value": n / best, "unit": "BSDF evals/s", "cores": threads, "kind": "port",
            "sample": f"the same {n}-query batch, FP32 oracle, best of {passes} passes "
                      f"({total:.1f} s wall on {threads} threads)"}


# ------------------------------------------------------------------------------------ surrogate

MFMA_BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_QUERY = 2 * (22 * 128 + 128 * 128 + 128 * 128 + 128 * 16)  # 75,264


def train_surrogate(ctx, n_train: int, epochs: int):
    """The surrogate trained on n_train exact samples: (surrogate, training record)."""
    import torch
    from pltfalcor_amd.surrogate import ExactSamples, Surrogate, train
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tr = ExactSamples(ctx, n_train, seed=21)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    s = Surrogate(ctx, max_batch=16384, seed=3)
    t0 = time.perf_counter()
    if epochs > 0:
        losses = train(s, tr, epochs=epochs, batch=16384, lr=2e-3)
    else:  # statistics only: the untrained network
        from pltfalcor_amd.surrogate import prepare
        prepare(s, tr)
        losses = [float("nan")]
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del tr
    return s, {"samples": n_train, "epochs": epochs, "seconds": train_s, "samples_per_s": epochs * n_train / train_s if epochs else None,
               "final_loss": losses[-1], "exact_sample_generation_s": gen_s}


def surrogate_run(ctx, s, training, n_test: int, reps: int, cpu_seconds: float = 0.0):
    import ctypes as C
    import numpy as np
    import torch
    from pltfalcor_amd.surrogate import ExactSamples
    te = ExactSamples(ctx, n_test, seed=22)
    n = te.n
    rng = np.random.default_rng(4)
    dop = rng.uniform(0, 1, n)
    v = rng.normal(0, 1, (n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    Sin = torch.tensor(np.c_[np.ones(n), 0.999 * dop[:, None] * v].astype(np.float32), device="cuda")
    for mode in (0, 1):
        s.eval(te, Sin, mode)
    torch.cuda.synchronize()
    walls = {}
    for mode in (0, 1):
        t0 = time.perf_counter()
        for _ in range(reps):
            out, fb = s.eval(te, Sin, mode)
        torch.cuda.synchronize()
        walls[mode] = (time.perf_counter() - t0) / reps
    # the fused inference kernel alone (raw outputs), HIP events on the launch stream
    Y = s.forward(te.X)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        s.lib.plt_surrogate_forward(s.handle, n, C.c_void_p(te.X.data_ptr()), C.c_void_p(Y.data_ptr()), None)
    e1.record()
    torch.cuda.synchronize()
    kern_ms = e0.elapsed_time(e1) / reps
    out, fb = s.eval(te, Sin, 0)
    ex, _ = s.eval(te, Sin, 1)
    o, e = out.cpu().numpy().astype(np.float64), ex.cpu().numpy().astype(np.float64)
    ok = e[:, 0] > 1e-6
    rel = np.abs(o[ok, 0] - e[ok, 0]) / e[ok, 0]
    relv = np.linalg.norm(o[ok] - e[ok], axis=1) / np.linalg.norm(e[ok], axis=1)
    fam = te.X.cpu().numpy()[:, 10:13].argmax(1)[ok]
    achieved = MLP_FLOP_PER_QUERY * n / (kern_ms * 1e-3) / 1e12
    res = {
        "value": n / walls[0], "unit": "surrogate BSDF evals/s",
        "exact_value": n / walls[1], "surrogate_over_exact": walls[1] / walls[0],
        "error_vs_exact": {"s0_rel_median": float(np.median(rel)), "s0_rel_mean": float(rel.mean()),
                           "s0_energy_weighted": float(np.abs(o[ok, 0] - e[ok, 0]).sum() / e[ok, 0].sum()),
                           "stokes_rel_median": float(np.median(relv)),
                           "s0_rel_median_per_family": {f: float(np.median(rel[fam == k]))
                                                         for k, f in enumerate(("coated opaque dielectric",
                                                                                "coated conductor",
                                                                                "grated conductor base"))},
                           "fallbacks": fb},
        "training": training,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None, "kernel": "mlp_infer_kernel",
                     "kernel_ms": kern_ms, "flop_per_query": MLP_FLOP_PER_QUERY, "queries_per_launch": n},
    }
    if cpu_seconds > 0:
        res["cpu_baseline"] = surrogate_cpu_baseline(s.get_params(), te.X[:1 << 15].cpu().numpy(), cpu_seconds)
    return res


C4_SPP = 16


def c4_run(ctx, sur, training):
    """The render with the learned surrogate in the loop (C4): the 1080p scene with the grating plate
    under a 5 mm glass slab, 16 spp (frames 0-15), depth 8, manifold NEE, rendered exactly and with
    the surrogate standing for the coated and grated scattering lobes (split solve pass both times),
    each accumulated in FP64. The Stokes vector per pixel comes from the camera polarizer at four
    angles (I = (S0 + cos t S1 + sin t S2) / 2 for t = 0, pi/2, pi, 3 pi/2), rendered the same way;
    the RMSE of each Stokes component is taken over pixels and RGB channels at matched seeds and given
    relative to the exact image's mean S0."""
    import numpy as np
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.postfx import AccumulatePass, AccumulatePrecision
    from pltfalcor_amd.scene import C4_GRATING, C4_SLAB, c4_scene
    scene = c4_scene(ctx, WIDTH / HEIGHT)
    opts = {"maxBounces": MAX_BOUNCES, "HWSS": 4}
    color = torch.empty((HEIGHT, WIDTH, 4), dtype=torch.float32, device="cuda")
    albedo, normal, out = torch.empty_like(color), torch.empty_like(color), torch.empty_like(color)
    acc = AccumulatePass(ctx, {"enabled": True, "precisionMode": AccumulatePrecision.Double})

    def render(use_sur: bool, theta=None):
        scene.set_camera(polarizer=0.0 if theta is None else 1.0, polarizer_theta=0.0 if theta is None else theta)
        p = PLTPT(ctx, opts)
        p.set_sample_pattern(PLTPT.HALTON, 16)
        p.set_launch_config(solve_schedule=PLTPT.SOLVE_SPLIT)
        p.set_scene(scene)
        if use_sur:
            p.set_surrogate(sur)
        p.execute(WIDTH, HEIGHT, color, albedo, normal)  # warm-up (allocations)
        p.set_frame(0)
        acc.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(C4_SPP):
            p.execute(WIDTH, HEIGHT, color, albedo, normal)
            acc.execute(color, out)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = p.surrogate_stats() if use_sur else None
        if use_sur and theta is None:
            # which inputs of the last frame's queries lie outside the training box, per input
            X = p.surrogate_queries()
            lo, hi = sur.stats[2], sur.stats[3]
            written = np.any(X != 0.0, axis=1)
            Xw = X[written]
            out_dim = (Xw < lo) | (Xw > hi)
            st["query_rows_written"] = int(Xw.shape[0])
            st["rows_outside_box"] = int(out_dim.any(axis=1).sum())
            st["rows_outside_box_per_input"] = {int(k): int(v) for k, v in enumerate(out_dim.sum(0)) if v}
        p.close()
        return out[..., :3].cpu().numpy().astype(np.float64), dt, st

    ex, t_ex, _ = render(False)
    su, t_su, st = render(True)
    pol = {}
    for k, th in enumerate((0.0, 0.5 * np.pi, np.pi, 1.5 * np.pi)):
        pol[(k, False)] = render(False, th)[0]
        pol[(k, True)] = render(True, th)[0]
    scene.set_camera(polarizer=0.0, polarizer_theta=0.0)
    stokes = {}
    for tag in (False, True):
        i0, i90, i180, i270 = (pol[(k, tag)] for k in range(4))
        stokes[tag] = (i0 + i180, i0 - i180, i90 - i270)
    s0_mean = float(stokes[False][0].mean())
    rmse = {f"S{k}": float(np.sqrt(np.mean((stokes[True][k] - stokes[False][k]) ** 2)) / s0_mean) for k in range(3)}
    # the same without the 0.1 % of pixels where the exact S0 is largest (fireflies of 16-spp images,
    # which dominate an RMSE), and the mean absolute error
    keep = stokes[False][0] <= np.quantile(stokes[False][0], 0.999)
    rmse_trim = {f"S{k}": float(np.sqrt(np.mean(((stokes[True][k] - stokes[False][k]) ** 2)[keep])) / s0_mean)
                 for k in range(3)}
    mae = {f"S{k}": float(np.mean(np.abs(stokes[True][k] - stokes[False][k]))) / s0_mean for k in range(3)}
    rmse_unpolarized = float(np.sqrt(np.mean((su - ex) ** 2)) / ex.mean())
    lum = lambda im: im @ np.array([0.2126, 0.7152, 0.0722])  # noqa: E731
    le, ls = lum(ex), lum(su)
    floor = 1e-3 * le.mean()
    rel = np.abs(ls - le) / (le + floor)
    acc.close()
    scene.close()
    eligible = st["rows"] // 4
    return {"workload": f"C4: the 1080p depth-8 scene with a {C4_GRATING['size']} m grated-conductor plate (pitch "
                        f"{C4_GRATING['pitch_um']} um, roughness {C4_GRATING['roughness']}) under a "
                        f"{C4_SLAB['thickness'] * 1e3:.0f} mm glass slab (IoR {C4_SLAB['ior']}), {C4_SPP} spp, "
                        "manifold NEE, split solve pass; surrogate for the coated and grated scattering lobes with "
                        "exact fallback outside its training box",
            "exact_path_samples_per_s": WIDTH * HEIGHT * C4_SPP / t_ex,
            "surrogate_path_samples_per_s": WIDTH * HEIGHT * C4_SPP / t_su,
            "surrogate_over_exact": t_ex / t_su,
            "stokes_rmse_rel_to_mean_s0": rmse, "stokes_rmse_rel_without_top_0.1pct_s0": rmse_trim,
            "stokes_mae_rel_to_mean_s0": mae, "unpolarized_rmse_rel_to_mean": rmse_unpolarized,
            "luminance_rel_err": {"median": float(np.median(rel)), "mean": float(rel.mean()),
                                  "image_mean": float(abs(ls.mean() - le.mean()) / le.mean())},
            "network_share_of_eligible_steps": 1.0 - st["exact_steps"] / max(1, eligible),
            "queries_outside_training_box": {k: st[k] for k in ("query_rows_written", "rows_outside_box",
                                                                "rows_outside_box_per_input")},
            "eligible_steps_per_frame": eligible, "training": training}


def surrogate_cpu_baseline(params, X, budget_s: float):
    sys.path.insert(0, str(ROOT / "tests"))
    import mlp_ref  # the CPU restatement of the network (test infrastructure)
    n, total, passes = X.shape[0], 0.0, 0
    while total < budget_s or passes < 2:
        t0 = time.perf_counter()
        mlp_ref.forward(params, X)
        total += time.perf_counter() - t0
        passes += 1
    return {"value": n * passes / total, "unit": "surrogate BSDF evals/s", "cores": os.cpu_count() or 1, "kind": "port",
            "sample": f"{n} queries x {passes} passes of the restated network (numpy float64, BLAS threads)"}


# ------------------------------------------------------------------------------------ path tracing


def segment_count(pltpt, frames, split: bool):
    """Traced path segments (bounce records written) per frame, read back from the pass, and the
    split solve pass's event counts per frame."""
    import numpy as np
    counts, stats = [], []
    for f in frames:
        pltpt.set_frame(f)
        pltpt.execute_frame()
        rec = pltpt.bounce_records(WIDTH, HEIGHT)
        flags = rec[:, 6, :]
        end = (flags & 0x30) != 0
        counts.append(int((np.argmax(end, axis=0) + 1).sum()))
        del rec, flags, end
        if split:
            stats.append(pltpt.solve_stats())
    return counts, stats


def lights_run(ctx, opts, steps: int, launch=None):
    """The default workload with the other light types of the reference's light sampling added: a
    256 x 128 sky environment map (importance-sampled) and four analytic lights (rect, disc,
    sphere, distant) beside the emissive quad. Returns path samples/s and per-frame kernel times."""
    import numpy as np
    import torch
    from pltfalcor_amd.lights import DiscLight, DistantLight, RectLight, SphereLight
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import PATH_2, Scene
    scene = Scene(ctx, aspect=WIDTH / HEIGHT, **PATH_2)
    v = np.linspace(0, 1, 128, dtype=np.float32)[:, None, None]
    sky = np.broadcast_to(np.concatenate([0.2 + 0.8 * (1 - v), 0.4 + 0.6 * (1 - v), np.ones_like(v)], 2), (128, 256, 3))
    sky = np.array(sky)
    sky[10:20, 100:120] = (40.0, 35.0, 30.0)  # a sun-like patch
    scene.set_env_map(sky, rotation=(0.0, 20.0, 0.0), intensity=0.3)
    emission = 2  # the scene's blackbody emission profile
    rect = RectLight()
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = [[1, 0, 0], [0, 0, -1], [0, 1, 0]]
    m[:3, 3] = (-4.0, 6.0, 3.0)
    rect.transform = m
    disc = DiscLight()
    m2 = m.copy()
    m2[:3, 3] = (4.0, 6.0, -3.0)
    disc.transform = m2
    sph = SphereLight()
    m3 = np.eye(4, dtype=np.float32)
    m3[:3, 3] = (0.0, 5.0, 5.0)
    sph.transform = m3
    sph.scaling = (0.5, 0.5, 0.5)
    sun = DistantLight()
    sun.direction = (0.4, -1.0, 0.3)
    for light in (rect, disc, sph, sun):
        light.setIntensity(emission)
    scene.set_lights([rect, disc, sph, sun])
    p = PLTPT(ctx, opts)
    p.set_sample_pattern(PLTPT.HALTON, 16)
    if launch:
        p.set_launch_config(**launch)
    p.set_scene(scene)
    color = torch.empty((HEIGHT, WIDTH, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    for _ in range(2):
        p.execute(WIDTH, HEIGHT, color, albedo, normal)
    torch.cuda.synchronize()
    p.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(steps * SPP):
        p.execute(WIDTH, HEIGHT, color, albedo, normal)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sample_ms, solve_ms, executes = p.kernel_times()
    p.close()
    scene.close()
    return {"value": WIDTH * HEIGHT * steps * SPP / wall, "unit": "path samples/s",
            "workload": "the default workload plus a 256x128 importance-sampled sky and rect, disc, sphere and "
                        "distant lights (light type chosen per NEE sample among environment, emissive, analytic), "
                        "same launch configuration",
            "sample_kernel_ms": sample_ms / executes, "solve_pass_ms": solve_ms / executes}


def postfx_timing(acc_double, acc_single, tm, color, albedo, out, ldr, reps: int):
    """Per-launch time of the passes after PLTPT at 1080p, from HIP events on the stream they run
    on: FP64 accumulation (16 B in + 32 B sum read + 32 B sum written + 16 B out per pixel), FP32
    accumulation (16 + 16 + 16 + 16 B) and the auto-exposure tone mapper (the luminance pass reads
    each source texel about once, 16 B, and the tone-mapping pass reads and writes 16 B each)."""
    import torch
    pixels = color.shape[0] * color.shape[1]
    res = {}
    for name, fn, nbytes in (("accumulate_double", lambda: acc_double.execute(color, out), 96 * pixels),
                             ("accumulate_single", lambda: acc_single.execute(albedo, out), 64 * pixels),
                             ("tonemap_auto_exposure", lambda: tm.execute(out, ldr), 48 * pixels)):
        for _ in range(5):
            fn()
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": ms, "algorithmic_bytes": nbytes, "GB/s": nbytes / (ms * 1e-3) / 1e9,
                     "frac_hbm_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return res


def path_cpu_baseline(scene_np, budget_s: float, opts):
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # the oracle is the CPU baseline (test infrastructure)
    threads = max(1, min(16, os.cpu_count() or 1))
    osc = O.OracleScene(scene_np)
    kw = dict(max_bounces=MAX_BOUNCES, threads=threads, do_rr=opts["doRussianRoulette"], do_mnee=opts["doMNEE"],
              mnee_max_occluders=opts["MNEEMaxOccluders"], mnee_max_iterations=opts["MNEEMaxIterations"],
              mnee_threshold=opts["MNEESolverThreshold"])
    # chunks of 64 rows in frame order, frames 0..SPP-1 in turn, until the budget is spent
    osc.render(WIDTH, HEIGHT, frame=0, first=WIDTH * (HEIGHT // 2), count=WIDTH, **kw)  # warm-up
    chunk = 64 * WIDTH
    total, paths, frame, first = 0.0, 0, 0, 0
    while total < budget_s:
        n = min(chunk, WIDTH * HEIGHT - first)
        t0 = time.perf_counter()
        osc.render(WIDTH, HEIGHT, frame=frame, first=first, count=n, **kw)
        total += time.perf_counter() - t0
        paths += n
        first += n
        if first >= WIDTH * HEIGHT:
            first, frame = 0, (frame + 1) % SPP
    osc.close()
    return {"value": paths / total, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"{paths} paths of the same workload rendered in 64-row chunks in frame order from frame 0 "
                      f"({paths / (WIDTH * HEIGHT):.2f} frames, {total:.1f} s on {threads} threads; FP32 oracle, "
                      "BVH prebuilt)"}


def bvh_run(ctx, steps: int):
    """The path workload scene's BVH (5,242,884 triangles): the host builder's binned SAH (made at
    scene creation) against the device builder (PLOC + 4-wide collapse), timed over `steps` rebuilds
    after one warm-up build, and 2^20 camera rays cast against each tree (hits must be identical)."""
    import numpy as np
    import torch
    from pltfalcor_amd.scene import PATH_2, Scene
    scene = Scene(ctx, aspect=WIDTH / HEIGHT, **PATH_2)
    host = scene.bvh_info()
    n = 1 << 20
    rng = np.random.default_rng(5)
    cam = np.array([0, 5, 16], np.float32)
    tgt = np.stack([rng.uniform(-10, 10, n), rng.uniform(0, 2, n), rng.uniform(-10, 10, n)], 1).astype(np.float32)
    d = tgt - cam
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = torch.from_numpy(np.ascontiguousarray(np.concatenate(
        [np.tile(cam, (n, 1)).T, d.T, np.zeros((1, n), np.float32), np.full((1, n), 3.4e38, np.float32)]),
        dtype=np.float32)).cuda()

    def cast():
        hits = scene.trace(rays)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(10):
            scene.trace(rays)
        torch.cuda.synchronize()
        return hits.cpu().numpy(), (time.perf_counter() - t) / 10 * 1e3

    host_hits, host_cast = cast()
    scene.build_bvh("device")
    wall, build = [], []
    for _ in range(max(1, steps)):
        t = time.perf_counter()
        build.append(scene.build_bvh("device")["build_ms"])
        wall.append((time.perf_counter() - t) * 1e3)
    dev = scene.bvh_info()
    dev_hits, dev_cast = cast()
    if not np.array_equal(host_hits, dev_hits):
        raise RuntimeError("the device tree's hits differ from the host tree's")
    ms = float(np.mean(build))
    scene.close()
    tris = scene_triangles(PATH_2)
    return {"value": tris / (ms * 1e-3), "triangles": tris, "ms": ms, "wall_ms": float(np.mean(wall)),
            "host_build_ms": host["build_ms"], "host": host, "device": dev,
            "camera_rays_cast_ms": {"host_tree": host_cast, "device_tree": dev_cast}}


def scene_triangles(desc) -> int:
    # ground quad (2), 20 * 4^s per icosphere, the light quad (2)
    return 4 + desc["sphere_count"] * 20 * 4 ** desc["subdivision"]


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus N>1 is not supported "
                          "(paths and queries would shard by pixel/query range across GPUs)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    torch.cuda.set_device(0)

    if args.workload == "bsdf":
        batch, r = bsdf_batch(args.steps, args.warmup)
        out = {"metric": METRIC, "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": r["workload"], "queries": N_QUERIES}, "roofline": r["roofline"],
               "cpu_baseline": None}
        if args.copy_calibration > 0:
            out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration)
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = bsdf_cpu_baseline(batch.queries.cpu().numpy(), args.cpu_seconds)
        print(json.dumps(out), flush=True)
        return

    if args.workload == "bvh":
        from pltfalcor_amd.bsdf import default_context
        r = bvh_run(default_context(0), args.steps)
        out = {"metric": METRIC, "value": r["value"], "unit": "triangles/s (BVH build)", "n_gpus": 1,
               "steps": args.steps, "warmup": 1, "ms_per_step": r["ms"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": "BVH build of the path workload scene (seed 1, 256 icospheres subdiv 5, "
                                      "5,242,884 triangles): device PLOC clustering + 4-wide collapse, against the "
                                      "host binned-SAH build",
                          "triangles": r["triangles"]},
               "roofline": None, "cpu_baseline": None, "bvh": r}
        print(json.dumps(out), flush=True)
        return

    if args.workload in ("surrogate", "c4"):
        from pltfalcor_amd.bsdf import default_context
        sctx = default_context(0)
        trained, training = train_surrogate(sctx, 1 << 24, args.train_epochs)
    if args.workload == "c4":
        r = c4_run(sctx, trained, training)
        trained.close()
        out = {"metric": METRIC, "value": r["surrogate_path_samples_per_s"], "unit": "path samples/s", "n_gpus": 1,
               "steps": 1, "warmup": 1, "ms_per_step": WIDTH * HEIGHT * C4_SPP / r["surrogate_path_samples_per_s"] * 1e3,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32 transport, bf16 surrogate (FP32 accumulation)", "data": "synthetic",
               "config": {"workload": r["workload"], "width": WIDTH, "height": HEIGHT, "spp": C4_SPP,
                          "max_bounces": MAX_BOUNCES},
               "roofline": None, "cpu_baseline": None, "c4": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "surrogate":
        r = surrogate_run(sctx, trained, training, 1 << 22, max(3, args.steps), 0.0 if args.no_cpu_baseline else args.cpu_seconds)
        trained.close()
        out = {"metric": METRIC, "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": max(3, args.steps),
               "warmup": 1, "ms_per_step": (1 << 22) / r["value"] * 1e3, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "bf16 (FP32 accumulation)", "data": "synthetic",
               "config": {"workload": "learned BSDF surrogate: MLP 22-128-128-128-16 ReLU trained on 2^24 exact "
                                      f"samples ({args.train_epochs} epochs, batch 16384, Adam) of the scattering lobes "
                                      "of the coated opaque dielectric, coated conductor and grated conductor; evaluated "
                                      "on 2^22 held-out queries with random partially polarised incident Stokes vectors",
                          "queries": 1 << 22},
               "roofline": r["roofline"], "cpu_baseline": r.get("cpu_baseline"),
               "surrogate": {k: r[k] for k in ("exact_value", "surrogate_over_exact", "error_vs_exact", "training")}}
        print(json.dumps(out), flush=True)
        return

    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.postfx import AccumulatePass, AccumulatePrecision, ToneMapper
    from pltfalcor_amd.scene import PATH_2, Scene

    ctx = default_context(0)
    t_build = time.perf_counter()
    scene = Scene(ctx, aspect=WIDTH / HEIGHT, **PATH_2)
    t_build = time.perf_counter() - t_build
    bvh = {"host": scene.bvh_info()}
    if args.bvh == "device":
        scene.build_bvh("device")
        bvh["device"] = scene.bvh_info()
    bvh["used"] = args.bvh
    # the reference's PLTPathTracer graph creates PLTPT with its defaults (manifold NEE on, 2 occluders,
    # 60 Newton iterations, threshold 5e-5); only the depth and the HWSS count are set here
    opts = {"maxBounces": MAX_BOUNCES, "HWSS": 4}
    overrides = {}
    for kv in args.option:
        k, _, v = kv.partition("=")
        overrides[k] = float(v)
    opts.update(overrides)
    pltpt = PLTPT(ctx, opts)
    pltpt.set_sample_pattern(PLTPT.HALTON, 16)  # the graph's GBufferRT jitter (samplePattern Halton, 16)
    pltpt.set_scene(scene)
    dev = torch.device("cuda")
    color = torch.empty((HEIGHT, WIDTH, 4), dtype=torch.float32, device=dev)
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    pltpt.execute_frame = lambda: pltpt.execute(WIDTH, HEIGHT, color, albedo, normal)  # noqa: E731
    # the passes after PLTPT in the reference graph: colour accumulated in FP64, albedo and normal in
    # FP32, tone mapping with auto exposure (the OptiX denoiser between them is not part of this build)
    acc_c = AccumulatePass(ctx, {"enabled": True, "precisionMode": AccumulatePrecision.Double})
    acc_a = AccumulatePass(ctx, {"enabled": True, "precisionMode": AccumulatePrecision.Single})
    acc_n = AccumulatePass(ctx, {"enabled": True, "precisionMode": AccumulatePrecision.Single})
    tm = ToneMapper(ctx, {"autoExposure": True, "exposureCompensation": 0.0})
    out_c, out_a, out_n, ldr = (torch.empty_like(color) for _ in range(4))

    def graph_tail():
        acc_c.execute(color, out_c)
        acc_a.execute(albedo, out_a)
        acc_n.execute(normal, out_n)
        tm.execute(out_c, ldr)

    def step():
        pltpt.set_frame(0)  # one step = frames 0..SPP-1 (SPP samples per pixel)
        for p in (acc_c, acc_a, acc_n):
            p.reset()
        for _ in range(SPP):
            pltpt.execute_frame()
            graph_tail()

    # launch configuration: the autotuner (before the warm-up, outside the timed region) times the
    # sample-pass schedules and block sizes, the solve megakernel against the split solve pass at
    # several event-area sizes, and graph replay, on four frames of this workload
    tune = None
    if args.solve == "split":
        pltpt.set_launch_config(solve_schedule=PLTPT.SOLVE_SPLIT)
    elif args.solve == "autotune":
        t0 = time.perf_counter()
        best, d_ms, b_ms = pltpt.autotune(WIDTH, HEIGHT, color, albedo, normal, frames=4)
        tune = {"chosen": best, "default_ms_per_frame": d_ms, "best_ms_per_frame": b_ms,
                "seconds": round(time.perf_counter() - t0, 3)}
    launch = pltpt.launch_config()
    split = launch["solve_schedule"] == PLTPT.SOLVE_SPLIT

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    pltpt.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sample_ms, light_ms, walk_ms, unwind_ms, mega_ms, executes = pltpt.kernel_times_split()
    pltpt.set_profiling(False)

    segs, sstats = segment_count(pltpt, range(SPP), split)  # outside the timed region
    pixels = WIDTH * HEIGHT
    regions = (pixels + 63) // 64
    seg_per_exec = sum(segs) / SPP
    if split and not launch["graph"]:
        avg = {k: v / executes for k, v in (("sample", sample_ms), ("light", light_ms), ("walk", walk_ms),
                                            ("unwind", unwind_ms), ("megakernel", mega_ms))}
    else:  # one solve launch sequence, timed as a whole
        avg = {"sample": sample_ms / executes, "megakernel": (light_ms + walk_ms + unwind_ms + mega_ms) / executes}
    per_frame = {k: sum(st[k] for st in sstats) / SPP for k in ("events", "walk_vertices", "event_vertices",
                                                                 "overflowed_paths", "walks", "failed_walks",
                                                                 "walk_rays")} if split else None
    # algorithmic bytes per launch of each kernel
    kb = {"sample": SAMPLE_BYTES_PER_SEGMENT * seg_per_exec + SAMPLE_BYTES_PER_PIXEL * pixels}
    if split:
        ev_b = EVENT_BYTES * per_frame["events"] + WALK_VERTEX_BYTES * per_frame["walk_vertices"] + REGION_BYTES * regions
        kb["light"] = SOLVE_BYTES_PER_SEGMENT * seg_per_exec + ev_b
        kb["walk"] = WALK_BYTES * per_frame["walks"] + 2 * WALK_VERTEX_BYTES * per_frame["walk_vertices"]
        # the rest of the unwind (the part beside the walk kernel is in neither figure's time)
        kb["unwind"] = SUM_BYTES_PER_EVENT * per_frame["events"] + REGION_BYTES * regions + SOLVE_BYTES_PER_PIXEL * pixels
        kb["megakernel"] = REGION_BYTES * regions  # the fallback: its region check, plus overflowed paths' work
    else:
        kb["megakernel"] = SOLVE_BYTES_PER_SEGMENT * seg_per_exec + SOLVE_BYTES_PER_PIXEL * pixels
    knames = {"sample": "sample_kernel", "light": "solve_sample_kernel<4, false>",
              "walk": "solve_walk_kernel (the non-walk events' solve_unwind_kernel<4> runs beside it on a second "
                      "stream)",
              "unwind": "solve_walk_unwind_kernel<4>, the join with the second stream, solve_sum_kernel",
              "megakernel": "solve_kernel<4, false>" + (" (fallback of the split pass)" if split and not launch["graph"] else "")}
    if split and launch["graph"]:
        knames["megakernel"] = "split solve pass replayed from a graph (light + walk + unwind + fallback)"
        kb["megakernel"] = SOLVE_BYTES_PER_SEGMENT * seg_per_exec + SOLVE_BYTES_PER_PIXEL * pixels
    profiles = {"sample": "pltpt_sample_1080p_d8_pmc.md", "light": "pltpt_solve_light_1080p_d8_pmc.md",
                "walk": "pltpt_solve_walk_1080p_d8_pmc.md", "unwind": None,
                "megakernel": None if split else "pltpt_solve_1080p_d8_pmc.md"}
    kdom = max(avg, key=lambda k: avg[k])
    kname, kms, kbytes = knames[kdom], avg[kdom], kb[kdom]
    traffic = traffic_from_profile(profiles[kdom]) if profiles[kdom] else None
    achieved = kbytes / (kms * 1e-3) / 1e9

    post = postfx_timing(acc_c, acc_a, tm, color, albedo, out_c, ldr, 50)
    lights = None if args.no_secondary else lights_run(ctx, opts, 2, launch)
    bsdf = None if args.no_secondary else bsdf_batch(500, 20)[1]
    sur, c4 = None, None
    if not args.no_secondary:
        from pltfalcor_amd.bsdf import default_context as _dc
        sctx = _dc(0)  # the surrogate's context: the Drude profiles its exact samples use
        trained, training = train_surrogate(sctx, 1 << 24, 6)
        sur = surrogate_run(sctx, trained, training, 1 << 22, 5)
        c4 = c4_run(ctx, trained, training)
        trained.close()
    out = {
        "metric": METRIC,
        "value": pixels * SPP * args.steps / wall,
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"PLTPT {WIDTH}x{HEIGHT}, {SPP} spp (frames 0-{SPP - 1}) per step, max depth "
                               f"{MAX_BOUNCES}, procedural scene seed 1: 20x20 ground + 256 icospheres subdiv 5 "
                               "(5,242,884 triangles), one 3x3 quad light (6500 K blackbody), 8 materials uniform "
                               "over dielectric / Drude conductor / coated opaque dielectric / grated conductor / "
                               "birefringent thin dielectric / 20-layer stack (FP64 steps) / coated conductor / "
                               "opaque dielectric; HWSS 4 + spectral MIS, NEE + MIS, RR; manifold NEE on (<= 2 "
                               "specular occluders, <= 60 Newton iterations, threshold 5e-5); Halton-16 camera "
                               "jitter; FP64/FP32 accumulation and auto-exposure tone mapping per frame",
                   "width": WIDTH, "height": HEIGHT, "spp": SPP, "max_bounces": MAX_BOUNCES,
                   "triangles": int(scene.arrays.n_triangles), "scene_build_s": round(t_build, 3),
                   "segments_per_frame": seg_per_exec, "option_overrides": overrides or None, "bvh": bvh},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kname, "kernel_ms": kms,
                     "algorithmic_bytes_per_launch": kbytes,
                     "kernels": {knames[k]: {"ms": avg[k], "algorithmic_bytes": kb[k],
                                             "GB/s": kb[k] / (avg[k] * 1e-3) / 1e9 if avg[k] > 0 else None}
                                 for k in avg},
                     "walk_rays_per_s": per_frame["walk_rays"] / (avg["walk"] * 1e-3) if split and avg.get("walk") else None,
                     "sample_kernel_ms": avg["sample"], "solve_pass_ms": sum(v for k, v in avg.items() if k != "sample")},
        "launch": {"config": launch, "autotune": tune, "split_solve_per_frame": per_frame},
        "cpu_baseline": None,
        "secondary": {"bsdf_eval_batch": {k: bsdf[k] for k in ("value", "unit", "roofline")},
                      "surrogate": {k: sur[k] for k in ("value", "unit", "exact_value", "surrogate_over_exact",
                                                        "error_vs_exact", "training", "roofline")},
                      "graph_tail": post, "lights": lights, "c4": c4}
                     if bsdf else {"graph_tail": post},
    }
    if args.copy_calibration > 0:
        out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = path_cpu_baseline(scene.numpy(), args.cpu_seconds, pltpt.get_scripting_dictionary())
    for p in (acc_c, acc_a, acc_n, tm, pltpt):
        p.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
