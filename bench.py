"""Throughput of the two Unblending passes on one MI355X.

A step is one unmix + refine of one synthetic image resident in HBM (the
default workload: 4096 x 4096, 7 layers -- Normal, Multiply, Screen,
Overlay, ColorBurn, HardLight, SoftLight -- guided filter r = 32,
eps = 1e-3).  Prints one JSON line with Mpixel/s, the roofline of the
dominant kernel (the first-pass solve, FP64), the filter's HBM rate, and the
oracle timed on the host cores on a band of the same image.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def solve_flops(evals: int, iterations: int, L: int) -> float:
    """Algorithmic FP64 operations of the solve kernels (add/sub/mul/div/sqrt
    = 1, counted from the device source): an objective + gradient evaluation
    with its trial point and Armijo product costs 277 L - 168; a search
    direction (free set, two-loop over 4 pairs, update) 128 L."""
    return float(evals) * (277 * L - 168) + float(iterations) * (128 * L)


def filter_bytes(n: int, channels: int) -> float:
    """Algorithmic bytes of both filter passes: pass A reads 3 + C planes and
    writes 4C, pass B reads 4C + 3 and writes C, 8 bytes each."""
    return 8.0 * (6 + 10 * channels) * n


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(wl, rows, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ub_oracle as O
    band = np.ascontiguousarray(wl.rgb[:rows])
    layers = [O.Layer(l.mode, tuple(l.mean), tuple(np.ravel(l.precision)), l.comp) for l in wl.layers]
    t0 = time.perf_counter()
    first, _ = O.unmix(band, layers, opaque=True, threads=threads)
    O.refine(band, layers, first, opaque=True, radius=wl.radius, eps=wl.eps, threads=threads)
    dt = time.perf_counter() - t0
    n = band.shape[0] * band.shape[1]
    return {"value": n / dt / 1e6, "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "sample": "oracle (FP64 C restatement), unmix + refine of rows 0-%d of the same image "
                      "(%d x %d px), %d threads, %.1f s" % (rows - 1, band.shape[1], rows, threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--size", type=str, default="")
    ap.add_argument("--cpu-rows", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solver", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus)
        sys.exit(2)

    import torch
    import unblending_amd as U
    from unblending_amd import device, synth

    size = tuple(int(v) for v in args.size.split("x")) if args.size else None
    t0 = time.perf_counter()
    wl = synth.config(args.config, size)
    log("generated %s in %.1f s" % (wl.name, time.perf_counter() - t0))
    h, w = wl.height, wl.width
    L = len(wl.layers)
    n = w * h
    infos = [U.LayerInfo(U.CompOp(*l.comp), U.BlendMode(l.mode), U.GaussianColorModel(l.mean, l.precision))
             for l in wl.layers]

    dev = torch.device("cuda:0")
    rgb = torch.from_numpy(wl.rgb).to(dev)
    first = torch.empty((L, h, w, 4), dtype=torch.float64, device=dev)
    out = torch.empty_like(first)
    alpha = torch.empty((L, h, w), dtype=torch.float64, device=dev)
    ws1 = device.workspace(w, h, L, False, device=dev)
    ws2 = device.workspace(w, h, L, True, device=dev)
    params = U.hip_params(radius=wl.radius, eps=wl.eps, solver=args.solver)

    def step(p=params, s1=None, s2=None):
        device.unmix(rgb, infos, True, out=first, alpha_planes=alpha, ws=ws1, params=p, stats=s1)
        device.refine(rgb, first, infos, True, False, alpha_planes=alpha, out=out, ws=ws2, params=p, stats=s2)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log("warmup step %d done" % i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / args.steps
    value = n / (ms * 1e-3) / 1e6
    log("timed %d steps: %.2f ms/step" % (args.steps, ms))

    # one instrumented step: HIP events around each kernel on the launch stream
    tparams = U.hip_params(radius=wl.radius, eps=wl.eps, solver=args.solver, timing=True)
    s1, s2 = U.HipStats(), U.HipStats()
    step(tparams, s1, s2)
    torch.cuda.synchronize()
    flops1 = solve_flops(s1.evals, s1.iterations, L)
    achieved = flops1 / (s1.solve_ms * 1e-3) / 1e12
    fbytes = filter_bytes(n, s2.filter_channels)
    fgbs = fbytes / (s2.filter_ms * 1e-3) / 1e9 if s2.filter_ms > 0 else 0.0

    rec = {
        "metric": "Mpixel/s unmix+refine, plain vs MLP warm-start at equal error; filter HBM GB/s",
        "value": round(value, 4),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, unblending_amd/synth.py)",
        "config": {"workload": wl.name, "width": w, "height": h, "layers": L,
                   "modes": [synth.MODE_NAMES[l.mode] for l in wl.layers], "filter_radius": wl.radius,
                   "filter_eps": wl.eps, "warm_start": "none", "solver": "plain"},
        "roofline": {"bound": "valu", "kernel": "first-pass solve (FP64)", "achieved": round(achieved, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": None, "launch_ms": round(s1.solve_ms, 3),
                     "flops_per_launch": flops1},
        "filter": {"achieved": round(fgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(fgbs / HBM_PEAK_GBS, 4), "ms": round(s2.filter_ms, 3),
                   "coeff_ms": round(s2.filter_coeff_ms, 3), "out_ms": round(s2.filter_out_ms, 3),
                   "bytes": fbytes, "channels": s2.filter_channels},
        "solve": {"unmix_ms": round(s1.solve_ms, 3), "refine_ms": round(s2.solve_ms, 3),
                  "unmix_evals_per_px": s1.evals / n, "unmix_iters_per_px": s1.iterations / n,
                  "unmix_outer_per_px": s1.outer / n, "refine_evals_per_px": s2.evals / n,
                  "refine_tflops": round(solve_flops(s2.evals, s2.iterations, L) / (s2.solve_ms * 1e-3) / 1e12, 3)},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        rec["cpu_baseline"] = cpu_baseline(wl, args.cpu_rows, threads)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
