"""Benchmark of the Hermite force + predictor hot path on one MI355X.

Workload (6th order, N = 65536): Plummer sphere from seed 2 (Aarseth-Henon-
Wielen sampler, radii truncated at 22.8, Henon units, equal masses),
eps = 1/1024.  The state gets its stored derivatives from the two bootstrap
force passes of the reference's init_step (SRC/nbodysystem.h:573-588) on the
GPU.  One step is one pass of the hot path over the whole system with inputs
resident in HBM: predict_all(t) of all N j-particles followed by the force on
all N i-particles (N^2 interactions, the reference's "Init N square"
evaluation, SRC/nbodysystem.h:586-594), for the split-precision kernel
(headline value) and the plain FP64 kernel (reported beside it).

Prints one JSON line (see the repository's bench contract).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import hermite_amd as H  # noqa: E402

ORDER, N, SEED, EPS = 6, 65536, 2, 1.0 / 1024
PEAK_VALU_TFLOPS = 78.6  # MI355X FP64 vector (= unpacked FP32 issue rate)


def state(order=ORDER, n=N, seed=SEED, eps2=EPS * EPS):
    m, x, v = H.plummer(n, seed)
    P = H.make_particles(order, m, x, v)
    g = H.Gravity(order, n, "fp64")
    try:
        g.set_jp_all(P)
        for _ in range(2):  # bootstrap passes of init_step for order > 4
            g.predict_all(0.0)
            F = g.calc_force_on_first_nact(n, eps2)
            P[:, 10:10 + F.shape[1]] = F
            g.set_jp_all(P)
    finally:
        g.close()
    return P


def run_variant(P, variant, steps, warmup, eps2):
    n = len(P)
    g = H.Gravity(ORDER, n, variant)
    g.set_jp_all(P)
    g.predict_all(0.0)
    g.synchronize()
    dt = 1.0 / 1024
    for k in range(warmup):
        g.predict_async(k * dt)
        g.force_async(n, eps2)
    g.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        g.predict_async(k * dt)
        g.force_async(n, eps2)
    g.synchronize()
    wall = time.perf_counter() - t0
    geo = g.last_geometry()
    kernel_ms = g.time_force(n, eps2, reps=max(3, steps))
    predict_ms = g.time_predict(0.0, reps=max(3, steps))
    g.close()
    inter = float(n) * n
    return {
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "interactions_per_s": inter * steps / wall,
        "gflops": inter * steps * H.FLOPS[ORDER] / wall * 1e-9,
        "force_kernel_ms": kernel_ms,
        "predict_kernel_ms": predict_ms,
        "geometry": geo,
    }


def cpu_baseline(P, eps2, budget_s=10.0):
    """The reference's own AVX2 back end (SRC/hermite6-avx.h, built from its
    sources with its own flags into oracle/_ref) on the same state: the force
    on a bounded sample of i-particles against all N j-particles."""
    try:
        import oracle
        ref = oracle.RefAVX(ORDER)
        kind = "reference"
    except (ImportError, FileNotFoundError):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 64)
    threads = min(threads, 64)
    nact = 2048
    _, info = ref.forces(P, 0.0, nact, eps2, reps=1, threads=threads, want_output=False)
    per = max(info["seconds"], 1e-6)
    reps = max(1, min(200, int(budget_s / per)))
    _, info = ref.forces(P, 0.0, nact, eps2, reps=reps, threads=threads, want_output=False)
    inter = float(nact) * len(P) * reps
    cpu = platform.processor() or platform.machine()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                cpu = ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return {
        "value": inter / info["seconds"],
        "unit": "interactions/s",
        "cores": info["threads"],
        "kind": kind,
        "sample": f"6th-order force of the reference AVX2 back end on {nact} i x {len(P)} j "
                  f"(same N=65536 state), {reps} reps, {info['seconds']:.2f} s, {cpu}",
        "gflops": inter / info["seconds"] * H.FLOPS[ORDER] * 1e-9,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; multi-GPU sharding "
                                   "(j-split with one reduction per block step) is design-only"}))
        sys.exit(2)
    eps2 = EPS * EPS
    P = state()
    res = {v: run_variant(P, v, args.steps, args.warmup, eps2) for v in ("split", "fp64")}
    s, f = res["split"], res["fp64"]
    flops_per_launch = float(N) * N * H.FLOPS[ORDER]
    achieved = flops_per_launch / (s["force_kernel_ms"] * 1e-3) * 1e-12
    cpu = None if args.no_cpu else cpu_baseline(P, eps2)
    line = {
        "metric": "interactions/s (6th-order Hermite force+predictor, split precision)",
        "value": s["interactions_per_s"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (split)",
        "data": "synthetic (seeded Plummer sphere)",
        "config": {"workload": "6th-order Hermite, Plummer N=65536 seed 2, eps=1/1024, "
                               "predict_all + force on all N (N^2 interactions per step)",
                   "order": ORDER, "n": N, "nact": N, "parallelism": "single GPU"},
        "gflops": s["gflops"],
        "fp64": {"interactions_per_s": f["interactions_per_s"], "gflops": f["gflops"],
                 "ms_per_step": f["ms_per_step"], "force_kernel_ms": f["force_kernel_ms"]},
        "split_over_fp64": s["interactions_per_s"] / f["interactions_per_s"],
        "predict_kernel_ms": s["predict_kernel_ms"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_VALU_TFLOPS, "traffic": None,
                     "kernel": "tile_kernel<6,SPLIT>", "flops_per_interaction": H.FLOPS[ORDER],
                     "fp64_frac": flops_per_launch / (f["force_kernel_ms"] * 1e-3) * 1e-12 / PEAK_VALU_TFLOPS},
        "geometry": s["geometry"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
