"""Benchmark of the Hermite force + predictor hot path on one MI355X.

Workload: 6th-order Hermite, Plummer sphere N = 65536 from seed 2 (Aarseth-
Henon-Wielen sampler, radii truncated at 22.8, Henon units, equal masses),
eps = 1/1024, eta = 0.2 (eta_s = 0.1 for the first steps), power-of-two
block timesteps with dt_max = 1/16.  init_step (SRC/nbodysystem.h:565-631)
runs once on the GPU through the host driver; the block-step integration
then runs GPU-resident (schedule, predict_all, force on the active list,
corrector and timestep on the device).  The timed steps are launched
directly on one stream with a HIP event pair around every force launch;
--graph replays them from a hipGraph instead, without the events.

One step is one dt_max interval of integration (the reference's
integrate_one_dtmax, SRC/nbodysystem.h:904-944): the timed region is
steps x 1/16 time units from t = 0 (default 16 steps: t = 0 -> 1), started
and ended by a stream synchronisation inside hhr_run.  Warmup steps run on a
separate context that is discarded, so the timed run starts from t = 0.

value = interactions/s of the plain FP64 kernel (num_step * N / wall), with
|dE/E| at the end of the interval.  Beside it: the split-precision kernel's
integration outcome, both kernels' all-pairs force rate (the reference's
"Init N square" evaluation), the split kernel's force error against the FP64
kernel, the force kernel's roofline (HIP events around every force launch of
the timed run) and the reference's own AVX2 back end and driver, built from
its sources, timed on the host cores (cpu_baseline).

Prints one JSON line on stdout; progress goes to stderr.
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

METRIC = "interactions/s and Gflops (60/97/144 flop conv.), split-prec vs FP64 at stated |ΔE/E|"
ORDER, N, SEED, EPS, ETA, ETA_S, DTMAX = 6, 65536, 2, 1.0 / 1024, 0.2, 0.1, 1.0 / 16
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak (also the unpacked FP32 issue rate)
BATCH = 16               # block steps per graph launch
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "r_force_kernel_6th_n65536_traffic.txt")


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def initial_state(eps2):
    """Particle records after init_step on the GPU (FP64 kernel)."""
    m, x, v = H.plummer(N, SEED)
    P = H.make_particles(ORDER, m, x, v)
    s = H.System(ORDER, P, 0.0, eps2, ETA, ETA_S, DTMAX, variant="fp64")
    try:
        s.init_step()
        return s.ptcl(), s.init_energy, P
    finally:
        s.close()


def integrate(variant, P0, e0, eps2, tend, timing, graph=False):
    r = H.Resident(ORDER, variant, [(P0, eps2, ETA, DTMAX, 0.0, 0.0)])
    try:
        if timing:
            r.set_timing(True)
        t0 = time.perf_counter()
        err = None
        try:
            r.run(tend, use_graph=graph, batch=BATCH)
        except H.HHError as ex:
            err = str(ex)
        wall = time.perf_counter() - t0
        st = r.stats()
        out = {"wall_s": wall, "num_step": st["num_step"], "num_bstep": st["num_bstep"], "tsys": st["tsys"],
               "error": err}
        if err is None:
            out["dE_E"] = (r.energy() - e0) / e0
        if timing:
            out["force_ms"], out["force_launches"] = r.force_timing()
        out["steps_launched"] = r.steps_launched
        return out
    finally:
        r.close()


def all_pairs(P0, eps2):
    """Force on all N from all N (inputs resident) for both kernels, and the
    split kernel's per-particle relative error against the FP64 kernel."""
    res, F = {}, {}
    for variant in ("fp64", "split"):
        g = H.Gravity(ORDER, N, variant)
        try:
            g.set_jp_all(P0)
            g.predict_all(0.0)
            g.time_force(N, eps2, 2)
            ms = g.time_force(N, eps2, 20)
            F[variant] = g.calc_force_on_first_nact(N, eps2)
            geo = g.last_geometry()
        finally:
            g.close()
        inter = float(N) * N
        res[variant] = {"force_ms": ms, "interactions_per_s": inter / (ms * 1e-3),
                        "gflops": inter * H.FLOPS[ORDER] / (ms * 1e-3) * 1e-9, "geometry": geo}
        log(f"all-pairs {variant}: {ms:.3f} ms  {res[variant]['interactions_per_s']:.3e} int/s")
    err = {}
    for k, name in enumerate(("acc", "jrk", "snp")):
        d = np.linalg.norm(F["split"][:, 3 * k:3 * k + 3] - F["fp64"][:, 3 * k:3 * k + 3], axis=1)
        ref = np.linalg.norm(F["fp64"][:, 3 * k:3 * k + 3], axis=1)
        rel = d / ref
        err[name] = {"max": float(rel.max()), "median": float(np.median(rel))}
    res["split_rel_error_vs_fp64"] = err
    res["split_over_fp64"] = res["split"]["interactions_per_s"] / res["fp64"]["interactions_per_s"]
    return res


def traffic_from_profile():
    """HBM bytes per force launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy of known size), if recorded."""
    try:
        with open(TRAFFIC_FILE) as f:
            for ln in f:
                if ln.startswith("bytes_per_launch:"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(P_raw, eps2, tend):
    """The reference's own driver and AVX2 back end (SRC/nbodysystem.h with
    SRC/hermite6-avx.h, built from its sources with its own flags into
    oracle/_ref) integrating the same initial state from t = 0."""
    try:
        import oracle
        ref = oracle.RefAVX(ORDER)
    except (ImportError, FileNotFoundError):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    info = ref.run(P_raw, eps2, ETA, ETA_S, DTMAX, tend, threads=threads)
    cpu = platform.processor() or platform.machine()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                cpu = ln.split(":", 1)[1].strip()
    except OSError:
        pass
    inter = float(info["num_step"]) * len(P_raw)
    return {
        "value": inter / info["seconds"],
        "unit": "interactions/s",
        "cores": info["threads"],
        "kind": "reference",
        "sample": f"reference driver + AVX2 back end, same N=65536 6th-order state, t=0->{tend:g} "
                  f"({info['num_bstep']} block steps, {info['num_step']} particle steps, "
                  f"{info['seconds']:.1f} s, dE/E {info['de']:.2e}), {cpu}",
        "gflops": inter / info["seconds"] * H.FLOPS[ORDER] * 1e-9,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="dt_max intervals timed (16 = t 0 -> 1)")
    ap.add_argument("--warmup", type=int, default=1, help="dt_max intervals on a discarded context")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-tend", type=float, default=DTMAX, help="time span of the CPU baseline run")
    ap.add_argument("--graph", action="store_true",
                    help="replay the timed steps from a hipGraph (no per-launch events: roofline from a second, "
                         "event-timed pass)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run 5 device-to-device copies of 256 MiB (the reference point that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU j-split is not "
              "built", file=sys.stderr)
        sys.exit(2)
    if args.calibrate:
        log("calibration: importing torch")
        import torch
        src = torch.ones(256 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copies done (5 x 256 MiB device-to-device)")
        del src, dst
    eps2 = EPS * EPS
    tend = args.steps * DTMAX
    log("init_step on the GPU")
    P0, e0, P_raw = initial_state(eps2)
    if args.warmup > 0:
        log(f"warmup: {args.warmup} interval(s) on a discarded context")
        integrate("fp64", P0, e0, eps2, args.warmup * DTMAX, timing=not args.graph, graph=args.graph)
    log(f"timed: FP64 resident integration t=0->{tend:g}")
    f = integrate("fp64", P0, e0, eps2, tend, timing=not args.graph, graph=args.graph)
    if f["error"]:
        raise SystemExit(f"FP64 integration failed: {f['error']}")
    if args.graph:
        log("event-timed pass for the roofline")
        ft = integrate("fp64", P0, e0, eps2, tend, timing=True)
        f["force_ms"], f["force_launches"] = ft["force_ms"], ft["force_launches"]
    inter = float(f["num_step"]) * N
    value = inter / f["wall_s"]
    log(f"FP64: {f['wall_s']:.3f} s, {f['num_bstep']} block steps, {value:.3e} int/s, dE/E {f['dE_E']:.3e}")
    log("split-precision integration, one interval")
    s = integrate("split", P0, e0, eps2, DTMAX, timing=False)
    log(f"split: stopped at t={s['tsys']:.3e} after {s['num_bstep']} block steps: {s['error']}")
    ap2 = all_pairs(P0, eps2)
    # roofline of the dominant kernel (the persistent force kernel)
    launches = max(f["force_launches"], 1)
    avg_ms = f["force_ms"] / launches
    flops_per_launch = inter * H.FLOPS[ORDER] / launches
    achieved = flops_per_launch / (avg_ms * 1e-3) * 1e-12
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: reference AVX2 driver t=0->{args.cpu_tend:g}")
        cpu = cpu_baseline(P_raw, eps2, args.cpu_tend)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * f["wall_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Plummer sphere)",
        "config": {"workload": f"6th-order Hermite block-step integration, Plummer N=65536 seed 2, eps=1/1024, "
                               f"eta=0.2, dt_max=1/16, t=0->{tend:g}, plain FP64 kernel, GPU-resident steps "
                               + ("replayed from a hipGraph" if args.graph else "launched on one stream"),
                   "order": ORDER, "n": N, "kernel": "fp64"},
        "gflops": value * H.FLOPS[ORDER] * 1e-9,
        "dE_E": f["dE_E"],
        "block_steps": f["num_bstep"],
        "mean_nact": f["num_step"] / f["num_bstep"],
        "us_per_block_step": 1e6 * f["wall_s"] / f["num_bstep"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profile(),
                     "kernel": "r_force_kernel<6,FP64>", "flops_per_interaction": H.FLOPS[ORDER],
                     "avg_launch_ms": avg_ms, "launches": f["force_launches"],
                     "force_share_of_wall": f["force_ms"] * 1e-3 / f["wall_s"]},
        "split": {"integration_error": s["error"], "tsys_reached": s["tsys"], "block_steps": s["num_bstep"],
                  "all_pairs_interactions_per_s": ap2["split"]["interactions_per_s"],
                  "all_pairs_over_fp64": ap2["split_over_fp64"],
                  "rel_error_vs_fp64": ap2["split_rel_error_vs_fp64"]},
        "all_pairs_fp64": {"interactions_per_s": ap2["fp64"]["interactions_per_s"], "gflops": ap2["fp64"]["gflops"],
                           "force_ms": ap2["fp64"]["force_ms"],
                           "frac": ap2["fp64"]["gflops"] * 1e-3 / PEAK_FP64_TFLOPS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
