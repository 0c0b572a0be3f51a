"""Benchmark of the Hermite force + predictor hot path on one MI355X.

Default workload (c2): 6th-order Hermite, Plummer sphere N = 65536 from
seed 2 (Aarseth-Henon-Wielen sampler, radii truncated at 22.8, Henon units,
equal masses), eps = 1/1024, eta = 0.2 (eta_s = 0.1 for the first steps),
power-of-two block timesteps with dt_max = 1/16.  --workload selects one of
the other configurations (hermite_amd/workloads.py): c1 (4th order,
N = 1024), c3 (8th order, N = 2^18, t = 0 -> 1/8), c4 (256 clusters of
256..8192 particles, 4th order) or c5 (6th order, N = 2^20 with Salpeter
masses, dt down to 2^-40, t = 0 -> 1/16).

init_step (SRC/nbodysystem.h:565-631) runs on the GPU through the host
driver; the block-step integration then runs GPU-resident (schedule,
predict_all, force on the active list, corrector and timestep on the
device).  One step is one dt_max interval of integration (the reference's
integrate_one_dtmax, SRC/nbodysystem.h:904-944): the timed region is
steps x dt_max time units from t = 0, started and ended by a stream
synchronisation inside hhr_run, with the block steps replayed from a
hipGraph of 64 steps (--direct: launched directly on one stream).  Warmup
intervals run on a separate context that is discarded, so the timed run
starts from t = 0.

value = interactions/s of the plain FP64 kernel over the timed run
(sum of nact * N over block steps / wall), with |dE/E| at its end.  A second,
timed pass of the same integration (the device clock written at every
kernel's start, in the same hipGraph replay) gives the force kernel's mean
launch time for the roofline and the time split by active count.  Beside
it: the split-precision kernel's integration outcome, both kernels'
all-pairs force rate (the reference's "Init N square" evaluation), the split
kernel's force error against the FP64 kernel, single force calls at
nact = 1, 4, ..., 16384 and N with predict_all on the initial state (device
time per call; --micro runs only these, --no-micro skips them), and the
reference's own AVX2 back end, built from its sources, timed on the host
cores (cpu_baseline).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import hermite_amd as H  # noqa: E402
from hermite_amd import workloads as W  # noqa: E402

METRIC = "interactions/s and Gflops (60/97/144 flop conv.), split-prec vs FP64 at stated |ΔE/E|"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak
BATCH = 64               # block steps per launch batch (one hipGraph)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "r_force_kernel_6th_n65536_traffic.txt")

# Kept for the tests and tools that import the default workload's constants.
ORDER, N, SEED, EPS, ETA, ETA_S, DTMAX = 6, 65536, 2, 1.0 / 1024, 0.2, 0.1, 1.0 / 16


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def initial_state(eps2):
    """Particle records of the default workload after init_step on the GPU
    (FP64 kernel), its initial energy, and the raw records."""
    m, x, v = H.plummer(N, SEED)
    P = H.make_particles(ORDER, m, x, v)
    s = H.System(ORDER, P, 0.0, eps2, ETA, ETA_S, DTMAX, variant="fp64")
    try:
        s.init_step()
        return s.ptcl(), s.init_energy, P
    finally:
        s.close()


def initial_states(w):
    """(records after init_step, eps2, E0, raw records) per cluster of w."""
    out = []
    for c in w.clusters:
        P = W.cluster_state(w.order, c)
        s = H.System(w.order, P, 0.0, c.eps ** 2, w.eta, w.eta_s, w.dtmax, w.dtmin, variant="fp64")
        try:
            s.init_step()
            out.append((s.ptcl(), c.eps ** 2, s.init_energy, P))
        finally:
            s.close()
    return out


def integrate(w, variant, states, tend, timing=False, graph=False, chunks=1, wall_limit=None, max_bsteps=0,
              t_energy=None):
    """Resident integration t = 0 -> tend, in `chunks` equal time spans
    (each hhr_run call ends with a stream synchronisation; a progress line
    per chunk).  wall_s counts the hhr_run calls only.  With t_energy
    (0 < t_energy <= tend, a multiple of dt_max) |dE/E| is also recorded at
    that time, between two hhr_run calls (None when the run ends before).  With wall_limit (s)
    the run stops after the first chunk that ends past it; with max_bsteps a
    chunk stops after that many block steps."""
    clus = [(P0, eps2, w.eta, w.dtmax, w.dtmin, 0.0) for P0, eps2, _, _ in states]
    r = H.Resident(w.order, variant, clus)

    def de_now():
        return np.array([abs((r.energy(k) - states[k][2]) / states[k][2]) for k in range(len(states))])

    try:
        if timing:
            # device clock stamps at kernel starts: the steps still replay
            # from the graph (direct launches with --direct)
            r.set_timing(True, stamps=True)
        if max_bsteps:
            r.set_step_limit(max_bsteps)
        ends = [tend * k / chunks for k in range(1, chunks + 1)]
        if t_energy is not None and 0 < t_energy < tend and t_energy not in ends:
            ends = sorted(ends + [t_energy])
        wall, err, de_at = 0.0, None, None
        try:
            for t in ends:
                t0 = time.perf_counter()
                r.run(t, use_graph=graph, batch=BATCH)
                wall += time.perf_counter() - t0
                if len(ends) > 1:
                    log(f"  {variant} t={t:g}: {wall:.1f} s")
                if t == t_energy:
                    de_at = de_now()
                if wall_limit is not None and t < tend and wall > wall_limit:
                    err = f"stopped by the benchmark after {wall_limit:g} s of wall time"
                    break
        except H.HHError as ex:
            err = str(ex)
        st = [r.stats(k) for k in range(len(states))]
        inter = float(sum(s["num_step"] * len(states[k][0]) for k, s in enumerate(st)))
        out = {"wall_s": wall, "interactions": inter, "num_step": sum(s["num_step"] for s in st),
               "num_bstep": max(s["num_bstep"] for s in st), "tsys": min(s["tsys"] for s in st),
               "dtmin_hits": sum(s["dtmin_hits"] for s in st), "error": err}
        if err is None:
            de = de_now()
            out["dE_E"] = float(de.max())
            if len(de) > 1:
                out["dE_E_median"] = float(np.median(de))
            if t_energy is not None and t_energy == tend:
                de_at = de
        if de_at is not None:
            out["dE_E_at_t_energy"] = float(de_at.max())
            if len(de_at) > 1:
                out["dE_E_median_at_t_energy"] = float(np.median(de_at))
        if timing:
            out["force_ms"], out["force_launches"] = r.force_timing()
            out["log"] = r.step_log()
        return out
    finally:
        r.close()


def all_pairs(order, P0, eps2):
    """Force on all N from all N (inputs resident) for both kernels, and the
    split kernel's per-particle relative error against the FP64 kernel."""
    n = len(P0)
    res, F = {}, {}
    for variant in ("fp64", "split"):
        g = H.Gravity(order, n, variant)
        try:
            g.set_jp_all(P0)
            g.predict_all(0.0)
            ms0 = g.time_force(n, eps2, 1)
            reps = int(max(1, min(20, 3000.0 / max(ms0, 1e-3))))
            ms = g.time_force(n, eps2, reps)
            F[variant] = g.calc_force_on_first_nact(n, eps2)
            geo = g.last_geometry()
        finally:
            g.close()
        inter = float(n) * n
        res[variant] = {"force_ms": ms, "interactions_per_s": inter / (ms * 1e-3),
                        "gflops": inter * H.FLOPS[order] / (ms * 1e-3) * 1e-9, "geometry": geo}
        log(f"all-pairs {variant}: {ms:.3f} ms  {res[variant]['interactions_per_s']:.3e} int/s")
    err = {}
    names = ("acc", "jrk", "snp", "crk")[:H.FSTRIDE[order] // 3]
    for k, name in enumerate(names):
        d = np.linalg.norm(F["split"][:, 3 * k:3 * k + 3] - F["fp64"][:, 3 * k:3 * k + 3], axis=1)
        ref = np.linalg.norm(F["fp64"][:, 3 * k:3 * k + 3], axis=1)
        rel = d / ref
        err[name] = {"max": float(rel.max()), "median": float(np.median(rel))}
    res["split_rel_error_vs_fp64"] = err
    res["split_over_fp64"] = res["split"]["interactions_per_s"] / res["fp64"]["interactions_per_s"]
    return res


MICRO_NACT = (1, 4, 16, 64, 256, 1024, 4096, 16384)
PR_WORDS = {4: 7, 6: 10, 8: 13}  # predicted j record (pos, vel[, acc[, jrk]]), doubles
PEAK_HBM_GBS = 8000.0


def micro(order, P0, eps2, budget_ms=300.0):
    """Single force calls (calc_force_on_first_nact without the copy back:
    force kernel + slice reduction, tuned geometry, inputs resident) on the
    first nact particles for nact = 1, 4, ..., 16384 and N, both kernels,
    and predict_all of every particle (no dirty slots): mean device time per
    call from HIP events around repeated calls."""
    n = len(P0)
    rows = {}
    pred = None
    for variant in ("fp64", "split"):
        g = H.Gravity(order, n, variant)
        try:
            g.set_jp_all(P0)
            g.predict_all(0.0)
            if variant == "fp64":
                ms0 = g.time_predict(0.0, 1)
                ms = g.time_predict(0.0, int(max(5, min(1000, budget_ms / max(ms0, 1e-3)))))
                nbytes = float(n) * (H.JPSTRIDE[order] + PR_WORDS[order]) * 8
                pred = {"ms": ms, "bytes": nbytes, "gbytes_per_s": nbytes / (ms * 1e-3) * 1e-9,
                        "frac_hbm": nbytes / (ms * 1e-3) * 1e-9 / PEAK_HBM_GBS}
                log(f"micro predict_all: {1e3 * ms:.1f} us, {pred['gbytes_per_s']:.0f} GB/s")
            for nact in [x for x in MICRO_NACT if x < n] + [n]:
                ms0 = g.time_force(nact, eps2, 1)
                ms = g.time_force(nact, eps2, int(max(3, min(1000, budget_ms / max(ms0, 1e-3)))))
                inter = float(nact) * n
                tf = inter * H.FLOPS[order] / (ms * 1e-3) * 1e-12
                r = rows.setdefault(nact, {"nact": nact})
                r[f"{variant}_us"] = 1e3 * ms
                r[f"{variant}_interactions_per_s"] = inter / (ms * 1e-3)
                r[f"{variant}_frac"] = tf / PEAK_FP64_TFLOPS
            log(f"micro force {variant}: " + ", ".join(f"{k}: {r[variant + '_us']:.1f} us" for k, r in rows.items()))
        finally:
            g.close()
    return {"predict_all": pred, "force": list(rows.values()),
            "note": "device time per call; frac = conventional flops / 78.6 TF/s for both kernels (the split "
                    "kernel's own mix allows 88.9 TF/s at 6th order)"}


def traffic_from_profile(w):
    """HBM bytes per force launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy of known size), if recorded for
    this workload."""
    if w.name != "c2":
        return None
    try:
        with open(TRAFFIC_FILE) as f:
            for ln in f:
                if ln.startswith("bytes_per_launch:"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def _cpu_info():
    """lscpu model, sockets, cores per socket and threads per core, and the
    CPUs this process may run on."""
    info = {"model": platform.processor() or platform.machine()}
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except OSError:
        pass
    try:
        info["cpus_allowed"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


def _cpu_threads(info):
    """Threads for the CPU baseline: min(physical cores, 64) (the reference's
    MAXTHREAD, SRC/hermite6-avx.h:8), further capped by the CPU share this
    job is given (OMP_NUM_THREADS, which the GPU box sets to its 16 cores
    per GPU, and the CPUs the process may run on)."""
    phys = info.get("sockets", 0) * info.get("cores_per_socket", 0) if isinstance(info.get("sockets"), int) \
        and isinstance(info.get("cores_per_socket"), int) else 0
    n = min(phys or (os.cpu_count() or 1), 64)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if share:
        n = min(n, share)
    if info.get("cpus_allowed"):
        n = min(n, info["cpus_allowed"])
    return max(1, n)


def _reps_for(secs, target_s=10.0, most=2000):
    """Repetitions that bring a CPU sample of secs seconds to about target_s."""
    return int(min(most, max(1, math.ceil(target_s / max(secs, 1e-3))))) if secs < 0.5 * target_s else 1


def cpu_baseline(w, states, cpu_tend, budget_s=20.0):
    """The reference's own driver and AVX2 back end (SRC/nbodysystem.h with
    SRC/hermite{4,6,8}-avx.h, built from its sources with its own flags into
    oracle/_ref) on the same initial states.  Single systems of up to 2^16
    particles integrate t = 0 -> cpu_tend; larger ones time about 10 s of
    force calls on 4096 active particles; the ensemble integrates its clusters in
    order until about budget_s seconds of CPU work."""
    try:
        import oracle
        ref = oracle.RefAVX(w.order)
    except (ImportError, FileNotFoundError):
        return None
    cpuinfo = _cpu_info()
    threads = _cpu_threads(cpuinfo)
    cpu = cpuinfo["model"]
    # thread placement of the reference's OpenMP runs (child processes)
    os.environ["OMP_PLACES"] = "cores"
    os.environ["OMP_PROC_BIND"] = "close"
    if len(states) == 1 and len(states[0][3]) <= 65536:
        P_raw, eps2 = states[0][3], states[0][1]
        info = ref.run(P_raw, eps2, w.eta, w.eta_s, w.dtmax, cpu_tend, threads=threads)
        reps = _reps_for(info["seconds"])
        if reps > 1:  # a short run is repeated from the same state
            info = ref.run(P_raw, eps2, w.eta, w.eta_s, w.dtmax, cpu_tend, threads=threads, reps=reps)
        inter = float(reps) * info["num_step"] * len(P_raw)
        sample = (f"reference driver + AVX2 back end, same N={len(P_raw)} state, t=0->{cpu_tend:g} "
                  f"({info['num_bstep']} block steps, {info['num_step']} particle steps, dE/E {info['de']:.2e})"
                  + (f" run {reps} times" if reps > 1 else "") + f" ({info['seconds']:.1f} s), {cpu}")
        secs, used = info["seconds"], info["threads"]
    elif len(states) == 1:
        P0, eps2 = states[0][0], states[0][1]
        nact = 4096
        _, info = ref.forces(P0, 0.0, nact, eps2, reps=1, threads=threads, want_output=False)
        reps = int(min(200, max(2, 10.0 / max(info["seconds"], 1e-3))))  # about 10 s of calls
        _, info = ref.forces(P0, 0.0, nact, eps2, reps=reps, threads=threads, want_output=False)
        inter = float(reps) * nact * len(P0)
        secs, used = info["seconds"], info["threads"]
        sample = (f"reference AVX2 back end, force on {nact} active particles from all N={len(P0)} of the same "
                  f"state, {reps} calls ({secs:.1f} s), {cpu}")
    else:
        def ensemble_pass(reps, limit):
            inter, secs, used, done = 0.0, 0.0, threads, 0
            for P0, eps2, _, P_raw in states:
                if len(P_raw) % 4:  # the AVX back end takes i in groups of 4 (SRC/hermite4-avx.h:289)
                    continue
                if limit is not None and done >= limit:
                    break
                info = ref.run(P_raw, eps2, w.eta, w.eta_s, w.dtmax, cpu_tend, threads=threads, reps=reps)
                inter += float(reps) * info["num_step"] * len(P_raw)
                secs += info["seconds"]
                used = info["threads"]
                done += 1
                if secs > budget_s:
                    break
            return inter, secs, used, done

        inter, secs, used, done = ensemble_pass(1, None)
        reps = _reps_for(secs)
        if reps > 1:  # the same clusters again, each run reps times
            inter, secs, used, done = ensemble_pass(reps, done)
        sample = (f"reference driver + AVX2 back end on the first {done} clusters of the ensemble whose N_k is "
                  f"a multiple of 4, each t=0->{cpu_tend:g}" + (f" run {reps} times" if reps > 1 else "")
                  + f" ({secs:.1f} s), {cpu}")
    out = {"value": inter / secs, "unit": "interactions/s", "cores": used, "kind": "reference",
           "sample": sample, "gflops": inter / secs * H.FLOPS[w.order] * 1e-9,
           "per_core": inter / secs / max(used, 1), "cpu": cpuinfo,
           "placement": "OMP_PLACES=cores OMP_PROC_BIND=close"}
    if len(states) == 1 and len(states[0][0]) <= 65536 and len(states[0][0]) % 4 == 0:
        # the reference's "Init N square" evaluation: all N on all N, once
        P0, eps2 = states[0][0], states[0][1]
        n = len(P0)
        _, ai = ref.forces(P0, 0.0, n, eps2, reps=1, threads=threads, want_output=False)
        out["all_pairs"] = {"interactions_per_s": float(n) * n / ai["seconds"], "seconds": ai["seconds"],
                            "gflops": float(n) * n * H.FLOPS[w.order] / ai["seconds"] * 1e-9}
    return out


def by_active_count(log_, n_j):
    """Step time split by the step's active count: force (force kernel start
    to corrector start) and the step's other launches, from device clock
    stamps."""
    na, fm, om = log_
    out = []
    for lo, hi in ((1, 64), (65, 1024), (1025, 16384), (16385, 1 << 40)):
        m = (na >= lo) & (na <= hi)
        if m.any():
            out.append({"nact": f"{lo}-{hi}" if hi < (1 << 40) else f">={lo}", "steps": int(m.sum()),
                        "force_ms": round(float(fm[m].sum()), 2), "other_ms": round(float(om[m].sum()), 2),
                        "interaction_share": round(float(na[m].sum() / max(na.sum(), 1)), 4)})
    return out


def kernel_rate(ft, order):
    """Mean stamp-timed force launch (ms: force kernel start to corrector
    start) and the conventional TF/s of a launch (interactions x 60/97/144
    flops per launch / its mean time)."""
    launches = max(ft["force_launches"], 1)
    avg_ms = ft["force_ms"] / launches
    flops_per_launch = ft["interactions"] * H.FLOPS[order] / launches
    return avg_ms, flops_per_launch / (avg_ms * 1e-3) * 1e-12


def opmix_from_profile(order):
    """Lane instructions per interaction of the split force kernel, from the
    rocprofv3 counter summary of this order if recorded: {"f64": FP64-rate
    instructions (FP64 arithmetic and conversions), "f32": FP32 ones}."""
    path = os.path.join(ROOT, "profiles", f"r_force_kernel_{order}th_split_opmix.md")
    try:
        with open(path) as fh:
            for ln in fh:
                if ln.startswith("opmix:"):
                    return json.loads(ln.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def split_roofline(w, avg_ms, achieved, launches):
    """Roofline of r_force_kernel<ORDER,SPLIT>: its conventional TF/s
    (60/97/144 flops per interaction) against the FP64 vector peak, the FP32
    vector peak, and the blend of the two that its measured instruction mix
    allows: with n64 FP64-rate and n32 FP32 lane instructions per
    interaction, issued at 78.6/2 and 157.3/2 T per second, the blended peak
    is flops / (n64 / 39.3 + n32 / 78.65) TF/s."""
    mix = opmix_from_profile(w.order)
    out = {"bound": "valu", "kernel": force_kernel_name(w, "SPLIT"), "achieved": achieved,
           "unit": "TFLOP/s", "avg_launch_ms": avg_ms, "launches": launches,
           "peak_fp64": PEAK_FP64_TFLOPS, "frac_fp64": achieved / PEAK_FP64_TFLOPS,
           "peak_fp32": PEAK_FP32_TFLOPS, "frac_fp32": achieved / PEAK_FP32_TFLOPS, "opmix": mix}
    if mix:
        blend = H.FLOPS[w.order] / (mix["f64"] / (PEAK_FP64_TFLOPS / 2) + mix["f32"] / (PEAK_FP32_TFLOPS / 2))
        out.update({"peak_blended": blend, "frac_blended": achieved / blend})
    return out


FUSED_MAX = 4096  # single clusters up to this size take the two-launch block step


def force_kernel_name(w, mode):
    """The force kernel the GPU-resident block steps launch: for one cluster
    of at most FUSED_MAX particles the two-launch step's
    r_fused_islice_kernel<ORDER,MODE,JPL> (up to 1024 particles, 2048 at
    4th order) or r_fused_force_kernel<ORDER,MODE,JPL> (JPL: j per lane, 2
    up to 1024 particles and at 8th order, else 4); r_force_kernel<ORDER,MODE>
    otherwise."""
    n = w.n_total
    if len(w.clusters) == 1 and n <= FUSED_MAX:
        jpl = 2 if (n <= 1024 or w.order == 8) else 4
        kern = "r_fused_islice_kernel" if (n <= 1024 or (w.order == 4 and n <= 2048)) else "r_fused_force_kernel"
        return f"{kern}<{w.order},{mode},{jpl}>"
    return f"r_force_kernel<{w.order},{mode}>"


def workload_text(w, tend, graph):
    kind = "hipGraph replay" if graph else "direct launches on one stream"
    if len(w.clusters) == 1 and w.n_total <= FUSED_MAX:
        kind += ", two launches per block step"
    if w.name == "c4":
        return (f"4th-order Hermite ensemble of {len(w.clusters)} Plummer clusters, N_k 256..8192 "
                f"(sum {w.n_total}), eps=4/N_k, eta=0.1, dt_max=1/16, t=0->{tend:g}, plain FP64 kernel, "
                f"all clusters in shared GPU-resident launches ({kind})")
    c = w.clusters[0]
    mass = "Salpeter masses" if c.mass == "salpeter" else "equal masses"
    return (f"{w.order}th-order Hermite block-step integration, Plummer N={c.n} seed {c.seed} ({mass}), "
            f"eps=1/{round(1 / c.eps)}, eta={w.eta:g}, dt_max=1/{round(1 / w.dtmax)}"
            + (f", dt_min=2^{round(np.log2(w.dtmin))}" if w.dtmin > 0 else "")
            + f", t=0->{tend:g}, plain FP64 kernel, GPU-resident steps ({kind})")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--steps", type=int, default=0,
                    help="dt_max intervals timed (default: the workload's whole span, 16 for c2 = t 0 -> 1)")
    ap.add_argument("--warmup", type=int, default=1, help="dt_max intervals on a discarded context")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-tend", type=float, default=0.0, help="time span of the CPU baseline run (default dt_max)")
    ap.add_argument("--direct", action="store_true",
                    help="launch the timed steps directly on the stream instead of replaying a hipGraph")
    ap.add_argument("--single-pass", action="store_true",
                    help="time the stamp-timed pass itself instead of a separate untimed run (default for the "
                         "long workloads c3 and c5; the stamps add under 1 %% to a run)")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision legs")
    ap.add_argument("--no-micro", action="store_true", help="skip the single-force-call microbenchmark")
    ap.add_argument("--micro", action="store_true",
                    help="only the microbenchmark: predict_all and single force calls at nact = 1 ... N on the "
                         "workload's initial state (one JSON line; single-cluster workloads)")
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
    w = W.workload(args.workload)
    graph = not args.direct
    steps = args.steps or int(round(w.tend / w.dtmax))
    tend = steps * w.dtmax
    cpu_tend = args.cpu_tend or w.dtmax
    log(f"workload {w.name}: init_step on the GPU ({len(w.clusters)} cluster(s), N = {w.n_total})")
    states = initial_states(w)
    if args.micro:
        if len(states) != 1:
            raise SystemExit("--micro takes a single-cluster workload (c1, c2, c3, c5)")
        m = micro(w.order, states[0][0], states[0][1])
        big = m["force"][-1]
        print(json.dumps({"metric": "single force calls and predict_all by active count", "value": big["fp64_interactions_per_s"],
                          "unit": "interactions/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic (seeded Plummer spheres)",
                          "config": {"workload": f"{w.name}: initial state after init_step, N = {w.n_total}",
                                     "name": w.name, "order": w.order, "n": w.n_total},
                          "micro": m}), flush=True)
        return
    if args.warmup > 0:
        log(f"warmup: {args.warmup} interval(s) on a discarded context")
        integrate(w, "fp64", states, min(args.warmup, steps) * w.dtmax, graph=graph)
    single = args.single_pass or w.name in ("c3", "c5")
    chunks = 16 if w.name in ("c3", "c5") else 1
    t_cfg = w.tend  # the workload's own end time: |dE/E| is reported there too
    log(f"timed: FP64 resident integration t=0->{tend:g}" + (" (stamp-timed)" if single else ""))
    f = integrate(w, "fp64", states, tend, timing=single, graph=graph, chunks=chunks, t_energy=t_cfg)
    if f["error"]:
        raise SystemExit(f"FP64 integration failed: {f['error']}")
    value = f["interactions"] / f["wall_s"]
    log(f"FP64: {f['wall_s']:.3f} s, {f['num_bstep']} block steps, {value:.3e} int/s, dE/E {f['dE_E']:.3e}")
    if single:
        ft = f
    else:
        log("stamp-timed pass of the same integration (roofline)")
        ft = integrate(w, "fp64", states, tend, timing=True, graph=graph)
    avg_ms, achieved = kernel_rate(ft, w.order)
    split, ap2, split_roof = None, None, None
    if not args.no_split:
        # same span, same checks; a run that collapses or crawls (steps
        # pinned at dt_min) is stopped at 3x the FP64 wall or 4x its block
        # steps and reported as such
        log(f"split-precision integration t=0->{tend:g}" + (" (stamp-timed)" if single else ""))
        s = integrate(w, "split", states, tend, timing=single, graph=graph, chunks=chunks, t_energy=t_cfg,
                      wall_limit=max(60.0, 3 * f["wall_s"]), max_bsteps=4 * f["num_bstep"] + 1024)
        log(f"split: t={s['tsys']:.3e} after {s['num_bstep']} block steps, {s['wall_s']:.3f} s: "
            f"{s['error'] or 'ok'}, dE/E {s.get('dE_E')}")
        split = {"integration_error": s["error"], "tsys_reached": s["tsys"], "block_steps": s["num_bstep"],
                 "wall_s": s["wall_s"], "interactions_per_s": s["interactions"] / s["wall_s"],
                 "over_fp64": (s["interactions"] / s["wall_s"]) / value,
                 "dE_E": s.get("dE_E"), "dE_E_median": s.get("dE_E_median"),
                 f"dE_E_at_t{t_cfg:g}": s.get("dE_E_at_t_energy"),
                 f"dE_E_at_t{t_cfg:g}_over_fp64": (None if s.get("dE_E_at_t_energy") is None
                                                  else s["dE_E_at_t_energy"] / f["dE_E_at_t_energy"])}
        if s["error"] is None:
            st_ = s if single else integrate(w, "split", states, tend, timing=True, graph=graph)
            s_ms, s_ach = kernel_rate(st_, w.order)
            split_roof = split_roofline(w, s_ms, s_ach, st_["force_launches"])
        if len(states) == 1:
            ap2 = all_pairs(w.order, states[0][0], states[0][1])
            split.update({"all_pairs_interactions_per_s": ap2["split"]["interactions_per_s"],
                          "all_pairs_over_fp64": ap2["split_over_fp64"],
                          "rel_error_vs_fp64": ap2["split_rel_error_vs_fp64"]})
    mic = None
    if not args.no_micro and len(states) == 1:
        mic = micro(w.order, states[0][0], states[0][1])
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: reference AVX2 back end (t=0->{cpu_tend:g} where it integrates)")
        cpu = cpu_baseline(w, states, cpu_tend)
    n_j = w.n_total
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * f["wall_s"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Plummer spheres)",
        "config": {"workload": workload_text(w, tend, graph and not single), "name": w.name, "order": w.order,
                   "n": n_j, "clusters": len(w.clusters), "kernel": "fp64"},
        "gflops": value * H.FLOPS[w.order] * 1e-9,
        "dE_E": f["dE_E"],
        f"dE_E_at_t{t_cfg:g}": f.get("dE_E_at_t_energy"),
        "block_steps": f["num_bstep"],
        "mean_nact": f["num_step"] / max(f["num_bstep"], 1),
        "us_per_block_step": 1e6 * f["wall_s"] / max(f["num_bstep"], 1),
        "dtmin_hits": f["dtmin_hits"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profile(w),
                     "kernel": force_kernel_name(w, "FP64"), "flops_per_interaction": H.FLOPS[w.order],
                     "avg_launch_ms": avg_ms, "launches": ft["force_launches"],
                     "force_share_of_timed_wall": ft["force_ms"] * 1e-3 / ft["wall_s"],
                     "by_active_count": by_active_count(ft["log"], n_j)},
        "split": split,
        "roofline_split": split_roof,
        "all_pairs_fp64": None if ap2 is None else {
            "interactions_per_s": ap2["fp64"]["interactions_per_s"], "gflops": ap2["fp64"]["gflops"],
            "force_ms": ap2["fp64"]["force_ms"], "frac": ap2["fp64"]["gflops"] * 1e-3 / PEAK_FP64_TFLOPS},
        "micro": mic,
        "cpu_baseline": cpu,
    }
    how = "graph replay" if graph else "direct launches"
    line["timing"] = (f"one stamp-timed pass (device clock at every kernel start, {how})" if single else
                      f"untimed run, then a stamp-timed pass (device clock at every kernel start, {how}) "
                      "for the roofline")
    if w.name == "c4":
        line["dE_E_median"] = f.get("dE_E_median")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
