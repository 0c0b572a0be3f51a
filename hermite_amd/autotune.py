"""Autotuner for the force launch geometry.

For one (order, variant, N) it times every candidate geometry of the force
call (i-parallel tiles with 1 or 2 i per lane and a j-split, or the
j-parallel kernel with a j-split) at active counts 1, 2, 4, ..., N with HIP
events on the device (hh_time_force), keeps the fastest per active count and
turns the winners into a table of rows (nact_hi, kernel, ni, jsplit): a row
covers the active counts up to the geometric midpoint between its measured
count and the next one.  The reference fixes its geometry at compile time
(NJBLOCK/NXREDUCE, SRC/hermite6-titan.h:7-14, and the AVX back end's fixed
SIMD grouping); here it is measured per device and per N.

Tables live in one JSON file keyed "order/variant/N" (tuning/mi355x.json
next to this module); Gravity contexts load the row set of the nearest tuned
N automatically (HERMITE_HIP_TUNING=off disables it, a path selects another
file).  The C++ drop-in takes the text form written by write_text through
hh_load_tuning_file.

    python -m hermite_amd.autotune --orders 4 6 8 --variants fp64 split \\
        --sizes 1024 4096 16384 65536 262144
    python -m hermite_amd.autotune --orders 6 --sizes 1048576 --max-nact 16384
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "mi355x.json")
INT_MAX = 2**31 - 1


def candidates(order, variant, nbody, nact):
    """Launch geometries (kernel, ni, jsplit) worth timing at this size."""
    ntiles = (nbody + 255) // 256
    splits = [0] + [s for s in (1, 2, 4, 8, 16, 32, 64, 128, 256) if s <= ntiles]
    iblocks = max(1, nact // 256)
    # more than ~8192 workgroups never pays; skip those splits
    splits = [s for s in splits if s == 0 or s * iblocks <= 8192]
    out = [(1, 1, s) for s in splits]
    if variant != "exact":
        out += [(1, 2, s) for s in splits if s * max(1, nact // 512) <= 8192]
        if nact <= 256:
            out += [(2, 1, s) for s in [0] + [s for s in (4, 8, 16, 32, 64, 128, 256) if s <= ntiles]]
    return out


def _state(order, nbody, seed=7):
    import hermite_amd as H
    m, x, v = H.plummer(nbody, seed)
    P = H.make_particles(order, m, x, v)
    rng = np.random.default_rng(seed)
    # plausible stored derivatives so the predictor does real work
    P[:, 10:] = 1e-3 * rng.standard_normal(P[:, 10:].shape)
    return P


def tune(order, variant, nbody, reps=None, device=0, log=None, max_nact=0):
    """Measure the candidates at active counts 1, 2, 4, ..., nbody (up to
    max_nact when it is set and below nbody; larger counts then keep the
    built-in geometry) and return (rows, measurements)."""
    import hermite_amd as H
    eps2 = (4.0 / nbody) ** 2
    g = H.Gravity(order, nbody, variant, device=device, tuning=False)
    try:
        g.set_jp_all(_state(order, nbody))
        g.predict_all(1e-3)
        nacts = sorted(set([1 << k for k in range(int(math.log2(nbody)) + 1)] + [nbody]))
        if 0 < max_nact < nbody:
            nacts = [x for x in nacts if x <= max_nact]
        best, meas = [], {}
        for nact in nacts:
            work = nact * nbody
            r = reps or max(3, min(50, int(2e9 / max(work, 1))))
            res = []
            for (k, ni, s) in candidates(order, variant, nbody, nact):
                g.set_tuning(k, s)
                g.set_tuning_ex(ni, 0)
                g.time_force(nact, eps2, 1)  # warm (allocates partials)
                ms = g.time_force(nact, eps2, r)
                res.append((ms, k, ni, s))
            res.sort()
            meas[nact] = [(round(ms * 1e3, 2), k, ni, s) for ms, k, ni, s in res]
            ms, k, ni, s = res[0]
            best.append((nact, k, ni, s, ms))
            if log:
                print(f"order={order} {variant} N={nbody} nact={nact}: best k{k} ni{ni} s{s} {ms * 1e3:.1f} us "
                      f"({work / (ms * 1e-3):.3e} int/s)", file=log, flush=True)
    finally:
        g.close()
    rows = []
    for q, (nact, k, ni, s, _) in enumerate(best):
        if q + 1 < len(best):
            hi = int(math.sqrt(nact * best[q + 1][0]))
        else:
            hi = INT_MAX if nact == nbody else nact
        if rows and rows[-1][1:] == (k, ni, s):
            rows[-1] = (hi, k, ni, s)
        else:
            rows.append((hi, k, ni, s))
    return rows, meas


def key(order, variant, nbody):
    return f"{order}/{variant}/{nbody}"


def load(path=None):
    path = path or os.environ.get("HERMITE_HIP_TUNING") or DEFAULT_PATH
    if path == "off" or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def table_for(order, variant, nbody, path=None):
    """Rows tuned for the nearest N (in log2) of this order and variant, or
    None when the file holds none within a factor of 4."""
    tables = load(path)
    best, dist = None, 2.0 + 1e-9
    for k, rows in tables.items():
        o, v, n = k.split("/")
        if int(o) != order or v != variant:
            continue
        d = abs(math.log2(int(n)) - math.log2(nbody))
        if d <= dist:
            best, dist = rows, d
    return None if best is None else [tuple(r) for r in best]


def write_text(rows, path):
    """Text form read by hh_load_tuning_file: "nact_hi kernel ni jsplit"."""
    with open(path, "w") as f:
        f.write("# nact_hi kernel ni jsplit\n")
        for r in rows:
            f.write("%d %d %d %d\n" % tuple(r))


# ----------------------------------------------------------- resident plan
# Work split of the GPU-resident block step (hhr_set_plan, hhr_set_split_waves):
# jpar_max, smax, task_target, jpar_target, wj4_max, wj2_max.
PLAN_DEFAULT = (64, 256, 8192, 256, 8192, 16384)
PLAN_NAMES = ("jpar_max", "smax", "task_target", "jpar_target", "wj4_max", "wj2_max")
PLAN_CHOICES = ((32, 64, 128), (128, 256, 512), (4096, 8192), (128, 256, 512), (2048, 8192, 16384),
                (8192, 16384, 32768))
PLAN_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "mi355x_plan.json")
# span timed per candidate: whole dt_max intervals, or a block-step limit
PLAN_SPAN = {"c1": (1.0, 0), "c2": (0.25, 0), "c3": (1 / 32, 0), "c4": (0.25, 0), "c5": (1 / 16, 384)}


def plan_key(order, variant, n_total, nclu):
    return f"{order}/{variant}/{n_total}/{nclu}"


def _resident_states(w):
    import hermite_amd as H
    from hermite_amd import workloads as W
    out = []
    for c in w.clusters:
        s = H.System(w.order, W.cluster_state(w.order, c), 0.0, c.eps ** 2, w.eta, w.eta_s, w.dtmax, w.dtmin,
                     variant="fp64")
        try:
            s.init_step()
            out.append((s.ptcl(), c.eps ** 2))
        finally:
            s.close()
    return out


def _time_plan(w, variant, states, plan, span, max_bsteps):
    """Interactions per second of a resident integration over the span."""
    import hermite_amd as H
    r = H.Resident(w.order, variant, [(P, e2, w.eta, w.dtmax, w.dtmin, 0.0) for P, e2 in states])
    try:
        r.set_plan(*plan[:4])
        r.set_split_waves(*plan[4:])
        if max_bsteps:
            r.set_step_limit(max_bsteps)
        t0 = time.perf_counter()
        try:
            r.run(span)
        except H.StepLimitReached:
            pass
        wall = time.perf_counter() - t0
        inter = sum(r.stats(k)["num_step"] * len(states[k][0]) for k in range(len(states)))
    finally:
        r.close()
    return inter / wall


def tune_plan(name, variant="fp64", log=None, passes=1):
    """Coordinate descent over PLAN_CHOICES on workload `name`, starting from
    PLAN_DEFAULT; each candidate is timed twice (best kept).  Returns the
    best plan, its rate, the default's rate and every measurement."""
    from hermite_amd import workloads as W
    w = W.workload(name)
    span, maxb = PLAN_SPAN[name]
    states = _resident_states(w)
    meas = {}

    def rate(plan):
        if plan not in meas:
            meas[plan] = max(_time_plan(w, variant, states, plan, span, maxb) for _ in range(2))
            if log:
                print(f"{name} {variant} plan {plan}: {meas[plan]:.4e} int/s", file=log, flush=True)
        return meas[plan]

    best = PLAN_DEFAULT
    base = rate(best)
    for _ in range(passes):
        for i, choices in enumerate(PLAN_CHOICES):
            for v in choices:
                cand = best[:i] + (v,) + best[i + 1:]
                if cand[4] > cand[5]:
                    continue
                if rate(cand) > rate(best):
                    best = cand
    return best, rate(best), base, meas, w


PLAN_HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "hh_plan_table.hpp")


PLAN_MIN_GAIN = 1.02  # below this the built-in plan is kept (run-to-run noise is 1-2 %)


def write_plan_header(plans, path=PLAN_HEADER):
    """The measured plans as the C++ table hhr_create chooses from.  A plan
    enters the table only if it beat the built-in one by PLAN_MIN_GAIN;
    the others are listed as comments (measured, built-in plan kept)."""
    rows, kept = [], []
    for k in sorted(plans, key=lambda k: [int(x) if x.isdigit() else x for x in k.split("/")]):
        order, variant, n, nclu = k.split("/")
        e = plans[k]
        p = e["plan"]
        vals = f"{p['jpar_max']}, {p['smax']}, {p['task_target']}, {p['jpar_target']}, {p['wj4_max']}, {p['wj2_max']}"
        if e["gain"] >= PLAN_MIN_GAIN:
            rows.append(f"    {{{order}, {n}, {nclu}, {{{vals}}}}},  // {e['workload']}, {variant}: "
                        f"{e['gain']:.3f}x the built-in plan")
        else:
            kept.append(f"//   order {order}, N {n}, {nclu} cluster(s) ({e['workload']}, {variant}): best {{{vals}}} "
                        f"measured {e['gain']:.3f}x the built-in plan")
    rows.append("    {0, 0, 0, {0, 0, 0, 0, 0, 0}},  // end of table")
    with open(path, "w") as f:
        f.write("// Resident block-step plans measured on MI355X by\n"
                "// `python -m hermite_amd.autotune --plan c1 c2 c3 c4 c5` (hermite_amd/tuning/mi355x_plan.json):\n"
                "// order, total particles, clusters, then hhr_set_plan's and hhr_set_split_waves'\n"
                "// parameters.  hhr_create takes the row of its order and kind (one cluster or an\n"
                f"// ensemble) nearest in total particles, within a factor of 4.  Only plans that beat\n"
                f"// the built-in one by {PLAN_MIN_GAIN:.2f}x or more are rows; for these the built-in plan\n"
                "// was kept (within run-to-run noise of the best found):\n" + "\n".join(kept) + "\n"
                "#pragma once\n\nnamespace hh {\n\nstruct PlanRow {\n    int order, n_total, nclu;\n"
                "    int v[6];  // jpar_max smax task_target jpar_target wj4_max wj2_max\n};\n\n"
                "static const PlanRow kPlanTable[] = {\n" + "\n".join(rows) + "\n};\n\n}  // namespace hh\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--orders", type=int, nargs="+", default=[4, 6, 8])
    ap.add_argument("--variants", nargs="+", default=["fp64", "split"])
    ap.add_argument("--sizes", type=int, nargs="+", default=[1024, 4096, 16384, 65536])
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--log", default=None, help="also write every measurement here (JSON)")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--max-nact", type=int, default=0,
                    help="largest active count measured (0: N); above it the built-in geometry applies")
    ap.add_argument("--plan", nargs="+", default=None, metavar="WORKLOAD",
                    help="instead: tune the resident plan on these workloads (c1 ... c5)")
    ap.add_argument("--plan-variants", nargs="+", default=["fp64"])
    ap.add_argument("--plan-out", default=PLAN_PATH)
    ap.add_argument("--plan-header", default=PLAN_HEADER, help="C++ table written from the plans ('' for none)")
    a = ap.parse_args(argv)
    if a.plan:
        plans = {}
        if os.path.exists(a.plan_out):
            with open(a.plan_out) as f:
                plans = json.load(f)
        for name in a.plan:
            for variant in a.plan_variants:
                best, r_best, r_base, meas, w = tune_plan(name, variant, log=sys.stdout)
                plans[plan_key(w.order, variant, w.n_total, len(w.clusters))] = {
                    "workload": name, "plan": dict(zip(PLAN_NAMES, best)), "rate": r_best, "default_rate": r_base,
                    "gain": r_best / r_base}
                print(f"{name} {variant}: best {best} {r_best:.4e} int/s, default {r_base:.4e} "
                      f"({r_best / r_base:.3f}x)", flush=True)
                with open(a.plan_out, "w") as f:
                    json.dump(plans, f, indent=1, sort_keys=True)
        if a.plan_header:
            write_plan_header(plans, a.plan_header)
        return
    tables = load(a.out) if os.path.exists(a.out) else {}
    allmeas = {}
    t0 = time.time()
    for order in a.orders:
        for variant in a.variants:
            for n in a.sizes:
                rows, meas = tune(order, variant, n, device=a.device, log=sys.stdout, max_nact=a.max_nact)
                tables[key(order, variant, n)] = [list(r) for r in rows]
                allmeas[key(order, variant, n)] = meas
                print(f"{key(order, variant, n)}: {rows}  ({time.time() - t0:.0f} s)", flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(tables, f, indent=1, sort_keys=True)
    if a.log:
        with open(a.log, "w") as f:
            json.dump({k: {str(n): v for n, v in m.items()} for k, m in allmeas.items()}, f)


if __name__ == "__main__":
    main()
