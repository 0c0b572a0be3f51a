"""The benchmark workloads themselves on the GPU, at their full sizes, within
a bounded GPU time (hermite_amd/workloads.py defines them):

  c1  4th order, N = 1024, seed 1, eps 1/64, eta 0.1, dt_min 2^-30, t 0 -> 1:
      the GPU-resident EXACT integration equals the oracle's in-place
      schedule bit for bit.
  c3  8th order, N = 2^18 (seed 3, eps 1/4096) and
  c5  6th order, N = 2^20 Salpeter (seed 5, eps 1/4096), each in the particle
      order after init_step: momentum conservation of every force family,
      the oracle's forces on 64 sampled i-particles, FP64 against split;
      c5 also a bounded resident run (dt_min 2^-40, 256 block steps) of both
      kernels: finite states, no step clamped to dt_min.
  c4  the 256-cluster ensemble, t 0 -> 1/16: EXACT, a sample of clusters
      each bit-identical to itself run alone; FP64 and split energies
      against the EXACT run's.
Parity with the reference at these sizes rests on the size-independent
properties above and on the oracle sample (the oracle is pinned to the
reference's own CPU build at smaller sizes, tests/test_oracle_ref.py)."""
import numpy as np
import pytest

import hermite_amd as H
import oracle as O
from hermite_amd import workloads as W

pytestmark = pytest.mark.gpu


def fam_err(F, G, order):
    out = []
    for k in range(O.FSTRIDE[order] // 3):
        a, b = F[:, 3 * k:3 * k + 3], G[:, 3 * k:3 * k + 3]
        out.append(np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300))
    return out


def init_on_gpu(w, c):
    """Records after init_step (host driver on the GPU, FP64 kernel)."""
    s = H.System(w.order, W.cluster_state(w.order, c), 0.0, c.eps ** 2, w.eta, w.eta_s, w.dtmax, w.dtmin,
                 variant="fp64")
    try:
        s.init_step()
        return s.ptcl(), s.init_energy
    finally:
        s.close()


def all_force(order, P, variant, eps2):
    g = H.Gravity(order, len(P), variant)
    try:
        g.set_jp_all(P)
        g.predict_all(0.0)
        return g.calc_force_on_first_nact(len(P), eps2)
    finally:
        g.close()


def test_c1_resident_exact_equals_oracle():
    w = W.workload("c1")
    c = w.clusters[0]
    eps2 = c.eps ** 2
    print("c1: oracle init_step and in-place integration t = 0 -> 1", flush=True)
    m, x, v = H.plummer(c.n, c.seed)
    o = O.System(w.order, O.make_particles(w.order, m, x, v), 0.0, eps2, w.eta, w.eta_s, w.dtmax)
    o.init_step()
    r = H.Resident(w.order, "exact", [(o.ptcl(), eps2, w.eta, w.dtmax, w.dtmin, 0.0)])
    try:
        o.integrate_until_inplace(w.tend)
        r.run(w.tend)
        st = r.stats()
        print(f"c1: {st['num_bstep']} block steps, {st['num_step']} particle steps", flush=True)
        assert st["dtmin_hits"] == 0  # so the oracle's schedule (no dt_min) is the same
        assert (st["num_step"], st["num_bstep"], st["tsys"]) == (o.num_step, o.num_bstep, o.tsys)
        assert np.array_equal(r.particles(), o.ptcl())
        assert r.energy() == o.energy()
    finally:
        r.close()


def check_full_size(name, sample=64):
    w = W.workload(name)
    c = w.clusters[0]
    eps2 = c.eps ** 2
    print(f"{name}: init_step on the GPU, N = {c.n}", flush=True)
    P, _ = init_on_gpu(w, c)
    ids = O.ids_of(P)
    assert not np.array_equal(np.sort(ids), ids), "init_step left the generator order"
    m = P[:, O.W_MASS]
    F = all_force(w.order, P, "fp64", eps2)
    for k in range(O.FSTRIDE[w.order] // 3):
        mom = np.abs((m[:, None] * F[:, 3 * k:3 * k + 3]).sum(axis=0)).max()
        scale = (m[:, None] * np.abs(F[:, 3 * k:3 * k + 3])).sum(axis=0).max()
        print(f"{name}: family {k} momentum residual {mom / scale:.2e}", flush=True)
        assert mom / scale < 1e-12
    idx = np.linspace(0, c.n - 1, sample).astype(int)
    R = O.force_list(w.order, O.predict(w.order, P, 0.0), idx, eps2)
    errs = fam_err(F[idx], R, w.order)
    print(f"{name}: fp64 vs oracle on {sample} i", [f"{e.max():.2e}" for e in errs], flush=True)
    assert errs[0].max() <= 1e-12 and errs[1].max() <= 1e-12
    for e in errs[2:]:
        assert e.max() <= 1e-10
    S = all_force(w.order, P, "split", eps2)
    es = fam_err(S, F, w.order)
    print(f"{name}: split vs fp64 max", [f"{e.max():.2e}" for e in es], "median",
          [f"{np.median(e):.2e}" for e in es], flush=True)
    n64 = 2 if w.order == 8 else 1  # families in FP64 arithmetic
    for e in es[:n64]:
        assert e.max() <= 1e-11
    for e in es[n64:]:
        assert np.median(e) <= 1e-5
    return w, c, P


def test_c3_full_size_properties():
    check_full_size("c3")


def test_c5_full_size_properties_and_bounded_run():
    w, c, P = check_full_size("c5")
    eps2 = c.eps ** 2
    out = {}
    for variant in ("fp64", "split"):
        r = H.Resident(w.order, variant, [(P, eps2, w.eta, w.dtmax, w.dtmin, 0.0)])
        try:
            r.set_step_limit(256)
            with pytest.raises(H.StepLimitReached):
                r.run(w.tend)
            st = r.stats()
            Pr = r.particles()
        finally:
            r.close()
        print(f"c5 {variant}: {st['num_bstep']} block steps to t = {st['tsys']:.3e}, dt_min hits "
              f"{st['dtmin_hits']}", flush=True)
        assert 256 <= st["num_bstep"] < 256 + 64 and 0 < st["tsys"] < w.tend
        assert st["dtmin_hits"] == 0  # no step reaches 2^-40 in the first steps
        assert np.all(np.isfinite(Pr[:, 4:10]))
        out[variant] = (st, Pr)
    # the split kernel's acceleration is FP64 arithmetic: the step sizes
    # agree (the FP32 higher derivatives can move a quantised step by one
    # level now and then, so not bit for bit)
    ratio = out["split"][0]["tsys"] / out["fp64"][0]["tsys"]
    print(f"c5: time reached split / fp64 {ratio:.3f}", flush=True)
    assert 0.5 <= ratio <= 2.0


def test_c4_ensemble_sample_equals_clusters_alone():
    w = W.workload("c4")
    print(f"c4: init_step of {len(w.clusters)} clusters on the GPU", flush=True)
    states = [init_on_gpu(w, c) + (c.eps ** 2,) for c in w.clusters]
    clus = [(P, eps2, w.eta, w.dtmax, w.dtmin, 0.0) for P, _, eps2 in states]
    t = w.dtmax
    ens = H.Resident(w.order, "exact", clus)
    try:
        ens.run(t)
        de_exact = np.array([abs((ens.energy(k) - e0) / e0) for k, (_, e0, _) in enumerate(states)])
        sample = [0, 1, 37, 101, 128, 200, 254, 255]
        for k in sample:
            alone = H.Resident(w.order, "exact", [clus[k]])
            try:
                alone.run(t)
                assert ens.stats(k) == alone.stats(0), k
                assert np.array_equal(ens.particles(k), alone.particles(0)), k
            finally:
                alone.close()
        print(f"c4: exact ensemble t = {t}: clusters {sample} equal their solo runs; |dE/E| max "
              f"{de_exact.max():.2e}", flush=True)
    finally:
        ens.close()
    for variant in ("fp64", "split"):
        r = H.Resident(w.order, variant, clus)
        try:
            r.run(t)
            de = np.array([abs((r.energy(k) - e0) / e0) for k, (_, e0, _) in enumerate(states)])
            nb = np.array([r.stats(k)["num_bstep"] for k in range(len(clus))])
        finally:
            r.close()
        print(f"c4 {variant}: |dE/E| max {de.max():.2e} median {np.median(de):.2e}", flush=True)
        assert np.all(np.isfinite(de))
        assert np.median(de) <= 2 * np.median(de_exact) + 1e-12
        assert de.max() <= 4 * de_exact.max() + 1e-12
        assert nb.min() > 0


def test_c5_recipe_energy_vs_oracle_at_small_n():
    """The c5 recipe (Salpeter masses 0.1-100, eps 1/4096, eta 0.2,
    dt_max 1/16, dt_min 2^-40) at N = 2048, t = 0 -> 1/16, where the oracle
    integrates in about a second: the FP64 and split resident runs take the
    oracle's block steps to 2 % and reach |dE/E| within twice the oracle's
    (the full-size run's 2.7e-4 comes from close encounters of the massive
    stars at N = 2^20, which this sample cannot show)."""
    w = W.workload("c5")
    c = W.Cluster(2048, 5, 1 / 4096, "salpeter", 5005)
    eps2 = c.eps ** 2
    o = O.System(w.order, W.cluster_state(w.order, c), 0.0, eps2, w.eta, w.eta_s, w.dtmax)
    o.init_step()
    P0, e0 = o.ptcl(), o.init_energy
    o.integrate_until_inplace(w.tend)
    de_o = abs((o.energy() - e0) / e0)
    for variant in ("fp64", "split"):
        r = H.Resident(w.order, variant, [(P0, eps2, w.eta, w.dtmax, w.dtmin, 0.0)])
        try:
            r.run(w.tend)
            st = r.stats()
            de = abs((r.energy() - e0) / e0)
        finally:
            r.close()
        print(f"c5 recipe N=2048 {variant}: {st['num_bstep']} block steps (oracle {o.num_bstep}), |dE/E| {de:.3e} "
              f"(oracle {de_o:.3e})", flush=True)
        assert abs(st["num_bstep"] / o.num_bstep - 1) <= 0.02
        assert st["dtmin_hits"] == 0
        assert de <= 2 * de_o + 1e-13


@pytest.mark.parametrize("name,tend", [("c2", 1.0), ("c3", 1 / 32)])
def test_split_integration_energy_within_twice_fp64(name, tend):
    """The north star's integration bound on the benchmark workloads: the
    split-precision resident run reaches the end time with |dE/E| within
    twice the FP64 run's, in the same number of block steps to 2 % (the
    6th-order N = 2^16 run to t = 1; the 8th-order N = 2^18 run to
    t = 1/32, a quarter of its span)."""
    w = W.workload(name)
    c = w.clusters[0]
    P, e0 = init_on_gpu(w, c)
    out = {}
    for variant in ("fp64", "split"):
        r = H.Resident(w.order, variant, [(P, c.eps ** 2, w.eta, w.dtmax, w.dtmin, 0.0)])
        try:
            r.run(tend)
            st = r.stats()
            out[variant] = (st["num_bstep"], abs((r.energy() - e0) / e0))
        finally:
            r.close()
        print(f"{name} {variant} t={tend:g}: {out[variant][0]} block steps, |dE/E| {out[variant][1]:.3e}", flush=True)
    assert abs(out["split"][0] / out["fp64"][0] - 1) <= 0.02
    assert out["split"][1] <= 2 * out["fp64"][1]
