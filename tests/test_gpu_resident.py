"""GPU-resident block steps (hhr_*): schedule, prediction, force and
corrector on the device, one or many clusters, hipGraph replay.

The resident driver keeps particles in place (no dt-level permutation), so
its EXACT variant is bit-identical to the oracle's in-place schedule
(orc_sys_integrate_one_block_inplace); that schedule reproduces the nact
known answers of the reference logs.  An ensemble of independent clusters
must give each cluster exactly what it gives when run alone."""
import numpy as np
import pytest

import hermite_amd as H
import oracle as O
from tests.test_golden import NACT_KAT

pytestmark = pytest.mark.gpu

ORDERS = (4, 6, 8)


def initialised(order, n, seed, eps2, eta, eta_s, dtmax, rmax=22.8):
    """Particles after the reference init_step (oracle driver)."""
    m, x, v = H.plummer(n, seed, rmax)
    s = O.System(order, O.make_particles(order, m, x, v), 0.0, eps2, eta, eta_s, dtmax)
    s.init_step()
    return s


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("use_graph", [False, True])
def test_resident_exact_bitwise_vs_oracle_inplace(order, use_graph):
    n, eps2 = 500, (1 / 32) ** 2
    eta, eta_s = {4: (0.1, 0.01), 6: (0.4, 0.1), 8: (0.6, 0.1)}[order]
    o = initialised(order, n, 41, eps2, eta, eta_s, 1 / 16)
    r = H.Resident(order, "exact", [(o.ptcl(), eps2, eta, 1 / 16, 0.0, 0.0)])
    print(f"resident exact order={order} graph={use_graph}", flush=True)
    for t in (1 / 16, 1 / 4):
        o.integrate_until_inplace(t)
        r.run(t, use_graph=use_graph, batch=4)
        st = r.stats()
        assert (st["num_step"], st["num_bstep"], st["tsys"]) == (o.num_step, o.num_bstep, o.tsys)
        assert np.array_equal(r.particles(), o.ptcl())
    assert r.energy() == o.energy()
    r.close()


@pytest.mark.parametrize("order", ORDERS)
def test_resident_exact_reproduces_reference_nact(order):
    eta, eta_s, want1, want2 = NACT_KAT[order]
    n = 1024
    eps2 = (4.0 / n) ** 2
    o = initialised(order, n, 19640916, eps2, eta, eta_s, 1 / 16, rmax=100.0)
    r = H.Resident(order, "exact", [(o.ptcl(), eps2, eta, 1 / 16, 0.0, 0.0)])
    r.run(1.0)
    s1 = r.stats()
    r.run(2.0)
    s2 = r.stats()
    r.close()
    assert "%e" % (s1["num_step"] / s1["num_bstep"]) == want1
    assert "%e" % ((s2["num_step"] - s1["num_step"]) / (s2["num_bstep"] - s1["num_bstep"])) == want2


@pytest.mark.parametrize("order", ORDERS)
def test_resident_fp64_statistics(order):
    """FP64 kernel: same mean nact as the reference logs, |dE/E| of the same
    size as the oracle's on the same schedule."""
    eta, eta_s, want1, _ = NACT_KAT[order]
    n = 1024
    eps2 = (4.0 / n) ** 2
    o = initialised(order, n, 19640916, eps2, eta, eta_s, 1 / 16, rmax=100.0)
    P0, e0 = o.ptcl(), o.init_energy
    o.integrate_until_inplace(1.0)
    de_o = abs((o.energy() - e0) / e0)
    r = H.Resident(order, "fp64", [(P0, eps2, eta, 1 / 16, 0.0, 0.0)])
    r.run(1.0)
    st = r.stats()
    de = abs((r.energy() - e0) / e0)
    print(f"order={order} resident fp64: nact={st['num_step'] / st['num_bstep']:.6e} (reference {want1}) "
          f"|dE/E|={de:.3e} (oracle {de_o:.3e})", flush=True)
    assert abs(st["num_step"] / st["num_bstep"] / float(want1) - 1) < 0.01
    assert de < 3 * de_o + 1e-13
    r.close()


@pytest.mark.parametrize("order", ORDERS)
def test_resident_split_integrates(order):
    """The split kernel keeps the acceleration (8th order: and the jerk) in
    FP64 and the higher families in FP32, within the noise each family
    tolerates (tests/test_noise_sensitivity.py), so every order integrates
    the reference's pl1k run (t = 0 -> 1): the same mean active count as the
    reference log to 1 %, and |dE/E| within twice the FP64 run's."""
    eta, eta_s, want1, _ = NACT_KAT[order]
    n = 1024
    eps2 = (4.0 / n) ** 2
    o = initialised(order, n, 19640916, eps2, eta, eta_s, 1 / 16, rmax=100.0)
    P0, e0 = o.ptcl(), o.init_energy
    de = {}
    for variant in ("fp64", "split"):
        r = H.Resident(order, variant, [(P0, eps2, eta, 1 / 16, 0.0, 0.0)])
        try:
            r.run(1.0)
            st = r.stats()
            de[variant] = abs((r.energy() - e0) / e0)
        finally:
            r.close()
        print(f"{variant} order={order}: nact={st['num_step'] / st['num_bstep']:.4e} (reference {want1}) "
              f"|dE/E|={de[variant]:.3e}", flush=True)
        assert abs(st["num_step"] / st["num_bstep"] / float(want1) - 1) < 0.01
    assert de["split"] <= 2 * de["fp64"] + 1e-12, de


def test_ensemble_equals_clusters_run_alone():
    """Uneven clusters in one resident context: every cluster's particles,
    step counts and energy equal those of the same cluster run alone
    (EXACT variant: no summation-order freedom)."""
    order, eta, eta_s, dtmax = 4, 0.1, 0.01, 1 / 16
    sizes = [256, 300, 700, 1024, 333, 512, 1500, 260]
    clusters, solo = [], []
    for k, n in enumerate(sizes):
        eps2 = (4.0 / n) ** 2
        o = initialised(order, n, 100 + k, eps2, eta, eta_s, dtmax)
        clusters.append((o.ptcl(), eps2, eta, dtmax, 0.0, 0.0))
    ens = H.Resident(order, "exact", clusters)
    ens.run(0.25)
    for k, c in enumerate(clusters):
        alone = H.Resident(order, "exact", [c])
        alone.run(0.25)
        assert ens.stats(k) == alone.stats(0), k
        assert np.array_equal(ens.particles(k), alone.particles(0)), k
        assert ens.energy(k) == alone.energy(0)
        alone.close()
    ens.close()


def test_ensemble_fp64_matches_oracle_statistically():
    order, eta, eta_s, dtmax = 4, 0.1, 0.01, 1 / 16
    sizes = [256, 1024, 4096, 512]
    clusters, oracles = [], []
    for k, n in enumerate(sizes):
        eps2 = (4.0 / n) ** 2
        o = initialised(order, n, 200 + k, eps2, eta, eta_s, dtmax)
        clusters.append((o.ptcl(), eps2, eta, dtmax, 0.0, 0.0))
        oracles.append(o)
    ens = H.Resident(order, "fp64", clusters)
    ens.run(0.25)
    for k, o in enumerate(oracles):
        e0 = o.init_energy
        o.integrate_until_inplace(0.25)
        st = ens.stats(k)
        assert abs(st["num_step"] / o.num_step - 1) < 0.02
        assert abs((ens.energy(k) - e0) / e0) < 3 * abs((o.energy() - e0) / e0) + 1e-12
    ens.close()


@pytest.mark.parametrize("use_graph,stamps,fused", [(False, False, True), (True, False, True), (False, True, True),
                                                     (True, True, True), (True, True, False), (False, True, False)])
def test_resident_force_timing_changes_nothing(use_graph, stamps, fused):
    """Step timing (events around direct launches, or the device clock
    written at every kernel's start, in replayed graphs too), on the
    two-launch and the three-launch step, leaves the integration
    bit-identical, counts every launch and logs every block step with a
    positive force and other time."""
    order, n, eps2 = 6, 500, (1 / 32) ** 2
    o = initialised(order, n, 43, eps2, 0.4, 0.1, 1 / 16)
    P0 = o.ptcl()
    print(f"resident timing graph={use_graph} stamps={stamps} fused={fused}", flush=True)
    a = H.Resident(order, "fp64", [(P0, eps2, 0.4, 1 / 16, 0.0, 0.0)])
    b = H.Resident(order, "fp64", [(P0, eps2, 0.4, 1 / 16, 0.0, 0.0)])
    if not fused:
        a.set_fused(0)
        b.set_fused(0)
    a.run(1 / 8, use_graph=use_graph, batch=8)
    b.set_timing(True, stamps=stamps)
    b.run(1 / 16, use_graph=use_graph, batch=8)
    b.run(1 / 8, use_graph=use_graph, batch=8)
    ms, launches = b.force_timing()
    na, fm, om = b.step_log()
    assert a.stats() == b.stats()
    assert np.array_equal(a.particles(), b.particles())
    assert launches == b.steps_launched and launches >= b.stats()["num_bstep"]
    assert 0 < ms < 1e4
    assert len(na) == b.stats()["num_bstep"] and (na > 0).all()
    assert (fm > 0).all() and (om > 0).all() and fm.astype(np.float64).sum() <= ms * 1.0001 + 1e-6
    a.close()
    b.close()


def test_resident_exact_large_cluster_vs_oracle_inplace():
    """A single cluster of more than 1024 tiles (N = 262400) takes the
    compaction path that reads the slot records once the block time is
    known and counts its workgroups done in two levels (64 group counters,
    then the groups): its first 24 block steps, EXACT, are bit-identical to
    the oracle's in-place schedule from the same initial records."""
    order, n = 4, 1025 * 256
    eps2, eta, eta_s, dtmax = (4.0 / n) ** 2, 0.1, 0.01, 1 / 16
    m, x, v = H.plummer(n, 71)
    print(f"resident large cluster n={n}", flush=True)
    s = H.System(order, H.make_particles(order, m, x, v), 0.0, eps2, eta, eta_s, dtmax, variant="exact")
    s.init_step()
    P0 = s.ptcl()
    s.close()
    r = H.Resident(order, "exact", [(P0, eps2, eta, dtmax, 0.0, 0.0)])
    r.set_step_limit(24)
    with pytest.raises(H.StepLimitReached):
        r.run(dtmax, use_graph=True, batch=8)
    st = r.stats()
    o = O.System(order, P0, 0.0, eps2, eta, eta_s, dtmax)
    o.integrate_until_inplace(st["tsys"])
    print(f"24 block steps to t = {st['tsys']:.6e}, {st['num_step']} particle steps", flush=True)
    assert st["num_bstep"] == 24 and o.num_bstep == 24 and o.num_step == st["num_step"]
    assert np.array_equal(r.particles(), o.ptcl())
    r.close()


def test_resident_exact_salpeter_bitwise_vs_oracle_inplace():
    """Unequal (Salpeter) masses, 6th order: the resident EXACT schedule is
    bit-identical to the oracle's in-place schedule."""
    n, eps2, eta, eta_s = 512, (1 / 256) ** 2, 0.3, 0.1
    m, x, v = H.plummer(n, 51)
    m = H.salpeter(n, 5151)
    x = x - (m[:, None] * x).sum(axis=0)
    v = v - (m[:, None] * v).sum(axis=0)
    s = O.System(6, O.make_particles(6, m, x, v), 0.0, eps2, eta, eta_s, 1 / 16)
    s.init_step()
    r = H.Resident(6, "exact", [(s.ptcl(), eps2, eta, 1 / 16, 0.0, 0.0)])
    print("resident exact salpeter", flush=True)
    s.integrate_until_inplace(1 / 16)
    r.run(1 / 16, use_graph=True, batch=8)
    st = r.stats()
    assert (st["num_step"], st["num_bstep"]) == (s.num_step, s.num_bstep)
    assert np.array_equal(r.particles(), s.ptcl())
    r.close()


def test_resident_dtmin_clamp():
    """A lower step clamp: steps the quantised Aarseth criterion would make
    shorter are clamped to dt_min and counted; no particle steps below it."""
    order, n = 6, 512
    eps2, eta, eta_s = (1 / 256) ** 2, 0.3, 0.1
    o = initialised(order, n, 53, eps2, eta, eta_s, 1 / 16)
    P0 = o.ptcl()
    dtmin = float(np.median(P0[:, 3]))  # half the particles start below it
    r = H.Resident(order, "fp64", [(P0, eps2, eta, 1 / 16, dtmin, 0.0)])
    r.run(1 / 16)
    st = r.stats()
    P = r.particles()
    assert st["dtmin_hits"] > 0
    assert P[:, 3].min() >= dtmin
    assert st["tsys"] == 1 / 16
    r.close()


@pytest.mark.parametrize("n", [3000, 9000])
def test_resident_fp64_reduce_paths(n):
    """Both slice-sum widths (a wave per particle for steps of at most 4096
    active particles, 8 lanes beyond) against the EXACT run of
    the same cluster: the same schedule in the first steps and energies that
    agree to rounding."""
    order, eps2, eta, eta_s = 4, (4.0 / n) ** 2, 0.1, 0.01
    o = initialised(order, n, 57, eps2, eta, eta_s, 1 / 16)
    P0, e0 = o.ptcl(), o.init_energy
    res = {}
    for variant in ("exact", "fp64"):
        r = H.Resident(order, variant, [(P0, eps2, eta, 1 / 16, 0.0, 0.0)])
        r.set_fused(0)  # the reduce kernel's paths, not the two-launch step
        r.set_timing(True)
        r.run(1 / 32)
        na, _, _ = r.step_log()
        res[variant] = (r.stats(), r.energy(), na)
        r.close()
    (se, ee, nae), (sf, ef, naf) = res["exact"], res["fp64"]
    print(f"n={n}: largest step {nae.max()} active, steps {se['num_bstep']} / {sf['num_bstep']}", flush=True)
    if n > 4096:
        assert nae.max() > 4096  # the 8-lane path ran
    assert abs(sf["num_step"] / se["num_step"] - 1) < 0.01
    assert abs((ef - ee) / e0) < 1e-9


@pytest.mark.parametrize("use_graph", [False, True])
def test_resident_step_limit_stops_and_resumes(use_graph):
    """A block-step limit stops run() after the batch that reaches it, at a
    completed block step; removing the limit and running on gives the state
    of an uninterrupted run, bit for bit."""
    order, n, eps2 = 6, 500, (1 / 32) ** 2
    o = initialised(order, n, 61, eps2, 0.4, 0.1, 1 / 16)
    P0 = o.ptcl()
    print(f"resident step limit graph={use_graph}", flush=True)
    a = H.Resident(order, "fp64", [(P0, eps2, 0.4, 1 / 16, 0.0, 0.0)])
    a.run(1 / 8, use_graph=use_graph, batch=8)
    full = a.stats()["num_bstep"]
    assert full > 40
    b = H.Resident(order, "fp64", [(P0, eps2, 0.4, 1 / 16, 0.0, 0.0)])
    b.set_step_limit(20)
    with pytest.raises(H.StepLimitReached):
        b.run(1 / 8, use_graph=use_graph, batch=8)
    st = b.stats()
    assert 20 <= st["num_bstep"] <= 24 and st["tsys"] < 1 / 8
    b.set_step_limit(0)
    b.run(1 / 8, use_graph=use_graph, batch=8)
    assert a.stats() == b.stats()
    assert np.array_equal(a.particles(), b.particles())
    with pytest.raises(H.HHError):
        b.set_step_limit(-1)
    a.close()
    b.close()


@pytest.mark.parametrize("n", [2, 3, 65, 257])
def test_resident_edge_sizes(n):
    """Clusters of 2, 3, 65 and 257 particles (less than a wave, just over
    a wave and a tile): the EXACT resident schedule is bit-identical to the
    oracle's in-place schedule, and an ensemble of all four sizes gives each
    cluster what it gives alone (FP64)."""
    eps2, eta, eta_s = (1 / 32) ** 2, 0.4, 0.1
    o = initialised(6, n, 71, eps2, eta, eta_s, 1 / 16)
    r = H.Resident(6, "exact", [(o.ptcl(), eps2, eta, 1 / 16, 0.0, 0.0)])
    print(f"resident edge n={n}", flush=True)
    o.integrate_until_inplace(1 / 8)
    r.run(1 / 8, use_graph=True, batch=4)
    st = r.stats()
    assert (st["num_step"], st["num_bstep"], st["tsys"]) == (o.num_step, o.num_bstep, o.tsys)
    assert np.array_equal(r.particles(), o.ptcl())
    r.close()
    if n != 257:
        return
    clus = [initialised(6, m, 71, eps2, eta, eta_s, 1 / 16).ptcl() for m in (2, 3, 65, 257)]
    ens = H.Resident(6, "fp64", [(P, eps2, eta, 1 / 16, 0.0, 0.0) for P in clus])
    ens.run(1 / 8, use_graph=True, batch=4)
    for k, P in enumerate(clus):
        one = H.Resident(6, "fp64", [(P, eps2, eta, 1 / 16, 0.0, 0.0)])
        one.set_fused(0)  # the ensemble's launches, not the two-launch step
        one.run(1 / 8, use_graph=True, batch=4)
        assert one.stats() == ens.stats(k)
        assert np.array_equal(one.particles(), ens.particles(k))
        one.close()
    ens.close()


def test_plan_and_split_wave_settings_are_independent():
    """hhr_set_plan leaves the split-wave thresholds alone: setting them
    before or after the plan gives the same integration, bit for bit, and
    whole-tile tasks (thresholds 0) sum j in another order than the default
    split tiles."""
    order, n, eps2 = 6, 3000, (1 / 64) ** 2
    o = initialised(order, n, 67, eps2, 0.2, 0.1, 1 / 16)
    P0 = o.ptcl()
    print("plan / split-wave settings", flush=True)
    out = []
    for how in ("split-then-plan", "plan-then-split", "default"):
        r = H.Resident(order, "fp64", [(P0, eps2, 0.2, 1 / 16, 0.0, 0.0)])
        r.set_fused(0)  # the planned launches, not the two-launch step
        if how == "split-then-plan":
            r.set_split_waves(0, 0)
            r.set_plan()
        elif how == "plan-then-split":
            r.set_plan()
            r.set_split_waves(0, 0)
        r.run(1 / 32, use_graph=True, batch=8)
        out.append((r.stats(), r.particles()))
        r.close()
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
    assert not np.array_equal(out[0][1], out[2][1])


def test_plan_parameters_out_of_range_are_rejected():
    """hhr_set_plan refuses task targets above 8192 and j-parallel groups for
    more than 4096 active particles, and the context still runs."""
    order, n, eps2 = 4, 300, (1 / 64) ** 2
    o = initialised(order, n, 68, eps2, 0.1, 0.01, 1 / 16)
    r = H.Resident(order, "fp64", [(o.ptcl(), eps2, 0.1, 1 / 16, 0.0, 0.0)])
    for bad in ((4097, 256, 8192, 256), (64, 256, 8193, 256), (64, 256, 8192, 8193), (64, 0, 8192, 256)):
        with pytest.raises(H.HHError):
            r.set_plan(*bad)
    r.set_plan(4096, 256, 8192, 256)
    r.run(1 / 16)
    assert r.stats()["num_bstep"] > 0
    r.close()


def test_large_ensemble_bitwise_vs_oracle_inplace():
    """2100 clusters of 3 to 12 particles: more clusters than the plan
    kernel has threads (several per thread) and than the force and reduce
    kernels keep in LDS (their cluster lookups search the records in global
    memory).  Every cluster's EXACT integration is bit-identical to the
    oracle's in-place schedule of that cluster alone."""
    order, eps2, eta, eta_s = 6, (1 / 32) ** 2, 0.4, 0.1
    sizes = [3 + k % 10 for k in range(2100)]
    systems = [initialised(order, n, 1000 + k, eps2, eta, eta_s, 1 / 16) for k, n in enumerate(sizes)]
    print(f"resident ensemble of {len(sizes)} clusters", flush=True)
    r = H.Resident(order, "exact", [(o.ptcl(), eps2, eta, 1 / 16, 0.0, 0.0) for o in systems])
    r.run(1 / 16, use_graph=True, batch=16)
    for k, o in enumerate(systems):
        o.integrate_until_inplace(1 / 16)
        st = r.stats(k)
        assert (st["num_step"], st["num_bstep"], st["tsys"]) == (o.num_step, o.num_bstep, o.tsys), k
        assert np.array_equal(r.particles(k), o.ptcl()), k
    r.close()


def test_ensemble_planned_in_force_launch_bitwise():
    """Ensembles of up to 256 clusters are planned inside the force launch
    (every workgroup plans the step's tasks itself, one of them does the
    bookkeeping); larger ones by the plan kernel.  256 and 257 clusters of 3
    to 12 particles, EXACT: each of the first 256 clusters ends with the
    same records and counts in both ensembles, and a sample equals the
    oracle's in-place schedule of that cluster alone."""
    order, eps2, eta, eta_s = 6, (1 / 32) ** 2, 0.4, 0.1
    sizes = [3 + k % 10 for k in range(257)]
    systems = [initialised(order, n, 3000 + k, eps2, eta, eta_s, 1 / 16) for k, n in enumerate(sizes)]
    clus = [(o.ptcl(), eps2, eta, 1 / 16, 0.0, 0.0) for o in systems]
    print("resident ensembles of 256 and 257 clusters", flush=True)
    a = H.Resident(order, "exact", clus[:256])
    b = H.Resident(order, "exact", clus)
    a.run(1 / 16, use_graph=True, batch=16)
    b.run(1 / 16, use_graph=True, batch=16)
    for k in range(256):
        assert a.stats(k) == b.stats(k), k
        assert np.array_equal(a.particles(k), b.particles(k)), k
    for k in range(0, 257, 32):
        o = systems[k]
        o.integrate_until_inplace(1 / 16)
        st = b.stats(k)
        assert (st["num_step"], st["num_bstep"], st["tsys"]) == (o.num_step, o.num_bstep, o.tsys), k
        assert np.array_equal(b.particles(k), o.ptcl()), k
    a.close()
    b.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_ensemble_planned_in_force_launch_timing(use_graph):
    """The stamp-timed instances of an ensemble planned inside its force
    launch (the step counter advances during that launch) leave the
    integration bit-identical and log every block step."""
    order, eta, eta_s, dtmax = 4, 0.1, 0.01, 1 / 16
    clusters = []
    for k, n in enumerate([300, 700, 1024]):
        eps2 = (4.0 / n) ** 2
        o = initialised(order, n, 400 + k, eps2, eta, eta_s, dtmax)
        clusters.append((o.ptcl(), eps2, eta, dtmax, 0.0, 0.0))
    print(f"ensemble timing graph={use_graph}", flush=True)
    a = H.Resident(order, "fp64", clusters)
    b = H.Resident(order, "fp64", clusters)
    a.run(1 / 8, use_graph=use_graph, batch=8)
    b.set_timing(True, stamps=True)
    b.run(1 / 8, use_graph=use_graph, batch=8)
    ms, launches = b.force_timing()
    na, fm, om = b.step_log()
    for k in range(3):
        assert a.stats(k) == b.stats(k), k
        assert np.array_equal(a.particles(k), b.particles(k)), k
    assert launches == b.steps_launched
    nb = [b.stats(k)["num_bstep"] for k in range(3)]
    assert max(nb) <= len(na) <= sum(nb)
    assert (na > 0).all() and (fm > 0).all() and (om > 0).all()
    assert 0 < ms < 1e4
    a.close()
    b.close()


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("variant", ["fp64", "split"])
def test_fused_step_equals_planned_step(order, variant):
    """The two-launch block step of small clusters (hhr_set_fused), with
    either force kernel (i-slices, or the active list built in every
    workgroup), against the three-launch step (compaction, planned force,
    reduce + correct) on the same cluster: the same block steps and
    particle steps, positions and energies equal to rounding (only the j
    summation order differs), at N = 257 (ragged), 1024, 2048 and 4096 (the
    largest fused size; one, two and four j per lane, and the size's
    default); with graph replay and direct launches, and a step-limited run
    that resumes."""
    eps2, eta, eta_s = (1 / 64) ** 2, {4: 0.1, 6: 0.3, 8: 0.5}[order], {4: 0.01, 6: 0.1, 8: 0.1}[order]
    for n, jpl in ((257, 1), (1024, 2), (2048, 0), (4096, 4)):
        o = initialised(order, n, 81, eps2, eta, eta_s, 1 / 16)
        P0, e0 = o.ptcl(), o.init_energy
        out = {}
        for kind in ("islice", "list", "planned"):
            r = H.Resident(order, variant, [(P0, eps2, eta, 1 / 16, 0.0, 0.0)])
            try:
                if kind == "planned":
                    r.set_fused(0)
                else:
                    r.set_fused(4096, 0, jpl, 4096 if kind == "islice" else -1)
                    r.set_step_limit(7)
                    with pytest.raises(H.StepLimitReached):
                        r.run(1 / 16, use_graph=False, batch=4)
                    r.set_step_limit(0)
                r.run(1 / 16, use_graph=True, batch=8)
                out[kind] = (r.stats(), r.particles(), r.energy())
            finally:
                r.close()
        sp, pp, ep = out["planned"]
        for kind in ("islice", "list"):
            sf, pf, ef = out[kind]
            dx = np.abs(pf[:, 4:7] - pp[:, 4:7]).max()
            print(f"two-launch ({kind}) order={order} {variant} n={n}: block steps {sf['num_bstep']} / {sp['num_bstep']}, "
                  f"max |dx| {dx:.2e}, dE {abs((ef - ep) / e0):.2e}", flush=True)
            assert sf["tsys"] == sp["tsys"] == 1 / 16
            assert abs(sf["num_step"] / sp["num_step"] - 1) < 0.01
            assert dx < (1e-8 if variant == "fp64" else 1e-6)
            assert abs((ef - ep) / e0) < (1e-11 if variant == "fp64" else 1e-9)
