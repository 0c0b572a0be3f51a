"""The benchmark workloads (hermite_amd/workloads.py): sizes, seeds,
parameters and generated states of the five configurations."""
import math

import numpy as np

import hermite_amd as H
from hermite_amd import workloads as W


def test_fixed_configurations():
    c1, c2, c3, c5 = (W.workload(n) for n in ("c1", "c2", "c3", "c5"))
    assert (c1.order, c1.clusters[0].n, c1.clusters[0].seed, c1.clusters[0].eps) == (4, 1024, 1, 1 / 64)
    assert (c1.eta, c1.eta_s, c1.dtmax, c1.dtmin, c1.tend) == (0.1, 0.01, 1 / 16, 2.0 ** -30, 1.0)
    assert (c2.order, c2.clusters[0].n, c2.clusters[0].eps, c2.eta, c2.dtmax, c2.tend) == (6, 65536, 1 / 1024, 0.2, 1 / 16, 1.0)
    assert (c3.order, c3.clusters[0].n, c3.clusters[0].eps, c3.eta, c3.dtmax, c3.tend) == (8, 262144, 1 / 4096, 0.4, 1 / 32, 1 / 8)
    assert (c5.order, c5.clusters[0].n, c5.clusters[0].mass, c5.dtmin, c5.tend) == (6, 1 << 20, "salpeter", 2.0 ** -40, 1 / 16)


def test_ensemble_sizes_log_uniform():
    w = W.workload("c4")
    ns = np.array([c.n for c in w.clusters])
    assert len(ns) == 256 and ns.min() >= 256 and ns.max() <= 8192
    # log-uniform: the mean of ln N is near the midpoint, quartiles spread
    lo, hi = math.log(256), math.log(8192)
    assert abs(np.log(ns).mean() - (lo + hi) / 2) < 0.25
    assert len({c.seed for c in w.clusters}) == 256
    assert all(abs(c.eps - 4.0 / c.n) < 1e-15 for c in w.clusters)
    assert W.workload("c4", nclu=8).clusters == w.clusters[:8]


def test_salpeter_state():
    w = W.workload("c5")
    c = W.Cluster(4096, w.clusters[0].seed, w.clusters[0].eps, "salpeter", w.clusters[0].mass_seed)
    P = W.cluster_state(6, c)
    m = P[:, 1]
    assert abs(m.sum() - 1.0) < 1e-12
    assert m.max() / m.min() <= 1000.0 * (1 + 1e-12) and m.max() / m.min() > 100
    # Salpeter: N(>m) ~ m^-1.35, most of the number in the low-mass end
    assert np.median(m) < 2.5 * m.min()
    assert np.abs((m[:, None] * P[:, 4:7]).sum(axis=0)).max() < 1e-14
    assert np.abs((m[:, None] * P[:, 7:10]).sum(axis=0)).max() < 1e-14


def test_plummer_state_matches_generator():
    c = W.workload("c2").clusters[0]
    small = W.Cluster(512, c.seed, c.eps)
    P = W.cluster_state(6, small)
    m, x, v = H.plummer(512, c.seed)
    assert np.array_equal(P[:, 4:7], x) and np.array_equal(P[:, 7:10], v) and np.array_equal(P[:, 1], m)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message (and
    touches no device)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
