"""Double-double 8th order (order code 9: SRC/hermite8dd.h, the EIGHTHDD build
of SRC/compile.cpp:47-48): positions as posH + posL, relative positions
(posHj - posHi) + (posLj - posLi), position increments added by
dd_accumulate.

CPU: the oracle against vectors of the reference's double-double build
(tests/golden/ref_fixtures.npz; live against oracle/_ref/ref_cpu9.so where it
is built), translation invariance far from the origin, the dump format.
GPU: EXACT bit-identical to the oracle (force, potential, host-driver
integration), FP64 within rounding, and the offset cluster that plain 8th
order cannot resolve.
"""
import os
import subprocess

import numpy as np
import pytest

import hermite_amd as H
import oracle as O
from tests._state import hermite_state

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "ref_fixtures.npz"))
OFFSET = np.sqrt(3.0) * 2.0 ** 32
EPS2 = (1 / 64) ** 2


def two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def to_dd(P8, offset=OFFSET):
    n = len(P8)
    P = np.zeros((n, 40))
    P[:, 0:4] = P8[:, 0:4]
    P[:, 4:7], P[:, 7:10] = two_sum(np.full((n, 3), offset), P8[:, 4:7])
    P[:, 10:37] = P8[:, 7:34]
    return P


def rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


# ---------------------------------------------------------------- CPU
def test_oracle_force_potential_correct_match_reference_vectors():
    P = FIX["P9"]
    assert np.array_equal(O.force(9, O.predict(9, P, 0.0), len(P), EPS2), FIX["F9"])
    assert np.array_equal(O.potential(9, P, EPS2), FIX["pot9"])
    for p, f, out in zip(FIX["corr9_p"], FIX["corr9_f"], FIX["corr9_out"]):
        assert np.array_equal(O.correct(9, p, f, 0.4 ** 10, 1 / 16), out)


def test_oracle_matches_reference_live(ref_available):
    if not ref_available or not os.path.exists(os.path.join(O.REF_DIR, "ref_cpu9.so")):
        pytest.skip("oracle/_ref/ref_cpu9.so not built")
    ref = O.Ref("cpu", 9)
    P = to_dd(hermite_state(8, 300, seed=23))
    for nact in (1, 5, 300):
        assert np.array_equal(O.force(9, O.predict(9, P, 0.0), nact, EPS2), ref.gravity_forces(P, 0.0, nact, EPS2))
    assert np.array_equal(O.potential(9, P, EPS2), ref.gravity_potential(P, EPS2))
    rng = np.random.default_rng(3)
    for i in range(50):
        p = P[i].copy()
        p[2], p[3] = 0.0, 2.0 ** -int(rng.integers(3, 25))
        f = rng.normal(size=12)
        assert np.array_equal(O.correct(9, p, f, 0.3 ** 10, 1 / 32), ref.correct(p, f, 0.3, 0.3 ** 10, 1 / 32))
        assert np.array_equal(O.init_dt(9, p, 0.1, 1 / 16), ref.init_dt(p, 0.1, 1 / 16))


def test_translation_far_from_origin():
    """At sqrt(3) 2^32 from the origin a double resolves positions to 2^-20;
    the hi + lo pair keeps the cluster's forces (relative to the same cluster
    at the origin), while one double per coordinate loses them."""
    P8 = hermite_state(8, 256, seed=29)
    F0 = O.force(8, O.predict(8, P8, 0.0), 256, EPS2)
    Fdd = O.force(9, O.predict(9, to_dd(P8), 0.0), 256, EPS2)
    P8far = P8.copy()
    P8far[:, 4:7] = P8far[:, 4:7] + OFFSET
    Ffar = O.force(8, O.predict(8, P8far, 0.0), 256, EPS2)
    assert rel(Fdd[:, 0:3], F0[:, 0:3]).max() < 1e-12
    assert rel(Fdd[:, 3:6], F0[:, 3:6]).max() < 1e-10
    assert rel(Ffar[:, 0:3], F0[:, 0:3]).max() > 1e-6


def test_dd_accumulate_keeps_the_sum():
    """The corrector's position update: posH + posL advances by the
    increment to double-double accuracy (closed form of the pair sum)."""
    P = to_dd(hermite_state(8, 16, seed=31))
    rng = np.random.default_rng(5)
    for i in range(16):
        p = P[i].copy()
        p[2], p[3] = 0.0, 2.0 ** -8
        q = O.correct(9, p, rng.normal(size=12), 0.3 ** 10, 1 / 16)
        # exact increment from the corrector's velocity terms is not needed:
        # the pair stays normalised (|lo| <= ulp(hi) / 2) and hi + lo changes
        # by an amount far below ulp(hi)
        assert np.all(np.abs(q[7:10]) <= np.spacing(q[4:7]) / 2)
        moved = (q[4:7] - p[4:7]) + (q[7:10] - p[7:10])
        assert np.all(np.abs(moved) < 1.0)


def test_dd_dump_rewrites_identically(tmp_path):
    P = to_dd(hermite_state(8, 8, seed=3))
    a = tmp_path / "a.dat"
    H.write_dump(str(a), H.ORDER_8DD, 4, 0.5, -0.25, -0.2500001, 123, 45, P)
    head = a.read_text().splitlines()[0]
    assert head == "8 4"  # the file's order is Particle::order, 8
    d = H.read_dump(str(a), H.ORDER_8DD)
    b = tmp_path / "b.dat"
    H.write_dump(str(b), H.ORDER_8DD, d["dumpid"], d["tsys"], d["e0"], d["eprev"], d["nstep_tot"], d["nbstep_tot"],
                 d["P"])
    assert a.read_text() == b.read_text()
    Q = P.copy()
    Q[:, 2] = 0.5  # tlast = tsys on restore
    Q[:, 3] = 0.0
    Q[:, 37:40] = 0.0
    assert np.array_equal(d["P"], Q)
    with pytest.raises(H.HHError):
        H.read_dump(str(a), 6)


# ---------------------------------------------------------------- GPU
def gpu_force(P, nact, variant, eps2=EPS2, pot=False):
    g = H.Gravity(H.ORDER_8DD, len(P), variant)
    try:
        g.set_jp_all(P)
        g.predict_all(0.0)
        return g.calc_force_on_first_nact(nact, eps2, with_pot=pot)
    finally:
        g.close()


@pytest.mark.gpu
def test_gpu_exact_bitwise_on_reference_vectors():
    P = FIX["P9"]
    print("dd exact force on the reference vectors", flush=True)
    assert np.array_equal(gpu_force(P, len(P), "exact"), FIX["F9"])
    g = H.Gravity(H.ORDER_8DD, len(P), "exact")
    try:
        g.set_jp_all(P)
        assert np.array_equal(g.calc_potential(EPS2), FIX["pot9"])
    finally:
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1000, 1536])
def test_gpu_dd_forces_ragged(n):
    P = to_dd(hermite_state(8, n, seed=41))
    P[:, 2] = -np.linspace(0.0, 1 / 64, n)  # prediction over a spread of dt
    R = O.force(9, O.predict(9, P, 0.0), n, EPS2)
    for nact in (1, 7, 64, 256, n):
        print(f"dd n={n} nact={nact}", flush=True)
        assert np.array_equal(gpu_force(P, nact, "exact"), R[:nact])
        F = gpu_force(P, nact, "fp64")
        for k, tol in ((0, 1e-12), (1, 1e-12), (2, 1e-10), (3, 1e-10)):
            assert rel(F[:, 3 * k:3 * k + 3], R[:nact, 3 * k:3 * k + 3]).max() < tol, (nact, k)


@pytest.mark.gpu
def test_gpu_dd_far_cluster_matches_cluster_at_origin():
    P8 = hermite_state(8, 512, seed=43)
    F0 = gpu_force_plain(P8)
    Fdd = gpu_force(to_dd(P8), 512, "fp64")
    assert rel(Fdd[:, 0:3], F0[:, 0:3]).max() < 1e-12


def gpu_force_plain(P8):
    g = H.Gravity(8, len(P8), "fp64")
    try:
        g.set_jp_all(P8)
        g.predict_all(0.0)
        return g.calc_force_on_first_nact(len(P8), EPS2)
    finally:
        g.close()


@pytest.mark.gpu
def test_gpu_dd_host_driver_bitwise_vs_oracle():
    """init_step + 4 dt_max intervals of the host driver with the EXACT
    double-double back end against the oracle's driver, bit for bit."""
    n = 256
    m, x, v = H.plummer(n, 47)
    P8 = H.make_particles(8, m, x, v)
    P = to_dd(P8)
    args = (0.0, (1 / 32) ** 2, 0.6, 0.1, 1 / 16)
    a = O.System(9, P, *args)
    a.init_step()
    s = H.System(H.ORDER_8DD, P, *args, variant="exact")
    try:
        print("dd host driver init_step", flush=True)
        s.init_step()
        assert s.init_energy == a.init_energy
        for _ in range(4):
            a.integrate_one_dtmax()
            s.integrate_one_dtmax()
            assert (s.num_step, s.num_bstep) == (a.num_step, a.num_bstep)
        assert np.array_equal(s.ptcl(), a.ptcl())
        assert s.energy() == a.energy()
    finally:
        s.close()


@pytest.mark.gpu
def test_gpu_dd_rejects_split_and_resident():
    with pytest.raises(H.HHError):
        H.Gravity(H.ORDER_8DD, 64, "split")
    P = to_dd(hermite_state(8, 64, seed=3))
    with pytest.raises(H.HHError):
        H.Resident(H.ORDER_8DD, "fp64", [(P, EPS2, 0.4, 1 / 16, 0.0, 0.0)])


@pytest.mark.gpu
def test_hermite_run_order_8dd(tmp_path):
    """The command-line driver with "order 8dd": runs from a snapshot far
    from the origin and writes a double-double dump."""
    n = 64
    m, x, v = H.plummer(n, 11)
    P = H.make_particles(8, m, x + OFFSET, v)
    H.write_snapshot(str(tmp_path / "snap0.dat"), 8, 0, 0.0, P, np.zeros(n))
    inp = ("order 8dd\neta 0.75\neta_s 0.1\nsnapin snap0.dat\ndumpout_base d%03d.dat\ntend 0.125\ndtmaxinv 16\n"
           "epsinv 64\nlog_interval 0.0625\ndump_interval 0.125\nvariant exact\n")
    runner = os.path.join(os.path.dirname(HERE), "hermite_amd", "hermite_run")
    r = subprocess.run([runner], input=inp, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "order : 8" in r.stdout
    d = H.read_dump(str(tmp_path / "d001.dat"), H.ORDER_8DD)
    assert d["tsys"] == 0.125 and len(d["P"]) == n
