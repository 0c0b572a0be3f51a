"""The oracle against the reference's own CPU build run live (oracle/_ref,
compiled from the reference sources).  Skipped where that build is absent;
tests/test_golden.py covers the same ground from committed vectors."""
import os

import numpy as np
import pytest

import hermite_amd as H
import oracle as O
from tests._state import hermite_state

ORDERS = (4, 6, 8)


@pytest.fixture(autouse=True)
def _need_ref(ref_available):
    if not ref_available:
        pytest.skip("oracle/_ref (reference CPU build) not present")


@pytest.mark.parametrize("order", ORDERS)
def test_pow_and_invfordt_match_reference(order):
    ref = O.Ref("cpu", order)
    rng = np.random.default_rng(1)
    for x in np.exp(rng.uniform(-60, 60, size=500)):
        for nn in (2, 6, 10):
            assert O.pow_one_nth_quant(nn, x) == ref.pow_one_nth_quant(nn, x)


@pytest.mark.parametrize("order", ORDERS)
def test_correct_and_init_dt_match_reference(order):
    ref = O.Ref("cpu", order)
    rng = np.random.default_rng(order)
    for _ in range(200):
        p = np.zeros(O.PSTRIDE[order])
        p[O.W_MASS] = 1e-3
        p[O.W_DT] = 2.0 ** -rng.integers(4, 20)
        p[O.W_POS:O.W_POS + 6] = rng.normal(size=6)
        p[O.W_ACC:O.W_ACC + 3 * order] = rng.normal(size=3 * order) * 10.0 ** rng.uniform(-2, 2, 3 * order)
        f = rng.normal(size=O.FSTRIDE[order])
        etapow = 0.2 ** (2 * (order - 3))
        a = O.correct(order, p, f, etapow, 1 / 16)
        b = ref.correct(p, f, 0.2, etapow, 1 / 16)
        assert np.array_equal(a, b)
        assert np.array_equal(O.init_dt(order, p, 0.1, 1 / 16), ref.init_dt(p, 0.1, 1 / 16))


@pytest.mark.parametrize("order", ORDERS)
def test_gravity_matches_reference_live(order):
    P = hermite_state(order, 500, seed=21)
    ref = O.Ref("cpu", order)
    eps2 = (1 / 256) ** 2
    for nact in (1, 100, 500):
        assert np.array_equal(O.force(order, O.predict(order, P, 0.0), nact, eps2),
                              ref.gravity_forces(P, 0.0, nact, eps2))
    assert np.array_equal(O.potential(order, P, eps2), ref.gravity_potential(P, eps2))


@pytest.mark.parametrize("order", ORDERS)
def test_driver_matches_reference_live(order):
    n = 256
    m, x, v = H.plummer(n, 31)
    P = O.make_particles(order, m, x, v)
    args = (0.0, (1 / 32) ** 2, {4: 0.1, 6: 0.4, 8: 0.6}[order], {4: 0.01, 6: 0.1, 8: 0.1}[order], 1 / 16)
    a = O.System(order, P, *args)
    b = O.Ref("cpu", order).system(P, *args)
    a.init_step()
    b.init_step()
    for _ in range(3):
        a.integrate_one_dtmax()
        b.integrate_one_dtmax()
        assert (a.num_step, a.num_bstep) == (b.num_step, b.num_bstep)
        assert np.array_equal(a.ptcl(), b.ptcl())
    assert a.energy() == b.energy()
