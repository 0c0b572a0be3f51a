"""Pins of the CPU oracle that do not need the reference build: closed forms,
exact polynomial cases, time-derivative consistency and invariants.  Each test
is chosen so that a plausible slip in the restatement (a wrong Taylor or
corrector coefficient, a sign in the chain rule, a swapped operand) fails."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

ORDERS = (4, 6, 8)


def test_pow_one_nth_quant_closed_form():
    """2^floor(log2(x)/n) for every exponent and random significands
    (SRC/pownth.h:3-19)."""
    rng = np.random.default_rng(0)
    for n in (2, 6, 10):
        for e in range(-1000, 1001, 7):
            for m in (1.0, 1.5, 1.999999, 1.0 + rng.random()):
                x = math.ldexp(m, e)
                want = math.ldexp(1.0, math.floor(e / n))
                assert O.pow_one_nth_quant(n, x) == want, (n, x)


def test_invfordt_exact():
    for k in range(-60, 11):
        dt = math.ldexp(1.0, k)
        h, hinv = O.invfordt(dt)
        assert h == dt / 2 and hinv == 2 / dt


def taylor_particle(order, rng, t0):
    """A particle whose stored derivatives are exact dyadic numbers."""
    P = np.zeros(O.PSTRIDE[order])
    P[O.W_MASS] = 0.25
    P[O.W_TLAST] = t0
    words = {4: 22, 6: 28, 8: 28}[order]  # pos .. highest derivative used by the predictor
    P[O.W_POS:words] = rng.integers(-64, 64, size=words - O.W_POS) / 32.0
    return P


@pytest.mark.parametrize("order", ORDERS)
def test_predictor_is_the_taylor_polynomial(order):
    """Prediction equals the exact Taylor polynomial in rational arithmetic
    (to FP64 rounding), position, velocity, acceleration and jerk."""
    rng = np.random.default_rng(order)
    P = taylor_particle(order, rng, -0.375)
    tsys = 0.125
    Q = O.predict(order, P[None, :], tsys)[0]
    dt = Fraction(tsys) - Fraction(P[O.W_TLAST])
    nder = {4: 4, 6: 6, 8: 8}[order]  # x, v, a, j, (s, c, (d4, d5))
    c = [[Fraction(P[O.W_POS + 3 * d + k]) for d in range(nder)] for k in range(3)]

    def deriv(k, m):  # m-th derivative of coordinate k at dt
        return sum(c[k][d] * dt ** (d - m) / math.factorial(d - m) for d in range(m, nder))

    outs = [(0, 0), (4, 1)] + ([(7, 2)] if order >= 6 else []) + ([(10, 3)] if order == 8 else [])
    for off, m in outs:
        for k in range(3):
            exact = float(deriv(k, m))
            assert abs(Q[off + k] - exact) <= 4e-16 * max(1.0, abs(exact)), (order, off, k)
    assert Q[3] == P[O.W_MASS]


def random_pred(order, n, rng):
    Q = np.zeros((n, O.QSTRIDE[order]))
    Q[:, 0:3] = rng.normal(size=(n, 3))
    Q[:, 3] = rng.random(n) / n + 0.1 / n
    Q[:, 4:7] = rng.normal(size=(n, 3)) * 0.5
    if order >= 6:
        Q[:, 7:10] = rng.normal(size=(n, 3))
    if order == 8:
        Q[:, 10:13] = rng.normal(size=(n, 3)) * 2
    return Q


@pytest.mark.parametrize("order", ORDERS)
def test_force_vs_long_double_closed_form(order):
    rng = np.random.default_rng(10 + order)
    Q = random_pred(order, 96, rng)
    eps2 = 1e-3
    F = O.force(order, Q, 96, eps2)
    L = O.force_ld(order, Q, 96, eps2)
    for k in range(O.FSTRIDE[order] // 3):
        a, b = F[:, 3 * k:3 * k + 3], L[:, 3 * k:3 * k + 3]
        rel = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
        assert rel.max() < (1e-13 if k < 2 else 1e-11), (order, k, rel.max())


@pytest.mark.parametrize("order", ORDERS)
def test_chain_rule_by_time_differences(order):
    """Along Taylor trajectories x(t), the force output family k+1 is the
    time derivative of family k.  A five-point difference of the oracle's own
    outputs must reproduce jrk (and snp, crk) to O(h^4)."""
    rng = np.random.default_rng(20 + order)
    n = 12
    P = np.zeros((n, O.PSTRIDE[order]))
    P[:, O.W_MASS] = 1.0 / n
    nder = {4: 4, 6: 6, 8: 8}[order]  # Taylor terms the predictor uses
    for d, s in enumerate((1.0, 0.5, 0.3, 0.2, 0.1, 0.05, 0.02, 0.01)[:nder]):
        P[:, O.W_POS + 3 * d:O.W_POS + 3 * d + 3] = rng.normal(size=(n, 3)) * s
    eps2 = 0.05

    def F(t):
        return O.force(order, O.predict(order, P, t), n, eps2)

    h = 1e-3
    t0 = 0.01
    Fs = [F(t0 + k * h) for k in (-2, -1, 1, 2)]
    D = (Fs[0] - 8 * Fs[1] + 8 * Fs[2] - Fs[3]) / (12 * h)
    F0 = F(t0)
    nfam = O.FSTRIDE[order] // 3
    # family k+1 = d/dt family k needs the trajectory's (k+1)-th derivative
    # among the predicted inputs: vel for jrk, acc for snp, jrk for crk.
    for k in range(nfam - 1):
        num, ana = D[:, 3 * k:3 * k + 3], F0[:, 3 * k + 3:3 * k + 6]
        rel = np.linalg.norm(num - ana, axis=1) / np.linalg.norm(ana, axis=1)
        assert rel.max() < 1e-6, (order, k, rel.max())


@pytest.mark.parametrize("order", ORDERS)
def test_momentum_conservation(order):
    rng = np.random.default_rng(30 + order)
    n = 200
    Q = random_pred(order, n, rng)
    F = O.force(order, Q, n, 1e-4)
    m = Q[:, 3]
    for k in range(O.FSTRIDE[order] // 3):
        s = (m[:, None] * F[:, 3 * k:3 * k + 3]).sum(axis=0)
        scale = (m[:, None] * np.abs(F[:, 3 * k:3 * k + 3])).sum(axis=0)
        assert np.all(np.abs(s) <= 1e-13 * scale), (order, k)


def test_two_body_closed_form():
    """acc of a static pair: m / (r^2 + eps^2)^(3/2) along the separation."""
    Q = np.zeros((2, 10))
    Q[0, 3], Q[1, 3] = 0.3, 0.7
    Q[1, 0] = 2.0
    eps2 = 0.25
    F = O.force(6, Q, 2, eps2)
    assert F[0, 0] == pytest.approx(0.7 * 2.0 / (4.25 ** 1.5), rel=1e-15)
    assert F[1, 0] == pytest.approx(-0.3 * 2.0 / (4.25 ** 1.5), rel=1e-15)
    assert np.all(F[:, 1:3] == 0) and np.all(F[:, 3:] == 0)


@pytest.mark.parametrize("order", ORDERS)
def test_potential_closed_form(order):
    rng = np.random.default_rng(40 + order)
    n = 50
    P = np.zeros((n, O.PSTRIDE[order]))
    P[:, O.W_MASS] = rng.random(n)
    P[:, O.W_POS:O.W_POS + 3] = rng.normal(size=(n, 3))
    eps2 = 0.01
    pot = O.potential(order, P, eps2)
    x, m = P[:, O.W_POS:O.W_POS + 3], P[:, O.W_MASS]
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1) + eps2
    w = m[None, :] / np.sqrt(d2)
    np.fill_diagonal(w, 0.0)
    assert np.allclose(pot, -w.sum(1), rtol=1e-13, atol=0)


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("deg_off", [1, 2])
def test_corrector_exact_for_polynomial_force(order, deg_off):
    """The two-point Hermite corrector of order p integrates velocity exactly
    for an acceleration polynomial of degree p-1 (and then reconstructs its
    higher derivatives at the end of the step exactly), and position exactly
    for degree p-2 (SRC/hermite4.h:118-153, SRC/hermite6.h:136-184,
    SRC/hermite8.h:166-222)."""
    rng = np.random.default_rng(50 + order + deg_off)
    deg = order - deg_off
    a = [Fraction(int(v), 16) for v in rng.integers(-32, 32, size=(deg + 1))]  # a^(k)(0), one axis
    x0, v0 = Fraction(3, 8), Fraction(-5, 16)
    dt = Fraction(1, 2)  # reconstruction scales by (2/dt)^k: keep rounding small

    def acc_der(k, t):  # k-th derivative of acc at t
        return sum(a[d] * t ** (d - k) / math.factorial(d - k) for d in range(k, deg + 1))

    P = np.zeros(O.PSTRIDE[order])
    P[O.W_MASS] = 1.0
    P[O.W_DT] = float(dt)
    P[O.W_POS] = float(x0)
    P[O.W_VEL] = float(v0)
    nforce = O.FSTRIDE[order] // 3
    for k in range(nforce):
        P[O.W_ACC + 3 * k] = float(acc_der(k, Fraction(0)))
        P[O.W_ACC + 3 * k + 1] = 1.0  # keep the step estimator's norms non-zero
    f = np.zeros(O.FSTRIDE[order])
    for k in range(nforce):
        f[3 * k] = float(acc_der(k, dt))
        f[3 * k + 1] = 1.0
    out = O.correct(order, P, f, 1.0, 1.0 / 16)
    assert out[O.W_TLAST] == float(dt)
    if deg_off == 2:
        x_exact = x0 + v0 * dt + sum(a[d] * dt ** (d + 2) / math.factorial(d + 2) for d in range(deg + 1))
        assert out[O.W_POS] == pytest.approx(float(x_exact), rel=1e-14)
        return
    v_exact = v0 + sum(a[d] * dt ** (d + 1) / math.factorial(d + 1) for d in range(deg + 1))
    assert out[O.W_VEL] == pytest.approx(float(v_exact), rel=1e-14)
    first = O.W_ACC + 3 * nforce
    for k in range(nforce, deg + 1):
        got = out[first + 3 * (k - nforce)]
        want = float(acc_der(k, dt))
        assert got == pytest.approx(want, rel=1e-9, abs=1e-9), (order, k)


@pytest.mark.parametrize("order", ORDERS)
def test_aarseth_step_closed_form(order):
    rng = np.random.default_rng(60 + order)
    p = np.zeros(O.PSTRIDE[order])
    nvec = order  # acc .. highest stored derivative
    p[O.W_ACC:O.W_ACC + 3 * nvec] = rng.normal(size=3 * nvec)
    etapow = 0.3 ** (2 * (order - 3))
    s = [float(np.dot(p[O.W_ACC + 3 * k:O.W_ACC + 3 * k + 3], p[O.W_ACC + 3 * k:O.W_ACC + 3 * k + 3])) for k in range(nvec)]
    if order == 4:
        u, l, pp = math.sqrt(s[0] * s[2]) + s[1], math.sqrt(s[1] * s[3]) + s[2], 2
    elif order == 6:
        u, l, pp = math.sqrt(s[0] * s[2]) + s[1], math.sqrt(s[3] * s[5]) + s[4], 6
    else:
        u, l, pp = math.sqrt(s[0] * s[2]) + s[1], math.sqrt(s[5] * s[7]) + s[6], 10
    want = math.ldexp(1.0, math.floor(math.log2(etapow * u / l) / pp))
    assert O.aarseth_step_quant(order, p, etapow) == want


def test_calc_dtlim_brute_force():
    dtmax = 1.0 / 16
    for k in range(1, 4096):
        t = k * 2.0 ** -16
        want = dtmax
        while (t / want) != math.floor(t / want):
            want /= 2
        assert O.calc_dtlim(t, dtmax) == want


@pytest.mark.parametrize("order", ORDERS)
def test_driver_invariants(order):
    """Block-step driver: after every block the particles are grouped by
    ascending dt, the active set is the prefix with tlast+dt == tnext,
    times stay on the dt grid, and energy is conserved to integrator level."""
    import hermite_amd

    n = 64
    m, x, v = hermite_amd.plummer(n, 70 + order)
    P = O.make_particles(order, m, x, v)
    s = O.System(order, P, 0.0, (1 / 16) ** 2, {4: 0.1, 6: 0.3, 8: 0.3}[order], {4: 0.01, 6: 0.1, 8: 0.1}[order], 1 / 16)
    s.init_step()
    e0 = s.init_energy
    for _ in range(200):
        s.integrate_one_block()
        Pn = s.ptcl()
        dts = Pn[:, O.W_DT]
        assert np.all(np.diff(dts) >= 0)
        assert np.all(np.mod(Pn[:, O.W_TLAST], dts) == 0)
    for _ in range(4):
        s.integrate_one_dtmax()
    assert abs((s.energy() - e0) / e0) < 1e-5


def test_energy_error_falls_with_order():
    """Same Plummer sphere, same eta: the energy error of the 6th-order scheme
    is far below the 4th-order one and the 8th-order below the 6th."""
    import hermite_amd

    errs = {}
    for order in ORDERS:
        m, x, v = hermite_amd.plummer(64, 77)
        P = O.make_particles(order, m, x, v)
        s = O.System(order, P, 0.0, (1 / 16) ** 2, 0.1, 0.01 if order == 4 else 0.1, 1 / 16)
        s.init_step()
        for _ in range(8):
            s.integrate_one_dtmax()
        errs[order] = abs((s.energy() - s.init_energy) / s.init_energy)
    assert errs[6] < errs[4] / 100 and errs[8] < errs[6] / 10, errs
