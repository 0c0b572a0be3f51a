"""How much per-pair error each force family tolerates, order by order.

The corrector reconstructs the higher derivatives from force differences
scaled by (2/dt)^k (crk..d5a, SRC/hermite6.h:168-173; d4a..d7a,
SRC/hermite8.h:203-211) and the quantised Aarseth step is computed from them
(SRC/hermite6.h:87-108).  Relative noise in the acceleration is amplified
most; noise in the jerk enters one power of dt later, snap two.  These tests
plant relative Gaussian noise of a chosen size in one force family at a time
(the oracle's forces, on the CPU) and integrate a Plummer sphere with the
parameters of the 6th-order N = 65536 workload (seed 2, eps 1/1024,
eta 0.2, dt_max 1/16, to t = 1/16) and of the 8th-order N = 2^18 workload
(seed 3, eps 1/4096, eta 0.4, dt_max 1/32, to t = 1/32), at N = 512.

Measured (block steps to the end time; "collapses" = the block time stops
advancing or the state turns non-finite):

  6th order, noise-free: 165 steps.
    acc 1e-7 collapses; acc 1e-9: 165.  jrk 1e-7: 167; 1e-6: 176; 1e-5: 306.
    snp 1e-5: 165; 1e-4: 177.
  8th order, noise-free: 29 steps.
    acc 1e-10 collapses; acc 1e-12: 29.  jrk 1e-9: 29; 1e-7: 64.
    snp 1e-7: 29; 1e-5: 64.  crk 1e-5: 30.  snp 1e-6 + crk 1e-5: 35.

An FP32 interaction body leaves relative errors of 1e-7 .. 1e-5 per
particle.  So the split-precision kernel keeps the acceleration (and, at
8th order, the jerk) in FP64 -- FP64 r^2, refined r^-1, m r^-3 and
accumulation -- and evaluates only the higher families in FP32.
"""
import numpy as np
import pytest

import hermite_amd as H
import oracle as O

C2 = dict(seed=2, eps=1 / 1024, eta=0.2, dtmax=1 / 16, tend=1 / 16)
C3 = dict(seed=3, eps=1 / 4096, eta=0.4, dtmax=1 / 32, tend=1 / 32)


def integrate_with_noise(order, noise, seed, eps, eta, dtmax, tend, n=512, steps_max=4000):
    """Block steps of the oracle's integration with relative noise
    noise[f] planted in force family f (acc, jrk, snp, crk); returns
    (reached the end with a finite state, block steps)."""
    m, x, v = H.plummer(n, seed)
    eps2 = eps ** 2
    s = O.System(order, O.make_particles(order, m, x, v), 0.0, eps2, eta, 0.1 if order > 4 else 0.01, dtmax)
    s.init_step()
    P = s.ptcl()
    rng = np.random.default_rng(0)
    etapow = eta ** (2 * (order - 3))
    scale = np.repeat(np.asarray(noise, float), 3)[: O.FSTRIDE[order]]
    tsys, nb = 0.0, 0
    while tsys < tend and nb < steps_max:
        tn = P[:, O.W_TLAST] + P[:, O.W_DT]
        tnext = tn.min()
        if not np.isfinite(tnext) or not tnext > tsys or not np.all(np.isfinite(P[:, 4:10])):
            return False, nb
        il = np.nonzero(tn == tnext)[0]
        dtlim = O.calc_dtlim(tnext, dtmax)
        F = O.force_list(order, O.predict(order, P, tnext), il, eps2)
        F *= 1 + scale * rng.normal(size=F.shape)
        for k, i in enumerate(il):
            P[i] = O.correct(order, P[i], F[k], etapow, dtlim)
        tsys, nb = tnext, nb + 1
    return bool(tsys >= tend and np.all(np.isfinite(P[:, 4:10]))), nb


@pytest.fixture(scope="module")
def baseline():
    return {6: integrate_with_noise(6, (0, 0, 0), **C2), 8: integrate_with_noise(8, (0, 0, 0, 0), **C3)}


def test_noise_free_runs(baseline):
    assert baseline[6] == (True, 165)
    assert baseline[8] == (True, 29)


def test_fourth_order_absorbs_fp32_noise_everywhere():
    ok, _ = integrate_with_noise(4, (1e-7, 1e-7), seed=2, eps=1 / 1024, eta=0.1, dtmax=1 / 16, tend=1 / 16)
    assert ok


# (order, noise per family, expected outcome: None = collapses, else the
#  largest allowed ratio of block steps to the noise-free run)
CASES = [
    (6, (1e-7, 0, 0), None),          # FP32-level acc error: fatal
    (6, (1e-9, 0, 0), 1.0),
    (6, (0, 1e-6, 0), 1.1),           # FP32 jrk: harmless
    (6, (0, 0, 1e-5), 1.0),           # FP32 snp: harmless
    (6, (1e-12, 1e-6, 1e-5), 1.1),    # the split kernel's error budget
    (8, (1e-10, 0, 0, 0), None),      # acc needs FP64 at 8th order
    (8, (1e-12, 0, 0, 0), 1.0),
    (8, (0, 1e-7, 0, 0), 3.0),        # FP32 jrk doubles the steps ...
    (8, (0, 1e-9, 0, 0), 1.0),        # ... so the jerk stays FP64 too
    (8, (0, 0, 1e-7, 0), 1.0),
    (8, (0, 0, 0, 1e-5), 1.1),        # FP32 crk: harmless
    (8, (1e-12, 1e-12, 1e-6, 1e-5), 1.25),  # the split kernel's error budget
]


@pytest.mark.parametrize("order,noise,ratio", CASES)
def test_force_noise_tolerance_per_family(baseline, order, noise, ratio):
    ok, nb = integrate_with_noise(order, noise, **(C2 if order == 6 else C3))
    if ratio is None:
        assert not ok, (order, noise, nb)
    else:
        assert ok, (order, noise, nb)
        assert nb <= ratio * baseline[order][1], (order, noise, nb, baseline[order][1])
