"""Why an FP32 interaction body cannot drive the 6th- and 8th-order schemes.

The corrector reconstructs the higher derivatives from force differences
scaled by (2/dt)^k (crk..d5a, SRC/hermite6.h:168-173; d4a..d7a,
SRC/hermite8.h:203-211) and the quantised Aarseth step is computed from them
(SRC/hermite6.h:87-108).  Random relative errors in the forces are amplified
by (2/dt)^5 (6th) or (2/dt)^7 (8th); the step then shrinks, which amplifies
them further.  This test plants relative Gaussian noise of a chosen size in
the oracle's forces (on the CPU) and integrates 1/16 time unit of a Plummer
sphere: 4th order absorbs 1e-7, 6th order absorbs 1e-9 but runs away at 1e-7,
the noise level of a per-pair FP32 body."""
import numpy as np
import pytest

import hermite_amd as H
import oracle as O


def integrate_with_noise(order, noise, n=512, eta=0.2, steps_max=4000):
    m, x, v = H.plummer(n, 2)
    eps2 = (1 / 1024) ** 2
    s = O.System(order, O.make_particles(order, m, x, v), 0.0, eps2, eta, 0.1 if order > 4 else 0.01, 1 / 16)
    s.init_step()
    P = s.ptcl()
    rng = np.random.default_rng(0)
    etapow = eta ** (2 * (order - 3))
    tsys, nb = 0.0, 0
    while tsys < 1 / 16 and nb < steps_max:
        tn = P[:, O.W_TLAST] + P[:, O.W_DT]
        tnext = tn.min()
        if not np.isfinite(tnext) or not tnext > tsys or not np.all(np.isfinite(P[:, 4:10])):
            return False, nb
        il = np.nonzero(tn == tnext)[0]
        dtlim = O.calc_dtlim(tnext, 1 / 16)
        F = O.force_list(order, O.predict(order, P, tnext), il, eps2)
        F *= 1 + noise * rng.normal(size=F.shape)
        for k, i in enumerate(il):
            P[i] = O.correct(order, P[i], F[k], etapow, dtlim)
        tsys, nb = tnext, nb + 1
    return tsys >= 1 / 16 and np.all(np.isfinite(P[:, 4:10])), nb


@pytest.mark.parametrize("order,noise,stable", [(4, 1e-7, True), (6, 1e-9, True), (6, 1e-7, False)])
def test_force_noise_tolerance(order, noise, stable):
    ok, nb = integrate_with_noise(order, noise, eta=0.1 if order == 4 else 0.2)
    assert ok == stable, (order, noise, nb)
