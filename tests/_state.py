"""Seeded Hermite states for the parity tests (Plummer spheres with
bootstrapped derivatives), built with the product's generator and the oracle."""
import numpy as np

import hermite_amd
import oracle

W_TLAST, W_ACC = 2, 10


def hermite_state(order: int, n: int, seed: int = 7, eps: float = 1 / 64, spread: float = 1 / 64, rmax=22.8):
    """Particle records of a Plummer sphere whose stored derivatives come from
    the oracle's force (two bootstrap passes for order > 4, as in
    SRC/nbodysystem.h:573-581), with higher derivatives of matching scale and
    tlast spread over [-spread, 0] so that prediction to t = 0 is non-trivial."""
    m, x, v = hermite_amd.plummer(n, seed, rmax)
    P = oracle.make_particles(order, m, x, v)
    eps2 = eps * eps
    for _ in range(2 if order > 4 else 1):
        F = oracle.force(order, oracle.predict(order, P, 0.0), n, eps2)
        for i in range(n):
            oracle.lib().orc_assign_force(order, oracle._ptr(P[i]), oracle._ptr(np.ascontiguousarray(F[i])))
    rng = np.random.default_rng(seed)
    nder = {4: 2, 6: 3, 8: 4}[order]  # derivatives beyond the force words
    base = W_ACC + 3 * {4: 2, 6: 3, 8: 4}[order]
    scale = np.linalg.norm(P[:, W_ACC:W_ACC + 3], axis=1)[:, None]
    for d in range(nder):
        P[:, base + 3 * d: base + 3 * d + 3] = rng.normal(size=(n, 3)) * scale * (4.0 ** (d + 1))
    P[:, W_TLAST] = -spread * rng.random(n)
    return P
