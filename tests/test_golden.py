"""The oracle against vectors written by the reference's own CPU build
(tests/golden/ref_fixtures.npz, made by tests/golden/make_golden.py) and
against the known answers the reference ships in its run logs.

These run anywhere (no reference checkout needed)."""
import os

import numpy as np
import pytest

import hermite_amd as H
import oracle as O

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.npz"))
ORDERS = (4, 6, 8)


@pytest.mark.parametrize("order", ORDERS)
def test_force_matches_reference_generic_bitwise(order):
    P = FIX[f"P{order}"]
    F = O.force(order, O.predict(order, P, 0.0), len(P), (1 / 64) ** 2)
    assert np.array_equal(F, FIX[f"F{order}"])


@pytest.mark.parametrize("order", ORDERS)
def test_force_matches_reference_avx(order):
    """The AVX2 back end sums in another order and uses an FP32-seeded series
    rsqrt (SRC/v4df_util.h:45-72): agreement to rounding level only."""
    P = FIX[f"P{order}"]
    F = O.force(order, O.predict(order, P, 0.0), len(P), (1 / 64) ** 2)
    A = FIX[f"Favx{order}"]
    for k in range(O.FSTRIDE[order] // 3):
        rel = np.linalg.norm(F[:, 3 * k:3 * k + 3] - A[:, 3 * k:3 * k + 3], axis=1) / \
            np.linalg.norm(A[:, 3 * k:3 * k + 3], axis=1)
        assert rel.max() < 1e-12, (k, rel.max())


@pytest.mark.parametrize("order", ORDERS)
def test_potential_matches_reference_bitwise(order):
    P = FIX[f"P{order}"]
    assert np.array_equal(O.potential(order, P, (1 / 64) ** 2), FIX[f"pot{order}"])


@pytest.mark.parametrize("order", ORDERS)
def test_driver_run_matches_reference_bitwise(order):
    """init_step + 4 dt_max of the reference NbodySystem: identical step
    counts, energies and final particle state."""
    eta, eta_s, dtmax, eps2 = FIX[f"run{order}_params"]
    m, x, v = H.plummer(128, 5)
    s = O.System(order, O.make_particles(order, m, x, v), 0.0, eps2, eta, eta_s, dtmax)
    s.init_step()
    for _ in range(4):
        s.integrate_one_dtmax()
    assert s.num_step == FIX[f"run{order}_num_step"]
    assert s.num_bstep == FIX[f"run{order}_num_bstep"]
    assert s.init_energy == FIX[f"run{order}_e0"]
    assert s.energy() == FIX[f"run{order}_e1"]
    assert np.array_equal(s.ptcl(), FIX[f"run{order}_ptcl"])


def test_generator_reproduces_gen_plum():
    """RUN/gen-plum.c with its seed 19640916 and R < 100: bit-identical."""
    m, x, v = H.plummer(1024, 19640916, 100.0)
    G = FIX["genplum"]
    assert np.array_equal(G[:, 0], m) and np.array_equal(G[:, 1:4], x) and np.array_equal(G[:, 4:7], v)


# Known answers shipped by the reference: mean active-particle count per
# block step, nact = num_step / num_bstep, over t in [0, 1] and [1, 2] for
# pl1k.dat (gen-plum, N = 1024) with RUN/inp.{4th,6th,8th} (eta 0.1/0.5/0.75,
# eta_s 0.01/0.1/0.1, dt_max 1/16, eps = 4/N).  GPU and MIC runs agree on
# these to all printed digits:
#   4th  RUN/gpu-4th.log:33,35  RUN/mic-4th.log:2,4   5.489229e+01 5.392111e+01
#   6th  RUN/gpu-6th.log:33,35  RUN/mic-6th.log:2,4   7.783597e+01 7.474305e+01
#   8th  RUN/gpu-8th.log:40,42  RUN/mic-8th.log:2,4   9.150482e+01 8.746174e+01
NACT_KAT = {4: (0.1, 0.01, "5.489229e+01", "5.392111e+01"),
            6: (0.5, 0.1, "7.783597e+01", "7.474305e+01"),
            8: (0.75, 0.1, "9.150482e+01", "8.746174e+01")}


@pytest.mark.parametrize("order", ORDERS)
def test_nact_fingerprint_of_reference_logs(order):
    eta, eta_s, want1, want2 = NACT_KAT[order]
    n = 1024
    m, x, v = H.plummer(n, 19640916, 100.0)
    eps = 4.0 / n
    s = O.System(order, O.make_particles(order, m, x, v), 0.0, eps * eps, eta, eta_s, 1 / 16)
    s.init_step()
    while s.tsys < 1.0:
        s.integrate_one_dtmax()
    ns1, nb1 = s.num_step, s.num_bstep
    while s.tsys < 2.0:
        s.integrate_one_dtmax()
    assert "%e" % (ns1 / nb1) == want1
    assert "%e" % ((s.num_step - ns1) / (s.num_bstep - nb1)) == want2
