"""hh_force_first_nact_static: the force entry point that allocates nothing
(j-slice partial sums in a static device array).  Called through the C-ABI
and checked against hh_force_first_nact on the same context: bit-identical
when both launches use the same j-slice count, within the FP64 tolerance
(acc, jrk <= 1e-12, snp, crk <= 1e-10 per particle) when the scratch caps
the slice count.  Two calls in a row, the second with a smaller active
list, as the scratch keeps the previous call's contents."""
import ctypes as C

import numpy as np
import pytest

import hermite_amd as H
import oracle as O

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def call(fn, g, nact, eps2, order):
    F = np.zeros((nact, O.FSTRIDE[order]))
    pot = np.zeros(nact)
    rc = fn(g.h, nact, eps2, _p(F), _p(pot))
    assert rc == 0, H.hip_lib().hh_last_error(g.h)
    return F, pot, g.last_geometry()


def fam_err(F, G, order):
    return [np.max(np.linalg.norm(F[:, 3 * k:3 * k + 3] - G[:, 3 * k:3 * k + 3], axis=1)
                   / np.linalg.norm(G[:, 3 * k:3 * k + 3], axis=1)) for k in range(O.FSTRIDE[order] // 3)]


@pytest.mark.parametrize("order,variant", [(6, "fp64"), (6, "split"), (8, "fp64"), (4, "exact")])
def test_static_entry_matches_allocating_entry(order, variant):
    L = H.hip_lib()
    assert L.hh_static_scratch_doubles() == 4 << 20
    n = 65536  # large enough that the scratch caps the slice count of the all-pairs call
    eps2 = (1 / 1024) ** 2
    m, x, v = H.plummer(n, 21)
    P = O.make_particles(order, m, x, v)
    P[:, 10:O.PSTRIDE[order] - 4] = np.random.default_rng(21).normal(size=(n, O.PSTRIDE[order] - 14))
    P[:, O.W_TLAST] = -np.random.default_rng(22).random(n) / 64
    g = H.Gravity(order, n, variant, tuning=False)
    try:
        g.set_jp_all(P)
        g.predict_all(0.0)
        # default geometry: the scratch caps the slice count of the large call
        for nact in (n, 300, 3):
            print(f"static entry order={order} {variant} nact={nact}", flush=True)
            Fs, ps, geo_s = call(L.hh_force_first_nact_static, g, nact, eps2, order)
            Fa, pa, geo_a = call(L.hh_force_first_nact, g, nact, eps2, order)
            if geo_s == geo_a:
                assert np.array_equal(Fs, Fa) and np.array_equal(ps, pa), (nact, geo_s)
            else:
                assert geo_s["jsplit"] < geo_a["jsplit"], (geo_s, geo_a)
                e = fam_err(Fs, Fa, order)
                tol = [1e-12, 1e-12, 1e-10, 1e-10] if variant == "fp64" else [1e-11, 1e-5, 1e-4, 1e-4]
                assert all(x <= t for x, t in zip(e, tol)), (nact, e)
                assert np.max(np.abs(ps - pa) / np.abs(pa)) <= 1e-12
        # same slice count forced on both: identical bits
        if variant != "exact":
            for kernel, js, nact in ((1, 4, 2000), (2, 16, 5)):
                g.set_tuning(kernel, js)
                Fs, ps, geo_s = call(L.hh_force_first_nact_static, g, nact, eps2, order)
                Fa, pa, geo_a = call(L.hh_force_first_nact, g, nact, eps2, order)
                assert geo_s == geo_a and geo_s["jsplit"] == js
                assert np.array_equal(Fs, Fa) and np.array_equal(ps, pa), (kernel, js, nact)
    finally:
        g.close()
