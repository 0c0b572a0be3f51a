"""Writes tests/golden/ref_fixtures.npz from the reference's own CPU build.

Run in the build container (it needs oracle/_ref, compiled from the
reference sources by oracle/Makefile):  python tests/golden/make_golden.py

Contents, per order o in (4, 6, 8):
  P{o}      input Particle records (N=128 Plummer, seed 5, derivatives from
            the reference generic force, higher derivatives seeded, tlast
            spread over [-1/64, 0])
  F{o}      reference generic back end force on all 128 (SRC/hermite{o}.h)
  Favx{o}   reference AVX2 back end force on all 128 (SRC/hermite{o}-avx.h)
  pot{o}    reference generic potential
  run{o}_*  reference NbodySystem (SRC/nbodysystem.h) from the same Plummer
            sphere: init_step + 4 dtmax (t = 0 -> 1/4), eps = 1/16:
            e0, e1, num_step, num_bstep and the final Particle records.
  genplum   m, x, y, z, vx, vy, vz of RUN/gen-plum.c's pl1k.dat (N=1024)

Double-double 8th order (order code 9, SRC/hermite8dd.h; the reference's
driver does not build for it, so its generic back end and Particle only):
  P9        the order-8 state P8 moved by sqrt(3) 2^32 in every coordinate,
            positions split into posH + posL by an exact two-sum
  F9, pot9  reference double-double force on all 128 and potential
  corr9_p, corr9_f, corr9_out   Particle::correct on 32 seeded records
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import hermite_amd as H  # noqa: E402
import oracle as O  # noqa: E402


def ref_state(order, n=128, seed=5, eps2=(1 / 64) ** 2):
    ref = O.Ref("cpu", order)
    m, x, v = H.plummer(n, seed)
    P = O.make_particles(order, m, x, v)
    for _ in range(2 if order > 4 else 1):
        F = ref.gravity_forces(P, 0.0, n, eps2)
        P[:, O.W_ACC:O.W_ACC + F.shape[1]] = F
    rng = np.random.default_rng(seed)
    nf = O.FSTRIDE[order] // 3
    base = O.W_ACC + 3 * nf
    scale = np.linalg.norm(P[:, O.W_ACC:O.W_ACC + 3], axis=1)[:, None]
    for d in range({4: 2, 6: 3, 8: 4}[order]):
        P[:, base + 3 * d:base + 3 * d + 3] = rng.normal(size=(n, 3)) * scale * (4.0 ** (d + 1))
    P[:, O.W_TLAST] = -(1 / 64) * rng.random(n)
    return P


DD_OFFSET = np.sqrt(3.0) * 2.0 ** 32


def two_sum(a, b):
    """Knuth's exact sum: a + b = s + e (SRC/hermite8dd.h:26-35)."""
    s = a + b
    bb = s - a
    e = (a - (s - bb)) + (b - bb)
    return s, e


def to_dd(P8, offset=DD_OFFSET):
    """Order-8 Particle records -> double-double records at x + offset."""
    n = len(P8)
    P = np.zeros((n, O.PSTRIDE[9]))
    P[:, 0:4] = P8[:, 0:4]
    hi, lo = two_sum(np.full((n, 3), offset), P8[:, O.W_POS:O.W_POS + 3])
    P[:, 4:7], P[:, 7:10] = hi, lo
    P[:, 10:37] = P8[:, O.W_VEL:O.W_VEL + 27]
    return P


def main():
    out = {}
    eps2 = (1 / 64) ** 2
    for order in (4, 6, 8):
        ref = O.Ref("cpu", order)
        P = ref_state(order)
        out[f"P{order}"] = P
        out[f"F{order}"] = ref.gravity_forces(P, 0.0, len(P), eps2)
        out[f"Favx{order}"], _ = O.RefAVX(order).forces(P, 0.0, len(P), eps2)
        out[f"pot{order}"] = ref.gravity_potential(P, eps2)
        m, x, v = H.plummer(128, 5)
        P0 = O.make_particles(order, m, x, v)
        eta = {4: 0.1, 6: 0.3, 8: 0.4}[order]
        eta_s = {4: 0.01, 6: 0.1, 8: 0.1}[order]
        s = ref.system(P0, 0.0, (1 / 16) ** 2, eta, eta_s, 1 / 16)
        s.init_step()
        for _ in range(4):
            s.integrate_one_dtmax()
        out[f"run{order}_e0"] = np.float64(s.init_energy)
        out[f"run{order}_e1"] = np.float64(s.energy())
        out[f"run{order}_num_step"] = np.int64(s.num_step)
        out[f"run{order}_num_bstep"] = np.int64(s.num_bstep)
        out[f"run{order}_ptcl"] = s.ptcl()
        out[f"run{order}_params"] = np.array([eta, eta_s, 1 / 16, (1 / 16) ** 2])
    ref9 = O.Ref("cpu", 9)
    P9 = to_dd(out["P8"])
    out["P9"] = P9
    out["F9"] = ref9.gravity_forces(P9, 0.0, len(P9), eps2)
    out["pot9"] = ref9.gravity_potential(P9, eps2)
    rng = np.random.default_rng(9)
    cp, cf, co = [], [], []
    for i in range(32):
        p = P9[i].copy()
        p[O.W_TLAST] = 0.0
        p[O.W_DT] = 2.0 ** -int(rng.integers(4, 20))
        f = rng.normal(size=12) * np.linalg.norm(p[13:16])
        cp.append(p)
        cf.append(f)
        co.append(ref9.correct(p, f, 0.4, 0.4 ** 10, 1 / 16))
    out["corr9_p"], out["corr9_f"], out["corr9_out"] = np.array(cp), np.array(cf), np.array(co)
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([os.path.join(O.REF_DIR, "gen-plum"), "1"], cwd=td, check=True, capture_output=True)
        rows = open(os.path.join(td, "pl1k.dat")).read().split("\n")[3:3 + 1024]
        out["genplum"] = np.array([[float.fromhex(t) for t in r.split()[1:8]] for r in rows])
    np.savez_compressed(os.path.join(HERE, "ref_fixtures.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_fixtures.npz"))


if __name__ == "__main__":
    main()
