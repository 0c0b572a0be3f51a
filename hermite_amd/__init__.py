"""MI355X-native Hermite force + predictor back end.

The product is the C-ABI in ``include/hermite_hip.h`` implemented by
``libhermite_hip.so`` (hand-written gfx950 kernels) and, above it, the C++
drop-in ``Gravity`` (``include/hermite_hip_gravity.h``) and block-step driver.
This module is a thin ctypes mirror of the same ``Gravity`` interface
(``set_jp``, ``predict_all``, ``calc_force_on_first_nact``,
``calc_potential``; reference ``SRC/hermite6-titan.h:4-133``) for tests and
benchmarks.  It never falls back to a CPU path: without the built library or
a gfx950 device every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .build import HIP_LIB, HOST_LIB, build  # noqa: F401

# Order code 9 (ORDER_8DD) is the double-double 8th-order build
# (SRC/hermite8dd.h): positions as posH + posL.
ORDER_8DD = 9
PSTRIDE = {4: 24, 6: 28, 8: 36, 9: 40}   # reference Particle, doubles
FSTRIDE = {4: 6, 6: 9, 8: 12, 9: 12}     # reference Force, doubles
JPSTRIDE = {4: 14, 6: 20, 8: 26, 9: 29}  # device j-particle record, doubles
FLOPS = {4: 60, 6: 97, 8: 144, 9: 144}   # conventional flops per interaction (SRC/hermite{4,6,8}.h)

VARIANTS = {"fp64": 0, "split": 1, "exact": 2}

_dp = C.POINTER(C.c_double)


class HHError(RuntimeError):
    pass


class StepLimitReached(HHError):
    """Resident.run stopped at its block-step limit (HH_E_LIMIT); the state
    is that of a completed block step and a later run continues."""


_hip = None
_host = None


def hip_lib():
    """Load libhermite_hip.so (raises if it has not been built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise HHError(f"{HIP_LIB} is missing: run hermite_amd.build() (no CPU fallback exists)")
        L = C.CDLL(HIP_LIB)
        vp = C.c_void_p
        L.hh_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
        L.hh_destroy.argtypes = [vp]
        L.hh_set_jp.argtypes = [vp, C.c_int, _dp]
        L.hh_jstore.argtypes = [vp]
        L.hh_jstore.restype = _dp
        L.hh_mark_dirty.argtypes = [vp, C.c_int, C.c_int]
        L.hh_predict_all.argtypes = [vp, C.c_double]
        L.hh_force_first_nact.argtypes = [vp, C.c_int, C.c_double, _dp, _dp]
        L.hh_force_first_nact_static.argtypes = [vp, C.c_int, C.c_double, _dp, _dp]
        L.hhr_set_fused.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.hh_static_scratch_doubles.restype = C.c_longlong
        L.hh_potential.argtypes = [vp, C.c_double, _dp]
        L.hh_last_error.argtypes = [vp]
        L.hh_last_error.restype = C.c_char_p
        L.hh_set_tuning.argtypes = [vp, C.c_int, C.c_int]
        L.hh_set_tuning_ex.argtypes = [vp, C.c_int, C.c_int]
        ip = C.POINTER(C.c_int)
        L.hh_set_tuning_table.argtypes = [vp, C.c_int, ip, ip, ip, ip]
        L.hh_load_tuning_file.argtypes = [vp, C.c_char_p]
        L.hh_time_force.argtypes = [vp, C.c_int, C.c_double, C.c_int, _dp]
        L.hh_time_predict.argtypes = [vp, C.c_double, C.c_int, _dp]
        L.hh_force_async.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        L.hh_predict_async.argtypes = [vp, C.c_double]
        L.hh_synchronize.argtypes = [vp]
        L.hh_last_geometry.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.hh_eval_rsqrt.argtypes = [vp, C.c_int, _dp, _dp]
        L.hhr_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), _dp, _dp, C.POINTER(vp)]
        L.hhr_destroy.argtypes = [vp]
        L.hhr_run.argtypes = [vp, C.c_double, C.c_int, C.c_int]
        L.hhr_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        L.hhr_last_error.argtypes = [vp]
        L.hhr_last_error.restype = C.c_char_p
        L.hhr_last_run_ms.argtypes = [vp]
        L.hhr_last_run_ms.restype = C.c_double
        L.hhr_steps_launched.argtypes = [vp]
        L.hhr_steps_launched.restype = C.c_longlong
        L.hhr_get_particles.argtypes = [vp, C.c_int, _dp]
        L.hhr_energy.argtypes = [vp, C.c_int, _dp]
        L.hhr_set_timing.argtypes = [vp, C.c_int]
        L.hhr_set_plan.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.hhr_set_split_waves.argtypes = [vp, C.c_int, C.c_int]
        L.hhr_set_step_limit.argtypes = [vp, C.c_longlong]
        L.hhr_force_timing.argtypes = [vp, _dp, C.POINTER(C.c_longlong)]
        L.hhr_step_log.argtypes = [vp, C.c_longlong, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                   C.POINTER(C.c_float)]
        L.hhr_step_log.restype = C.c_longlong
        _hip = L
    return _hip


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise HHError(f"{HOST_LIB} is missing: run hermite_amd.build()")
        L = C.CDLL(HOST_LIB)
        L.hh_glibc_rand.argtypes = [C.c_uint, C.c_int, C.POINTER(C.c_int)]
        L.hh_plummer.argtypes = [C.c_int, C.c_uint, C.c_double, _dp, _dp, _dp]
        L.hh_salpeter.argtypes = [C.c_int, C.c_uint, C.c_double, C.c_double, C.c_double, _dp]
        vp = C.c_void_p
        L.hhd_create.restype = vp
        L.hhd_create.argtypes = [C.c_int, C.c_int, _dp] + [C.c_double] * 6 + [C.c_int, C.c_int]
        L.hhd_destroy.argtypes = [vp]
        L.hhd_status.argtypes = [vp]
        L.hhd_gravity.argtypes = [vp]
        L.hhd_gravity.restype = vp
        L.hhd_error.argtypes = [vp]
        L.hhd_error.restype = C.c_char_p
        for name in ("hhd_init_step", "hhd_integrate_one_block", "hhd_integrate_one_dtmax"):
            getattr(L, name).argtypes = [vp]
        L.hhd_integrate_until.argtypes = [vp, C.c_double]
        for name in ("hhd_energy", "hhd_init_energy", "hhd_tsys"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = C.c_double
        for name in ("hhd_num_step", "hhd_num_bstep", "hhd_dtmin_hits"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = C.c_long
        L.hhd_get_ptcl.argtypes = [vp, _dp]
        L.hhd_profile.argtypes = [vp, _dp]
        lp, ip = C.POINTER(C.c_long), C.POINTER(C.c_int)
        L.hh_read_snapshot.argtypes = [C.c_char_p, C.c_int, ip, lp, _dp, _dp]
        L.hh_write_snapshot.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_long, C.c_double, _dp, _dp]
        L.hh_read_dump.argtypes = [C.c_char_p, C.c_int, ip, lp, _dp, _dp, _dp, lp, lp, _dp]
        L.hh_write_dump.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_long, C.c_double, C.c_double, C.c_double,
                                    C.c_long, C.c_long, _dp]
        L.hh_io_error.restype = C.c_char_p
        _host = L
    return _host


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


# ------------------------------------------------------------------ inputs
def plummer(n: int, seed: int, rmax: float = 22.8):
    """Plummer sphere (mass, pos, vel) in Henon units, RUN/gen-plum.c method."""
    m = np.zeros(n)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    host_lib().hh_plummer(n, seed, rmax, _p(m), _p(x), _p(v))
    return m, x, v


def salpeter(n: int, seed: int, alpha: float = 2.35, mlo: float = 0.1, mhi: float = 100.0):
    m = np.zeros(n)
    host_lib().hh_salpeter(n, seed, alpha, mlo, mhi, _p(m))
    return m


# ------------------------------------------------------------------ formats
def _io(rc):
    if rc != 0:
        raise HHError(host_lib().hh_io_error().decode())


def read_snapshot(path: str, order: int):
    """Masaki snapshot (SRC/nbodysystem.h:379-403) -> (snapid, tsys, records)."""
    L = host_lib()
    sid, n, t = C.c_int(), C.c_long(), C.c_double()
    _io(L.hh_read_snapshot(path.encode(), order, C.byref(sid), C.byref(n), C.byref(t), None))
    P = np.zeros((n.value, PSTRIDE[order]))
    _io(L.hh_read_snapshot(path.encode(), order, C.byref(sid), C.byref(n), C.byref(t), _p(P)))
    return sid.value, t.value, P


def write_snapshot(path: str, order: int, snapid: int, tsys: float, P, pot):
    """write_snapshot_masaki (SRC/nbodysystem.h:359-378)."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    pot = np.ascontiguousarray(pot, dtype=np.float64)
    _io(host_lib().hh_write_snapshot(path.encode(), order, snapid, len(P), tsys, _p(P), _p(pot)))


def read_dump(path: str, order: int):
    """restore (SRC/nbodysystem.h:428-446) -> dict(dumpid, tsys, e0, eprev,
    nstep_tot, nbstep_tot, P)."""
    L = host_lib()
    did, n, t, e0, ep, a, b = C.c_int(), C.c_long(), C.c_double(), C.c_double(), C.c_double(), C.c_long(), C.c_long()
    args = [C.byref(x) for x in (did, n, t, e0, ep, a, b)]
    _io(L.hh_read_dump(path.encode(), order, *args, None))
    P = np.zeros((n.value, PSTRIDE[order]))
    _io(L.hh_read_dump(path.encode(), order, *args, _p(P)))
    return {"dumpid": did.value, "tsys": t.value, "e0": e0.value, "eprev": ep.value, "nstep_tot": a.value,
            "nbstep_tot": b.value, "P": P}


def write_dump(path: str, order: int, dumpid: int, tsys: float, e0: float, eprev: float, nstep_tot: int,
               nbstep_tot: int, P):
    """dump (SRC/nbodysystem.h:405-427)."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    _io(host_lib().hh_write_dump(path.encode(), order, dumpid, len(P), tsys, e0, eprev, nstep_tot, nbstep_tot,
                                 _p(P)))


def make_particles(order: int, mass, pos, vel, ids=None, tlast: float = 0.0) -> np.ndarray:
    """Reference Particle records (SRC/hermite6.h:37-47): derivatives zero."""
    n = len(mass)
    P = np.zeros((n, PSTRIDE[order]))
    idv = np.arange(n, dtype=np.int64) if ids is None else np.asarray(ids, dtype=np.int64)
    P[:, 0] = idv.view(np.float64)
    P[:, 1] = mass
    P[:, 2] = tlast
    P[:, 4:7] = pos
    if order == ORDER_8DD:  # posH = pos, posL = 0 (SRC/hermite8dd.h:64-78)
        P[:, 10:13] = vel
    else:
        P[:, 7:10] = vel
    return P


# ------------------------------------------------------------------ Gravity
class Gravity:
    """ctypes mirror of the reference's ``struct Gravity`` on the HIP back end.

    tuning: True loads the autotuned launch table for the nearest tuned N of
    this order and variant (hermite_amd/tuning/mi355x.json), False keeps the
    built-in geometry choice."""

    def __init__(self, order: int, nbody: int, variant: str = "fp64", device: int = 0, tuning: bool = True):
        L = hip_lib()
        self.order, self.nbody, self.variant = order, nbody, variant
        h = C.c_void_p()
        rc = L.hh_create(order, nbody, VARIANTS[variant], device, C.byref(h))
        if rc != 0:
            raise HHError(f"hh_create({order}, {nbody}, {variant}) failed ({rc}): "
                          f"{L.hh_last_error(None).decode()}")
        self.h = h
        self.tuning_rows = None
        if tuning and variant != "exact":
            from .autotune import table_for
            rows = table_for(order, variant, nbody)
            if rows:
                self.set_tuning_table(rows)
                self.tuning_rows = rows

    def _check(self, rc, what):
        if rc != 0:
            raise HHError(f"{what} failed ({rc}): {hip_lib().hh_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            hip_lib().hh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # set_jp(addr, Particle) -------------------------------------------------
    def set_jp(self, addr: int, rec: np.ndarray):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        self._check(hip_lib().hh_set_jp(self.h, addr, _p(rec)), "hh_set_jp")

    def set_jp_all(self, P: np.ndarray):
        """set_jp(i, P[i]) for every i, written straight into the pinned j-store."""
        n = len(P)
        js = np.ctypeslib.as_array(hip_lib().hh_jstore(self.h), shape=(self.nbody * JPSTRIDE[self.order],))
        js = js.reshape(self.nbody, JPSTRIDE[self.order])
        js[:n, 0:3] = P[:, 4:7]
        js[:n, 3] = P[:, 1]
        js[:n, 7] = P[:, 2]
        if self.order == ORDER_8DD:  # posL after posH in the Particle, last in the j-store
            js[:n, 4:7] = P[:, 10:13]
            js[:n, 8:26] = P[:, 13:31]
            js[:n, 26:29] = P[:, 7:10]
        else:
            js[:n, 4:7] = P[:, 7:10]
            js[:n, 8:14] = P[:, 10:16]
            if self.order >= 6:
                js[:n, 14:20] = P[:, 16:22]
            if self.order == 8:
                js[:n, 20:26] = P[:, 22:28]
        self._check(hip_lib().hh_mark_dirty(self.h, 0, n), "hh_mark_dirty")

    def predict_all(self, tsys: float):
        self._check(hip_lib().hh_predict_all(self.h, tsys), "hh_predict_all")

    def calc_force_on_first_nact(self, nact: int, eps2: float, with_pot: bool = False):
        F = np.zeros((nact, FSTRIDE[self.order]))
        pot = np.zeros(nact) if with_pot else None
        self._check(hip_lib().hh_force_first_nact(self.h, nact, eps2, _p(F), _p(pot) if with_pot else None),
                    "hh_force_first_nact")
        return (F, pot) if with_pot else F

    def calc_potential(self, eps2: float) -> np.ndarray:
        pot = np.zeros(self.nbody)
        self._check(hip_lib().hh_potential(self.h, eps2, _p(pot)), "hh_potential")
        return pot

    # measurement -------------------------------------------------------------
    def set_tuning(self, kernel: int = 0, jsplit: int = 0):
        self._check(hip_lib().hh_set_tuning(self.h, kernel, jsplit), "hh_set_tuning")

    def set_tuning_ex(self, i_per_thread: int = 0, split_fold: int = 0):
        self._check(hip_lib().hh_set_tuning_ex(self.h, i_per_thread, split_fold), "hh_set_tuning_ex")

    def set_tuning_table(self, rows):
        """rows: [(nact_hi, kernel, ni, jsplit), ...] ascending in nact_hi;
        an empty list restores the built-in choice."""
        n = len(rows)
        cols = [(C.c_int * max(n, 1))(*[int(r[k]) for r in rows]) for k in range(4)]
        self._check(hip_lib().hh_set_tuning_table(self.h, n, *cols), "hh_set_tuning_table")

    def load_tuning_file(self, path: str):
        self._check(hip_lib().hh_load_tuning_file(self.h, path.encode()), "hh_load_tuning_file")

    def time_force(self, nact: int, eps2: float, reps: int = 10) -> float:
        ms = C.c_double()
        self._check(hip_lib().hh_time_force(self.h, nact, eps2, reps, C.byref(ms)), "hh_time_force")
        return ms.value

    def time_predict(self, tsys: float, reps: int = 10) -> float:
        ms = C.c_double()
        self._check(hip_lib().hh_time_predict(self.h, tsys, reps, C.byref(ms)), "hh_time_predict")
        return ms.value

    def predict_async(self, tsys: float):
        self._check(hip_lib().hh_predict_async(self.h, tsys), "hh_predict_async")

    def force_async(self, nact: int, eps2: float, with_pot: bool = False):
        self._check(hip_lib().hh_force_async(self.h, nact, eps2, int(with_pot)), "hh_force_async")

    def synchronize(self):
        self._check(hip_lib().hh_synchronize(self.h), "hh_synchronize")

    def last_geometry(self):
        k, s, w = C.c_int(), C.c_int(), C.c_int()
        hip_lib().hh_last_geometry(self.h, C.byref(k), C.byref(s), C.byref(w))
        return {"kernel": k.value, "jsplit": s.value, "workgroups": w.value}

    def eval_rsqrt(self, x):
        """1/sqrt(x) as the FP64 kernels compute it, on this context's device."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._check(hip_lib().hh_eval_rsqrt(self.h, len(x), _p(x), _p(y)), "hh_eval_rsqrt")
        return y


# ------------------------------------------------------------------ driver
class System:
    """Block-timestep driver (the reference's NbodySystem semantics,
    SRC/nbodysystem.h) running on the HIP back end; native C++ in
    libhermite_host.so."""

    PHASES = ("force", "pot", "predict", "correct", "sort", "scan", "set_jp", "io", "total")

    def __init__(self, order, P, tsys, eps2, eta, eta_s, dtmax, dtmin=0.0, variant="fp64", device=0, tuning=True):
        L = host_lib()
        self.order, self.n = order, len(P)
        P = np.ascontiguousarray(P, dtype=np.float64)
        self.h = L.hhd_create(order, self.n, _p(P), tsys, eps2, eta, eta_s, dtmax, dtmin, VARIANTS[variant], device)
        if not self.h:
            raise HHError("hhd_create: bad arguments")
        self._check(L.hhd_status(self.h), "hhd_create")
        if tuning and variant != "exact":
            from .autotune import table_for
            rows = table_for(order, variant, self.n)
            if rows:
                n = len(rows)
                cols = [(C.c_int * n)(*[int(r[k]) for r in rows]) for k in range(4)]
                rc = hip_lib().hh_set_tuning_table(L.hhd_gravity(self.h), n, *cols)
                if rc != 0:
                    raise HHError(f"hh_set_tuning_table failed ({rc})")

    def _check(self, rc, what):
        if rc < 0:
            raise HHError(f"{what} failed ({rc}): {host_lib().hhd_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            host_lib().hhd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init_step(self):
        self._check(host_lib().hhd_init_step(self.h), "init_step")

    def integrate_one_block(self) -> int:
        return self._check(host_lib().hhd_integrate_one_block(self.h), "integrate_one_block")

    def integrate_one_dtmax(self):
        self._check(host_lib().hhd_integrate_one_dtmax(self.h), "integrate_one_dtmax")

    def integrate_until(self, tend):
        self._check(host_lib().hhd_integrate_until(self.h, tend), "integrate_until")

    def energy(self) -> float:
        e = host_lib().hhd_energy(self.h)
        self._check(host_lib().hhd_status(self.h), "energy")
        return e

    @property
    def init_energy(self):
        return host_lib().hhd_init_energy(self.h)

    @property
    def tsys(self):
        return host_lib().hhd_tsys(self.h)

    @property
    def num_step(self):
        return host_lib().hhd_num_step(self.h)

    @property
    def num_bstep(self):
        return host_lib().hhd_num_bstep(self.h)

    @property
    def dtmin_hits(self):
        return host_lib().hhd_dtmin_hits(self.h)

    def ptcl(self) -> np.ndarray:
        out = np.zeros((self.n, PSTRIDE[self.order]))
        host_lib().hhd_get_ptcl(self.h, _p(out))
        return out

    def profile(self) -> dict:
        out = np.zeros(len(self.PHASES))
        host_lib().hhd_profile(self.h, _p(out))
        return dict(zip(self.PHASES, out.tolist()))


# ------------------------------------------------------------------ resident
class Resident:
    """GPU-resident block-step integration of one or many independent clusters
    (hhr_* in libhermite_hip.so): schedule, prediction, force and corrector
    all on the device, optionally replayed from a hipGraph.

    clusters: list of (P, eps2, eta, dtmax, dtmin, tsys) where P are the
    cluster's Particle records after init_step (e.g. System(...).ptcl())."""

    def __init__(self, order, variant, clusters, device=0):
        L = hip_lib()
        self.order, self.variant = order, variant
        self.ns = [len(c[0]) for c in clusters]
        recs = np.ascontiguousarray(np.concatenate([np.asarray(c[0], dtype=np.float64) for c in clusters]))
        params = np.ascontiguousarray(np.array([[c[1], c[2], c[3], c[4], c[5]] for c in clusters], dtype=np.float64))
        ns = (C.c_int * len(self.ns))(*self.ns)
        h = C.c_void_p()
        rc = L.hhr_create(order, VARIANTS[variant], device, len(self.ns), ns, _p(recs), _p(params), C.byref(h))
        if rc != 0:
            raise HHError(f"hhr_create failed ({rc}): {L.hhr_last_error(None).decode()}")
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            raise HHError(f"{what} failed ({rc}): {hip_lib().hhr_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            hip_lib().hhr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, tend, use_graph=True, batch=64):
        rc = hip_lib().hhr_run(self.h, tend, int(use_graph), batch)
        if rc == -5:
            raise StepLimitReached(f"hhr_run: {hip_lib().hhr_last_error(self.h).decode()}")
        self._check(rc, "hhr_run")
        return hip_lib().hhr_last_run_ms(self.h)

    def set_step_limit(self, max_block_steps=0):
        """Make run() stop (StepLimitReached) once it has launched this many
        block steps without reaching tend; 0 removes the limit."""
        self._check(hip_lib().hhr_set_step_limit(self.h, int(max_block_steps)), "hhr_set_step_limit")

    def stats(self, k=0):
        a, b, d = C.c_longlong(), C.c_longlong(), C.c_longlong()
        t = C.c_double()
        self._check(hip_lib().hhr_stats(self.h, k, C.byref(a), C.byref(b), C.byref(t), C.byref(d)), "hhr_stats")
        return {"num_step": a.value, "num_bstep": b.value, "tsys": t.value, "dtmin_hits": d.value}

    def energy(self, k=0):
        e = C.c_double()
        self._check(hip_lib().hhr_energy(self.h, k, C.byref(e)), "hhr_energy")
        return e.value

    def particles(self, k=0):
        out = np.zeros((self.ns[k], PSTRIDE[self.order]))
        self._check(hip_lib().hhr_get_particles(self.h, k, _p(out)), "hhr_get_particles")
        return out

    @property
    def steps_launched(self):
        return hip_lib().hhr_steps_launched(self.h)

    def set_plan(self, jpar_max=64, smax=256, task_target=8192, jpar_target=256):
        """Work split of the persistent force kernel (see hhr_set_plan)."""
        self._check(hip_lib().hhr_set_plan(self.h, jpar_max, smax, task_target, jpar_target), "hhr_set_plan")

    def set_split_waves(self, wj4_max=8192, wj2_max=16384):
        """Active-count thresholds below which i-parallel tasks split each
        j-tile over 4 (2) waves (see hhr_set_split_waves)."""
        self._check(hip_lib().hhr_set_split_waves(self.h, wj4_max, wj2_max), "hhr_set_split_waves")

    def set_fused(self, max_n=4096, grid=0, jpl=0, islice_max=0):
        """Two-launch block steps for a single cluster of at most max_n
        particles (0 disables; grid, jpl and islice_max 0: defaults,
        islice_max -1: the active-list force kernel only; see
        hhr_set_fused)."""
        self._check(hip_lib().hhr_set_fused(self.h, int(max_n), int(grid), int(jpl), int(islice_max)),
                    "hhr_set_fused")

    def set_timing(self, on=True, stamps=False):
        """Time every block step (resets the totals).  stamps=False: HIP
        events around direct launches (run() then launches without a
        graph); stamps=True: the device clock written at every kernel's
        start, inside replayed graphs as well."""
        mode = (2 if stamps else 1) if on else 0
        self._check(hip_lib().hhr_set_timing(self.h, mode), "hhr_set_timing")

    def force_timing(self):
        """(summed force-kernel device ms, launches) since set_timing."""
        ms, n = C.c_double(), C.c_longlong()
        self._check(hip_lib().hhr_force_timing(self.h, C.byref(ms), C.byref(n)), "hhr_force_timing")
        return ms.value, n.value

    def step_log(self):
        """(active count, force ms, other ms) arrays of every timed step
        with active particles since set_timing."""
        L = hip_lib()
        n = L.hhr_step_log(self.h, 0, None, None, None)
        if n < 0:
            raise HHError(f"hhr_step_log failed ({n})")
        na = np.zeros(n, dtype=np.int32)
        fm = np.zeros(n, dtype=np.float32)
        om = np.zeros(n, dtype=np.float32)
        if n:
            L.hhr_step_log(self.h, n, na.ctypes.data_as(C.POINTER(C.c_int)),
                           fm.ctypes.data_as(C.POINTER(C.c_float)), om.ctypes.data_as(C.POINTER(C.c_float)))
        return na, fm, om
