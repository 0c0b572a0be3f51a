"""CPU oracle for the Hermite force + predictor path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package; the product package ``hermite_amd`` never does.

``liboracle.so`` is the C restatement in ``hermite_oracle.c`` (each function
cites the reference file:line it follows).  ``_ref/`` holds the reference's own
CPU back ends compiled from its sources by ``oracle/Makefile`` (generic FP64,
bit-comparable; AVX2 with the reference's flags, the timed CPU baseline).

Record layouts are documented at the top of ``hermite_oracle.c``; arrays are
numpy float64 of shape (n, stride), the particle id stored as int64 bits in
word 0.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

# order 9: the double-double 8th-order build (SRC/hermite8dd.h)
PSTRIDE = {4: 24, 6: 28, 8: 36, 9: 40}
QSTRIDE = {4: 7, 6: 10, 8: 13, 9: 17}
FSTRIDE = {4: 6, 6: 9, 8: 12, 9: 12}
FLOPS = {4: 60, 6: 97, 8: 144, 9: 144}

# word offsets inside a Particle record
W_ID, W_MASS, W_TLAST, W_DT, W_POS, W_VEL, W_ACC, W_JRK, W_SNP, W_CRK = 0, 1, 2, 3, 4, 7, 10, 13, 16, 19
W_D4A, W_D5A, W_D6A, W_D7A = 22, 25, 28, 31

_dp = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
    if os.path.isdir("/root/reference/SRC"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_pow_one_nth_quant.restype = C.c_double
        L.orc_pow_one_nth_quant.argtypes = [C.c_int, C.c_double]
        L.orc_invfordt.argtypes = [C.c_double, _dp, _dp]
        L.orc_predict.argtypes = [C.c_int, C.c_int, _dp, C.c_double, _dp]
        L.orc_force.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_double, _dp]
        L.orc_force_ld.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_double, _dp]
        L.orc_force_list.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.POINTER(C.c_int), C.c_double, _dp]
        L.orc_potential.argtypes = [C.c_int, C.c_int, _dp, C.c_double, _dp]
        L.orc_aarseth_step_quant.restype = C.c_double
        L.orc_aarseth_step_quant.argtypes = [C.c_int, _dp, C.c_double]
        L.orc_assign_force.argtypes = [C.c_int, _dp, _dp]
        L.orc_init_dt.argtypes = [C.c_int, _dp, C.c_double, C.c_double]
        L.orc_correct.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double]
        L.orc_calc_dtlim.restype = C.c_double
        L.orc_calc_dtlim.argtypes = [C.c_double, C.c_double]
        L.orc_sys_new.restype = C.c_void_p
        L.orc_sys_new.argtypes = [C.c_int, C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_sys_free.argtypes = [C.c_void_p]
        L.orc_sys_init_step.argtypes = [C.c_void_p]
        L.orc_sys_integrate_one_block.argtypes = [C.c_void_p]
        L.orc_sys_integrate_one_block.restype = C.c_int
        L.orc_sys_integrate_one_dtmax.argtypes = [C.c_void_p]
        L.orc_sys_integrate_until_inplace.argtypes = [C.c_void_p, C.c_double]
        for name in ("orc_sys_energy", "orc_sys_tsys", "orc_sys_init_energy"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_void_p]
        for name in ("orc_sys_num_step", "orc_sys_num_bstep"):
            getattr(L, name).restype = C.c_long
            getattr(L, name).argtypes = [C.c_void_p]
        for name in ("orc_sys_ptcl", "orc_sys_pot", "orc_sys_dtbuf"):
            getattr(L, name).restype = _dp
            getattr(L, name).argtypes = [C.c_void_p]
        _lib = L
    return _lib


# ---------------------------------------------------------------- records
def make_particles(order: int, mass, pos, vel, ids=None, tlast: float = 0.0) -> np.ndarray:
    """Particle records as the reference constructs them (SRC/hermite6.h:37-47):
    derivatives zero, tlast and dt zero."""
    n = len(mass)
    P = np.zeros((n, PSTRIDE[order]), dtype=np.float64)
    idv = np.arange(n, dtype=np.int64) if ids is None else np.asarray(ids, dtype=np.int64)
    P[:, W_ID] = idv.view(np.float64)
    P[:, W_MASS] = mass
    P[:, W_TLAST] = tlast
    P[:, W_POS:W_POS + 3] = pos
    # double-double (order 9): posH = pos, posL = 0, vel after posL
    # (SRC/hermite8dd.h:64-78)
    P[:, (10 if order == 9 else W_VEL):(10 if order == 9 else W_VEL) + 3] = vel
    return P


def ids_of(P: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(P[:, W_ID]).view(np.int64)


# ---------------------------------------------------------------- wrappers
def pow_one_nth_quant(n: int, x: float) -> float:
    return lib().orc_pow_one_nth_quant(n, x)


def invfordt(dt: float):
    h, hinv = C.c_double(), C.c_double()
    lib().orc_invfordt(dt, C.byref(h), C.byref(hinv))
    return h.value, hinv.value


def predict(order: int, P: np.ndarray, tsys: float) -> np.ndarray:
    P = np.ascontiguousarray(P)
    Q = np.zeros((len(P), QSTRIDE[order]))
    lib().orc_predict(order, len(P), _ptr(P), tsys, _ptr(Q))
    return Q


def force(order: int, Q: np.ndarray, nact: int, eps2: float) -> np.ndarray:
    Q = np.ascontiguousarray(Q)
    F = np.zeros((nact, FSTRIDE[order]))
    lib().orc_force(order, len(Q), _ptr(Q), nact, eps2, _ptr(F))
    return F


def force_ld(order: int, Q: np.ndarray, nact: int, eps2: float) -> np.ndarray:
    Q = np.ascontiguousarray(Q)
    F = np.zeros((nact, FSTRIDE[order]))
    lib().orc_force_ld(order, len(Q), _ptr(Q), nact, eps2, _ptr(F))
    return F


def force_list(order: int, Q: np.ndarray, ilist, eps2: float) -> np.ndarray:
    Q = np.ascontiguousarray(Q)
    il = np.ascontiguousarray(ilist, dtype=np.int32)
    F = np.zeros((len(il), FSTRIDE[order]))
    lib().orc_force_list(order, len(Q), _ptr(Q), len(il), il.ctypes.data_as(C.POINTER(C.c_int)), eps2, _ptr(F))
    return F


def potential(order: int, P: np.ndarray, eps2: float) -> np.ndarray:
    P = np.ascontiguousarray(P)
    pot = np.zeros(len(P))
    lib().orc_potential(order, len(P), _ptr(P), eps2, _ptr(pot))
    return pot


def correct(order: int, p: np.ndarray, f: np.ndarray, etapow: float, dtlim: float) -> np.ndarray:
    p = np.array(p, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    lib().orc_correct(order, _ptr(p), _ptr(f), etapow, dtlim)
    return p


def init_dt(order: int, p: np.ndarray, eta_s: float, dtmax: float) -> np.ndarray:
    p = np.array(p, dtype=np.float64)
    lib().orc_init_dt(order, _ptr(p), eta_s, dtmax)
    return p


def aarseth_step_quant(order: int, p: np.ndarray, etapow: float) -> float:
    p = np.ascontiguousarray(p, dtype=np.float64)
    return lib().orc_aarseth_step_quant(order, _ptr(p), etapow)


def calc_dtlim(tnext: float, dtmax: float) -> float:
    return lib().orc_calc_dtlim(tnext, dtmax)


class System:
    """Block-timestep driver restated from SRC/nbodysystem.h (TITAN path)."""

    def __init__(self, order, P, tsys, eps2, eta, eta_s, dtmax):
        self.order = order
        self.n = len(P)
        self._P = np.ascontiguousarray(P, dtype=np.float64)
        self.h = lib().orc_sys_new(order, self.n, _ptr(self._P), tsys, eps2, eta, eta_s, dtmax)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sys_free(self.h)
            self.h = None

    def init_step(self):
        lib().orc_sys_init_step(self.h)

    def integrate_one_block(self) -> int:
        return lib().orc_sys_integrate_one_block(self.h)

    def integrate_one_dtmax(self):
        lib().orc_sys_integrate_one_dtmax(self.h)

    def integrate_until_inplace(self, tend):
        """Block steps without the dt-level permutation (the GPU-resident
        driver's schedule) until tsys reaches tend."""
        lib().orc_sys_integrate_until_inplace(self.h, tend)

    def energy(self) -> float:
        return lib().orc_sys_energy(self.h)

    @property
    def tsys(self):
        return lib().orc_sys_tsys(self.h)

    @property
    def init_energy(self):
        return lib().orc_sys_init_energy(self.h)

    @property
    def num_step(self):
        return lib().orc_sys_num_step(self.h)

    @property
    def num_bstep(self):
        return lib().orc_sys_num_bstep(self.h)

    def ptcl(self) -> np.ndarray:
        p = lib().orc_sys_ptcl(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n * PSTRIDE[self.order],)).reshape(self.n, -1).copy()


# ---------------------------------------------------------------- reference
class RefAVX:
    """The reference's AVX2 back end (SRC/hermite{4,6,8}-avx.h) and driver,
    built with the reference's own flags as the stand-alone ``_ref/ref_avxN``
    executable and run as a child process."""

    def __init__(self, order: int):
        self.order = order
        self.exe = os.path.join(REF_DIR, f"ref_avx{order}")
        if not os.path.exists(self.exe):
            raise FileNotFoundError(self.exe)

    def _run(self, args, env=None):
        import json
        import subprocess

        out = subprocess.run([self.exe] + [str(a) for a in args], check=True, capture_output=True,
                             text=True, env=env)
        lines = [ln for ln in out.stdout.splitlines() if ln.startswith("{")]
        return json.loads(lines[-1]) if lines else None

    def forces(self, P, tsys, nact, eps2, reps=1, threads=None, want_output=True):
        """Returns (forces, timing dict).  The reference needs nact % 4 == 0
        for its 4-wide i-groups to stay inside [0, nact)."""
        import tempfile

        env = dict(os.environ)
        if threads:
            env["OMP_NUM_THREADS"] = str(threads)
        P = np.ascontiguousarray(P, dtype=np.float64)
        with tempfile.TemporaryDirectory() as td:
            fin, fout = os.path.join(td, "p.bin"), os.path.join(td, "f.bin")
            P.tofile(fin)
            info = self._run(["force", fin, len(P), repr(float(tsys)), nact, repr(float(eps2)), reps,
                              fout if want_output else "-"], env)
            F = np.fromfile(fout).reshape(nact, FSTRIDE[self.order]) if want_output else None
        return F, info

    def potential(self, P, eps2):
        import tempfile

        P = np.ascontiguousarray(P, dtype=np.float64)
        with tempfile.TemporaryDirectory() as td:
            fin, fout = os.path.join(td, "p.bin"), os.path.join(td, "pot.bin")
            P.tofile(fin)
            self._run(["pot", fin, len(P), repr(float(eps2)), fout])
            return np.fromfile(fout)

    def run(self, P, eps2, eta, eta_s, dtmax, tend, threads=None, reps=1):
        """The reference driver on the AVX2 back end from t = 0 (init_step,
        then integrate_one_dtmax until tend; only the integration is timed),
        reps times from the same state.  Returns {e0, e1, de, num_step,
        num_bstep, seconds, reps, threads}: seconds summed over the reps,
        the rest of one run."""
        import tempfile

        env = dict(os.environ)
        if threads:
            env["OMP_NUM_THREADS"] = str(threads)
        P = np.ascontiguousarray(P, dtype=np.float64)
        with tempfile.TemporaryDirectory() as td:
            fin = os.path.join(td, "p.bin")
            P.tofile(fin)
            return self._run(["run", fin, len(P)] + [repr(float(a)) for a in (eps2, eta, eta_s, dtmax, tend)]
                             + [int(reps)], env)


class Ref:
    """The reference's generic FP64 back end (SRC/hermite{4,6,8}.h) and its
    NbodySystem driver (SRC/nbodysystem.h), built into oracle/_ref with
    -ffp-contract=off and loaded in-process."""

    def __init__(self, kind: str, order: int):
        path = os.path.join(REF_DIR, f"ref_{kind}{order}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L, self.order, self.kind = L, order, kind
        L.ref_gravity_new.restype = C.c_void_p
        L.ref_gravity_new.argtypes = [C.c_int]
        L.ref_gravity_free.argtypes = [C.c_void_p]
        L.ref_set_jp.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ref_predict_all.argtypes = [C.c_void_p, C.c_double]
        L.ref_force.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
        L.ref_potential.argtypes = [C.c_void_p, C.c_double, _dp]
        L.ref_correct.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double]
        L.ref_init_dt.argtypes = [_dp, C.c_double, C.c_double]
        L.ref_pow_one_nth_quant.restype = C.c_double
        L.ref_pow_one_nth_quant.argtypes = [C.c_int, C.c_double]
        L.ref_invfordt.argtypes = [C.c_double, _dp, _dp]
        if hasattr(L, "ref_sys_new"):  # the double-double build has no driver
            L.ref_sys_new.restype = C.c_void_p
            L.ref_sys_new.argtypes = [C.c_int, _dp] + [C.c_double] * 6
            for name in ("ref_sys_free", "ref_sys_init_step", "ref_sys_integrate_one_dtmax"):
                getattr(L, name).argtypes = [C.c_void_p]
            for name in ("ref_sys_energy", "ref_sys_init_energy", "ref_sys_tsys"):
                getattr(L, name).restype = C.c_double
                getattr(L, name).argtypes = [C.c_void_p]
            for name in ("ref_sys_num_step", "ref_sys_num_bstep"):
                getattr(L, name).restype = C.c_long
                getattr(L, name).argtypes = [C.c_void_p]
            L.ref_sys_get_ptcl.argtypes = [C.c_void_p, _dp]
            L.ref_sys_get_pot.argtypes = [C.c_void_p, _dp]
        assert L.ref_particle_bytes() == 8 * PSTRIDE[order]

    # Gravity interface -------------------------------------------------
    def gravity_forces(self, P: np.ndarray, tsys: float, nact: int, eps2: float) -> np.ndarray:
        n = len(P)
        P = np.ascontiguousarray(P)
        g = self.L.ref_gravity_new(n)
        try:
            for i in range(n):
                self.L.ref_set_jp(g, i, P[i].ctypes.data_as(_dp))
            self.L.ref_predict_all(g, tsys)
            F = np.zeros((max(nact, (nact + 3) // 4 * 4), FSTRIDE[self.order]))
            self.L.ref_force(g, nact, eps2, _ptr(F))
        finally:
            self.L.ref_gravity_free(g)
        return F[:nact].copy()

    def gravity_potential(self, P: np.ndarray, eps2: float) -> np.ndarray:
        n = len(P)
        P = np.ascontiguousarray(P)
        g = self.L.ref_gravity_new(n)
        try:
            for i in range(n):
                self.L.ref_set_jp(g, i, P[i].ctypes.data_as(_dp))
            pot = np.zeros(n)
            self.L.ref_potential(g, eps2, _ptr(pot))
        finally:
            self.L.ref_gravity_free(g)
        return pot

    def correct(self, p, f, eta, etapow, dtlim):
        p = np.array(p, dtype=np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64)
        self.L.ref_correct(_ptr(p), _ptr(f), eta, etapow, dtlim)
        return p

    def init_dt(self, p, eta_s, dtmax):
        p = np.array(p, dtype=np.float64)
        self.L.ref_init_dt(_ptr(p), eta_s, dtmax)
        return p

    def pow_one_nth_quant(self, n, x):
        return self.L.ref_pow_one_nth_quant(n, x)

    # NbodySystem ----------------------------------------------------------
    def system(self, P, tsys, eps2, eta, eta_s, dtmax, tend=1.0):
        return _RefSystem(self, P, tsys, eps2, eta, eta_s, dtmax, tend)


class _RefSystem:
    def __init__(self, ref, P, tsys, eps2, eta, eta_s, dtmax, tend):
        self.ref, self.n = ref, len(P)
        P = np.ascontiguousarray(P, dtype=np.float64)
        self.h = ref.L.ref_sys_new(self.n, _ptr(P), tsys, eps2, eta, eta_s, dtmax, tend)

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.L.ref_sys_free(self.h)
            self.h = None

    def init_step(self):
        self.ref.L.ref_sys_init_step(self.h)

    def integrate_one_dtmax(self):
        self.ref.L.ref_sys_integrate_one_dtmax(self.h)

    def energy(self):
        return self.ref.L.ref_sys_energy(self.h)

    @property
    def init_energy(self):
        return self.ref.L.ref_sys_init_energy(self.h)

    @property
    def tsys(self):
        return self.ref.L.ref_sys_tsys(self.h)

    @property
    def num_step(self):
        return self.ref.L.ref_sys_num_step(self.h)

    @property
    def num_bstep(self):
        return self.ref.L.ref_sys_num_bstep(self.h)

    def ptcl(self):
        P = np.zeros((self.n, PSTRIDE[self.ref.order]))
        self.ref.L.ref_sys_get_ptcl(self.h, _ptr(P))
        return P
