"""Initial conditions and run parameters of the five benchmark workloads.

Every workload is a seeded Plummer sphere in Henon units drawn by the
Aarseth-Henon-Wielen method of RUN/gen-plum.c:116-181 (radii truncated at
22.8 Plummer scale lengths, centre-of-mass correction of gen-plum.c:187-226),
with the integration parameters of the reference's parameter files
(RUN/inp.{4th,6th,8th}: eta, eta_s, dt_max):

  c1  4th order, N = 1024, seed 1, eps = 1/64, eta = 0.1 (eta_s 0.01),
      dt_max = 1/16, dt_min = 2^-30, t = 0 -> 1
  c2  6th order, N = 65536, seed 2, eps = 1/1024, eta = 0.2 (eta_s 0.1),
      dt_max = 1/16, t = 0 -> 1
  c3  8th order, N = 262144, seed 3, eps = 1/4096, eta = 0.4 (eta_s 0.1),
      dt_max = 1/32, t = 0 -> 1/8
  c4  ensemble of 256 clusters, 4th order, N_k = round(exp(U)) with U
      uniform on [ln 256, ln 8192] (glibc rand() stream of seed 4),
      cluster k a Plummer sphere from seed 4000 + k, eps_k = 4/N_k,
      eta = 0.1 (eta_s 0.01), dt_max = 1/16, t = 0 -> 1
  c5  6th order, N = 2^20, Plummer positions and velocities from seed 5 with
      Salpeter masses (slope 2.35 on [0.1, 100], rescaled to total mass 1,
      drawn from the stream of seed 5005), centre of mass re-corrected with
      those masses, eps = 1/4096, eta = 0.2 (eta_s 0.1), dt_max = 1/16,
      dt_min = 2^-40, t = 0 -> 1/16

eta_s is the reference's initial-step parameter (SRC/hermite{4,6,8}.h
init_dt); the workloads do not name it, so the reference's own files supply
it (0.01 at 4th order, 0.1 at 6th and 8th).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Cluster:
    n: int
    seed: int
    eps: float
    mass: str = "equal"       # "equal" or "salpeter"
    mass_seed: int = 0


@dataclass
class Workload:
    name: str
    order: int
    eta: float
    eta_s: float
    dtmax: float
    dtmin: float
    tend: float
    clusters: list = field(default_factory=list)

    @property
    def n_total(self) -> int:
        return sum(c.n for c in self.clusters)


def _ensemble_sizes(nclu: int, seed: int, lo: int = 256, hi: int = 8192):
    from . import host_lib
    import ctypes as C

    r = (C.c_int * nclu)()
    host_lib().hh_glibc_rand(seed, nclu, r)
    a, b = math.log(lo), math.log(hi)
    return [int(round(math.exp(a + (v / 2147483648.0) * (b - a)))) for v in r]


def workload(name: str, nclu: int | None = None) -> Workload:
    """The named workload; nclu limits the c4 ensemble to its first clusters."""
    if name == "c1":
        return Workload("c1", 4, 0.1, 0.01, 1 / 16, 2.0 ** -30, 1.0, [Cluster(1024, 1, 1 / 64)])
    if name == "c2":
        return Workload("c2", 6, 0.2, 0.1, 1 / 16, 0.0, 1.0, [Cluster(65536, 2, 1 / 1024)])
    if name == "c3":
        return Workload("c3", 8, 0.4, 0.1, 1 / 32, 0.0, 1 / 8, [Cluster(262144, 3, 1 / 4096)])
    if name == "c4":
        sizes = _ensemble_sizes(256, 4)
        if nclu is not None:
            sizes = sizes[:nclu]
        return Workload("c4", 4, 0.1, 0.01, 1 / 16, 0.0, 1.0,
                        [Cluster(nk, 4000 + k, 4.0 / nk) for k, nk in enumerate(sizes)])
    if name == "c5":
        return Workload("c5", 6, 0.2, 0.1, 1 / 16, 2.0 ** -40, 1 / 16,
                        [Cluster(1 << 20, 5, 1 / 4096, "salpeter", 5005)])
    raise ValueError(f"unknown workload {name!r}")


def cluster_state(order: int, c: Cluster) -> np.ndarray:
    """Reference Particle records (derivatives zero) of one cluster."""
    from . import make_particles, plummer, salpeter

    m, x, v = plummer(c.n, c.seed)
    if c.mass == "salpeter":
        m = salpeter(c.n, c.mass_seed)
        mt = m.sum()
        x = x - (m[:, None] * x).sum(axis=0) / mt
        v = v - (m[:, None] * v).sum(axis=0) / mt
    return make_particles(order, m, x, v)
