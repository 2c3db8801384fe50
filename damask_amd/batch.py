"""Many independent RVEs on one GPU.

Each RVE is a GridMechSpectralBasic (one device handle with its own HIP
stream).  ``dmk_batch_run`` / ``dmk_batch_adjoint`` drive all of them from a
pool of native host threads, so the kernels and host synchronisations of
different RVEs overlap on the device.  ``random_batch`` generates the batched
workload: per RVE a cubic grid edge, a grain count, uniform random
orientations, its own xi_0 and n, and uniaxial tension along a random
direction (the load frame rotated by a uniform random rotation).
"""
from __future__ import annotations

import copy
import ctypes as C
import os
import time

import numpy as np

from . import _lib, adjoint, config, workloads
from .grid import ADJ_NOUT, GridMechSpectralBasic, loadcase_struct


def random_batch(n_rve: int, master_seed: int = 0, edges=(8, 12, 16, 20, 24, 32), grains=(10, 50),
                 xi_0=(20.0e6, 45.0e6), n_exp=(10.0, 30.0), strain: float = 0.02, increments: int = 40):
    """[(grid, material, loadcase)] of n_rve RVEs, every draw from one seeded generator."""
    rng = np.random.default_rng(master_seed)
    out = []
    for _ in range(n_rve):
        edge = int(rng.choice(edges))
        ng = int(rng.integers(grains[0], grains[1] + 1))
        phase = copy.deepcopy(workloads.FCC_PHASE)
        pl = phase["mechanical"]["plastic"]
        pl["xi_0_sl"] = [float(rng.uniform(*xi_0))]
        pl["n_sl"] = [float(rng.uniform(*n_exp))]
        grid, mat = workloads.polycrystal([edge] * 3, ng, int(rng.integers(2**31 - 1)), phase=phase)
        _, lcs = config.parse_load(workloads.tension_x(strain / 1.0e-3, increments))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        lc = lcs[0]
        lc.q_R = q if q[0] >= 0.0 else -q
        out.append((grid, config.parse_material(mat), lc))
    return out


def mixed_load_batch(n_rve: int = 64, master_seed: int = 0, edge: int = 48, grains: int = 300,
                     increments: int = 50, t: float = 20.0, phase: dict | None = None):
    """The large batched configuration: n_rve RVEs of edge^3 cells with
    `grains` seeded Voronoi grains and uniform random orientations each, the
    24-system BCC phenopowerlaw phase (Schmid law), each loaded by its own
    seeded random symmetric deformation rate (equivalent rate 1e-3/s, trace
    within +-1e-4/s) over `increments` increments of t / increments."""
    rng = np.random.default_rng(master_seed)
    out = []
    for _ in range(n_rve):
        grid, mat = workloads.polycrystal([edge] * 3, grains, int(rng.integers(2**31 - 1)),
                                          phase=phase or workloads.BCC_PHASE)
        rate = workloads.random_symmetric_rate(rng)
        _, lcs = config.parse_load(workloads.deformation_rate(rate, t, increments))
        out.append((grid, config.parse_material(mat), lcs[0]))
    return out


class Batch:
    """Solvers of a batch of RVEs and the batched forward / adjoint drivers."""

    def __init__(self, items, numerics: config.Numerics | None = None):
        self.items = items
        self.lib = _lib.load()
        self.solvers = [GridMechSpectralBasic(g, m, numerics) for g, m, _ in items]
        self.cells = np.asarray([s.N for s in self.solvers], np.int64)
        self._h = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        self.threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def tape(self, segment: int = 0, capacity: int | None = None):
        """Start the adjoint tapes; segment > 0: checkpoints every segment
        increments and recomputation in the reverse sweep (see adjoint_tape).
        capacity: increments a tape holds, allocated up front (default 2 N + 16
        per load case, room for cutbacks); a run that records more fails."""
        for s, (_, _, lc) in zip(self.solvers, self.items):
            s.adjoint_tape(capacity if capacity is not None else 2 * lc.N + 16, segment)

    def run(self, threads: int | None = None, lockstep: bool = False) -> dict:
        """All load cases; with lockstep the RVEs in flight (one per thread)
        run their residual evaluations as shared rounds, one launch per kernel
        over the round (dmk_batch_run_lockstep; DMK_LOCKSTEP=stress shares the
        stress integration only)."""
        lcs = (_lib.LoadCase * len(self.items))(*[loadcase_struct(lc) for _, _, lc in self.items])
        n = len(self.solvers)
        rc = np.zeros(n, np.int32)
        wall = np.zeros(n)
        fn = self.lib.dmk_batch_run_lockstep if lockstep else self.lib.dmk_batch_run
        t0 = time.perf_counter()
        r = fn(self._h, n, lcs, self.cells.ctypes.data_as(C.POINTER(C.c_int64)), threads or self.threads,
               _lib.iptr(rc), _lib.dptr(wall))
        t = time.perf_counter() - t0
        _lib.check(r, "dmk_batch_run_lockstep" if lockstep else "dmk_batch_run")
        return {"wall_s": t, "rve_wall_s": wall}

    def curves(self) -> list[np.ndarray]:
        return [adjoint._P33(s.adjoint_info()["P_av"]) for s in self.solvers]

    def gradients(self, P_bar: list[np.ndarray], rtol: float = 1e-6, threads: int | None = None,
                  lockstep: bool = False):
        """Per RVE: gradient of sum(P_bar * P_av) with respect to its phase's parameters.  With
        lockstep the sweeps in flight (one per thread) run their adjoint operators and dot
        products as shared rounds (dmk_batch_adjoint_lockstep); the gradients are the same."""
        n = len(self.solvers)
        pb = [adjoint._P9(p) for p in P_bar]
        tb = [np.zeros((1, ADJ_NOUT)) for _ in range(n)]
        PB = (C.POINTER(C.c_double) * n)(*[_lib.dptr(p) for p in pb])
        TB = (C.POINTER(C.c_double) * n)(*[_lib.dptr(t) for t in tb])
        rc = np.zeros(n, np.int32)
        wall = np.zeros(n)
        t0 = time.perf_counter()
        name = "dmk_batch_adjoint_lockstep" if lockstep else "dmk_batch_adjoint"
        r = getattr(self.lib, name)(self._h, n, PB, float(rtol), TB,
                                    self.cells.ctypes.data_as(C.POINTER(C.c_int64)), threads or self.threads,
                                    _lib.iptr(rc), _lib.dptr(wall))
        t = time.perf_counter() - t0
        _lib.check(r, name)
        grads = [adjoint.reduce_cotangent(s.material.phases[0], tb[i][0]) for i, s in enumerate(self.solvers)]
        return grads, {"wall_s": t, "rve_wall_s": wall}

    def close(self):
        for s in self.solvers:
            s.close()
