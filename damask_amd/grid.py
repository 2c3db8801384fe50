"""Python front end of the grid mechanical solver, mirroring the reference's
solver plug-in interface (``src/grid/grid_mech_spectral_basic.f90:91-96``):

* ``GridMechSpectralBasic(grid, material, numerics)`` -- ``*_init`` (:103-237)
* ``.forward(cutBack, guess, Delta_t, Delta_t_old, t_remaining, loadcase)`` (:284-355)
* ``.solution(incInfo)`` -> ``SolutionState`` (:243-277)
* ``.updateCoords()`` (:361-372)
* ``.restart_data()`` / ``.load_restart()`` (:378-416, :198-232)

and the ``DAMASK_grid`` increment/cutback loop (``src/grid/DAMASK_grid.f90:288-460``)
as ``DamaskGrid``.  All work runs in libdamask_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .config import LoadCase, Material, Numerics
from .geom import Grid

FIELDS = {"F": 0, "P": 1, "F_e": 2, "F_p": 3, "L_p": 4, "S": 5, "F_i": 6, "O": 7, "xi_sl": 8,
          "gamma_sl": 9, "dPdF": 10, "F_lastinc": 11, "F_tau": 12, "F_tau_lastinc": 13}
SCALARS = {"F_aim": 0, "P_av": 1, "P_aim": 2, "C_volAvg": 3, "C_minMaxAvg": 4, "err": 5, "F_aimDot": 6,
           "F_aim_lastinc": 7, "C_volAvgLastInc": 8, "C_minMaxAvgLastInc": 9, "err_curl": 10,
           "C_minMaxAvgRestart": 11, "fft_path": 12}
# restart data of grid_mech_spectral_basic (restartWrite, :378-416) and of the
# phase (phase_mechanical.f90:1248-1280)
RESTART_FIELDS = ("F", "F_lastinc", "F_p", "L_p", "S", "xi_sl", "gamma_sl")
RESTART_SCALARS = ("P_aim", "F_aim", "F_aim_lastinc", "F_aimDot")
RESTART_STIFFNESS = ("C_volAvg", "C_volAvgLastInc", "C_minMaxAvg")


@dataclasses.dataclass
class SolutionState:
    """``tSolutionState`` (``spectral_utilities.f90:77-83``)."""
    iterationsNeeded: int = 0
    converged: bool = True
    stagConverged: bool = True
    residual_evaluations: int = 0


def f33(a) -> np.ndarray:
    """(3,3) [i, j] -> 9 doubles in Fortran order."""
    return np.ascontiguousarray(np.asarray(a, float).T.reshape(9))


def phase_desc(p) -> _lib.PhaseDesc:
    d = _lib.PhaseDesc()
    d.lattice = {"cF": 0, "cI": 1}[p.lattice]
    d.n_fam = len(p.N_sl)
    for i, v in enumerate(p.N_sl):
        d.N_sl[i] = v
    d.C_11, d.C_12, d.C_44 = p.C_11, p.C_12, p.C_44
    for name, src in (("dot_gamma_0", p.dot_gamma_0_sl), ("n", p.n_sl), ("a", p.a_sl), ("h_0", p.h_0_sl_sl),
                      ("xi_0", p.xi_0_sl), ("xi_inf", p.xi_inf_sl), ("c_1", p.c_1), ("c_2", p.c_2),
                      ("f_sat", p.f_sat_sl_tw)):
        arr = getattr(d, name)
        for i, v in enumerate(src):
            arr[i] = v
    d.n_h = len(p.h_sl_sl)
    for i, v in enumerate(p.h_sl_sl):
        d.h_sl_sl[i] = v
    d.n_nS = len(p.a_nS)
    for i, v in enumerate(p.a_nS):
        d.a_nS[i] = v
    d.atol_xi, d.atol_gamma = p.atol_xi, p.atol_gamma
    return d


def numerics_struct(n: Numerics, cref_rule: int = 1, scheme: int = 0) -> _lib.Numerics:
    s = _lib.Numerics()
    for f, _ in _lib.Numerics._fields_:
        if f == "cref_rule":
            s.cref_rule = cref_rule
        elif f == "scheme":
            s.scheme = scheme
        elif hasattr(n, f):
            v = getattr(n, f)
            setattr(s, f, int(v) if isinstance(getattr(s, f), int) else float(v))
    return s


def loadcase_struct(lc: LoadCase) -> _lib.LoadCase:
    L = _lib.LoadCase()
    L.def_type = lc.def_type
    for k, (v, m) in enumerate(zip(f33(lc.def_values), f33(lc.def_mask.astype(float)))):
        L.def_values[k] = v
        L.def_mask[k] = int(m)
    L.stress_type = lc.stress_type
    for k, (v, m) in enumerate(zip(f33(lc.stress_values), f33(lc.stress_mask.astype(float)))):
        L.stress_values[k] = v
        L.stress_mask[k] = int(m)
    for k in range(4):
        L.q_R[k] = lc.q_R[k]
    L.t, L.r, L.N = lc.t, lc.r, lc.N
    L.f_out, L.estimate_rate = min(lc.f_out, 2**31 - 1), int(lc.estimate_rate)
    return L


# per-phase parameter cotangents of dmk_adjoint_backward: xi_0, n, xi_inf,
# h_0, a, dot_gamma_0 per slip system (24 each), h_sl-sl per interaction type (32)
ADJ_NSYS = 24
ADJ_NOUT = 6 * ADJ_NSYS + 32


class GridMechSpectralBasic:
    """One RVE on the GPU behind the spectral-basic plug-in interface."""
    taping = False
    SCHEME = 0
    RESTART_FIELDS = RESTART_FIELDS

    def __init__(self, grid: Grid, material: Material, numerics: Numerics | None = None, cref_rule: int = 1):
        self.lib = _lib.load()
        numerics = numerics or Numerics()
        numerics.validate()
        self.grid = grid
        self.material = material
        self.maps = material.cell_maps(grid.material)
        self.N = grid.n_cells
        self.numerics = numerics
        descs = (_lib.PhaseDesc * len(material.phases))(*[phase_desc(p) for p in material.phases])
        cells = np.ascontiguousarray(grid.cells, np.int32)
        size = np.ascontiguousarray(grid.size, float)
        phase = np.ascontiguousarray(self.maps["phase"], np.int32)
        q0 = np.ascontiguousarray(self.maps["O"], float)
        num = numerics_struct(numerics, cref_rule, self.SCHEME)
        h = C.c_void_p()
        _lib.check(self.lib.dmk_mech_init(_lib.iptr(cells), _lib.dptr(size), len(material.phases), descs,
                                          _lib.iptr(phase), _lib.dptr(q0), C.byref(num), C.byref(h)),
                   "dmk_mech_init")
        self.h = h
        self.nsl = self.lib.dmk_nsl(h)

    # --- plug-in interface -------------------------------------------------
    def forward(self, cutBack: bool, guess: bool, Delta_t: float, Delta_t_old: float, t_remaining: float,
                loadcase: LoadCase) -> None:
        bc = loadcase_struct(loadcase)
        _lib.check(self.lib.dmk_mech_forward(self.h, int(cutBack), int(guess), Delta_t, Delta_t_old, t_remaining,
                                             C.byref(bc)), "dmk_mech_forward")

    def materialpoint_forward(self) -> None:
        _lib.check(self.lib.dmk_mech_materialpoint_forward(self.h), "dmk_mech_materialpoint_forward")

    def solution(self, incInfo: str = "") -> SolutionState:
        it, conv, nf = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self.lib.dmk_mech_solution(self.h, C.byref(it), C.byref(conv), C.byref(nf)), "dmk_mech_solution")
        return SolutionState(it.value, bool(conv.value), True, nf.value)

    def updateCoords(self):
        nn = int(np.prod(np.asarray(self.grid.cells) + 1))
        x_p = np.zeros(3 * self.N)
        x_n = np.zeros(3 * nn)
        _lib.check(self.lib.dmk_mech_update_coords(self.h, _lib.dptr(x_p), _lib.dptr(x_n)), "dmk_mech_update_coords")
        return x_p.reshape(-1, 3), x_n.reshape(-1, 3)

    # --- data access ---------------------------------------------------------
    def field(self, name: str) -> np.ndarray:
        """Per-cell field, shape (N, 3, 3) with [e, i, j] = T_ij (O: (N,4), states: (N, nsl))."""
        w = {"O": 4, "xi_sl": self.nsl, "gamma_sl": self.nsl, "dPdF": 81}.get(name, 9)
        out = np.zeros(self.N * w)
        _lib.check(self.lib.dmk_mech_get_field(self.h, FIELDS[name], _lib.dptr(out)), "dmk_mech_get_field")
        out = out.reshape(self.N, w)
        if w == 9:
            return out.reshape(self.N, 3, 3).transpose(0, 2, 1).copy()
        return out

    def set_field(self, name: str, value: np.ndarray) -> None:
        v = np.asarray(value, float)
        if v.ndim == 3:
            v = v.transpose(0, 2, 1)
        v = np.ascontiguousarray(v.reshape(-1))
        _lib.check(self.lib.dmk_mech_set_field(self.h, FIELDS[name], _lib.dptr(v)), "dmk_mech_set_field")

    def scalar(self, name: str) -> np.ndarray:
        n = {"C_volAvg": 81, "C_minMaxAvg": 81, "C_volAvgLastInc": 81, "C_minMaxAvgLastInc": 81,
             "C_minMaxAvgRestart": 81, "err": 2, "err_curl": 1, "fft_path": 5}.get(name, 9)
        out = np.zeros(n)
        _lib.check(self.lib.dmk_mech_get_scalars(self.h, SCALARS[name], _lib.dptr(out)), "dmk_mech_get_scalars")
        if n == 9:
            return out.reshape(3, 3).T.copy()
        if n == 81:
            return out.reshape(3, 3, 3, 3, order="F")
        return out

    def set_scalar(self, name: str, value) -> None:
        v = np.asarray(value, float)
        v = np.ascontiguousarray(v.reshape(-1, order="F"))
        _lib.check(self.lib.dmk_mech_set_scalars(self.h, SCALARS[name], _lib.dptr(v)), "dmk_mech_set_scalars")

    def constitutive(self, F: np.ndarray, Delta_t: float):
        """homogenization_mechanical_response for an F field (N,3,3) -> (P, P_av, C_volAvg, C_minMaxAvg, status)."""
        Ff = np.ascontiguousarray(np.asarray(F, float).transpose(0, 2, 1).reshape(-1))
        P = np.zeros_like(Ff)
        Pav, Cv, Cm = np.zeros(9), np.zeros(81), np.zeros(81)
        st = C.c_int32()
        _lib.check(self.lib.dmk_constitutive(self.h, Delta_t, _lib.dptr(Ff), _lib.dptr(P), _lib.dptr(Pav),
                                             _lib.dptr(Cv), _lib.dptr(Cm), C.byref(st)), "dmk_constitutive")
        return (P.reshape(-1, 3, 3).transpose(0, 2, 1), Pav.reshape(3, 3).T, Cv.reshape(3, 3, 3, 3, order="F"),
                Cm.reshape(3, 3, 3, 3, order="F"), st.value)

    def counters(self) -> dict:
        out = np.zeros(4, np.int64)
        _lib.check(self.lib.dmk_mech_counters(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "dmk_mech_counters")
        return {"updates": int(out[0]), "newton": int(out[1]), "state": int(out[2]), "residuals": int(out[3])}

    def cell_counters(self) -> dict:
        """Per-cell Newton iterations, state iterations and sub-steps tried in the last response."""
        out = {k: np.zeros(self.N, np.int32) for k in ("newton", "fpi", "substeps")}
        _lib.check(self.lib.dmk_mech_cell_counters(self.h, _lib.iptr(out["newton"]), _lib.iptr(out["fpi"]),
                                                   _lib.iptr(out["substeps"])), "dmk_mech_cell_counters")
        return out

    KERNEL_GROUPS = ("stress", "tangent", "reduce", "fft_fwd", "gamma_div", "fft_inv", "c_avg")

    def profile(self, level: int = 2) -> None:
        """Per-kernel HIP-event timing: 0 off, 1 the stress kernel only, 2 every group."""
        _lib.check(self.lib.dmk_mech_profile(self.h, int(level)), "dmk_mech_profile")

    def kernel_times(self) -> dict:
        out = {}
        for k, name in enumerate(self.KERNEL_GROUPS):
            ms = C.c_double()
            n = C.c_int64()
            _lib.check(self.lib.dmk_mech_kernel_time(self.h, k, C.byref(ms), C.byref(n)), "dmk_mech_kernel_time")
            out[name] = (ms.value, n.value)
        return out

    def run_loadcase(self, lc: LoadCase) -> dict:
        """All increments of one load case in the native driver loop."""
        bc = loadcase_struct(lc)
        stats = np.zeros(4 * lc.N)
        P = np.zeros(9 * lc.N)
        Fa = np.zeros(9 * lc.N)
        t = np.zeros(1)
        nu = C.c_int64(0)
        rc = self.lib.dmk_grid_run_loadcase(self.h, C.byref(bc), _lib.dptr(t), _lib.dptr(stats), _lib.dptr(P),
                                            _lib.dptr(Fa), C.byref(nu))
        if rc not in (0, 950):
            _lib.check(rc, "dmk_grid_run_loadcase")
        return {"rc": rc, "t": float(t[0]), "stats": stats.reshape(-1, 4), "P_av": P.reshape(-1, 9),
                "F_aim": Fa.reshape(-1, 9), "n_updates": nu.value}

    # --- restart ------------------------------------------------------------
    def restart_write(self) -> dict:
        """The data restartWrite stores (grid_mech_spectral_basic.f90:378-416)."""
        d = {k: self.field(k) for k in self.RESTART_FIELDS}
        d.update({k: self.scalar(k) for k in RESTART_SCALARS + RESTART_STIFFNESS})
        # stored as C_minMaxAvg: the stiffness the Gamma operator was built from (:388,408)
        d["C_minMaxAvg"] = self.scalar("C_minMaxAvgRestart")
        return d

    def restart_read(self, d: dict, restart_inc: int) -> None:
        """Restart read of the solver init (grid_mech_spectral_basic.f90:198-232): fields
        and aims, a dt = 0 response at the restored F, then the stiffness averages; the
        native driver then skips increments 1..restart_inc."""
        for k in self.RESTART_FIELDS:
            self.set_field(k, d[k])
        for k in RESTART_SCALARS:
            self.set_scalar(k, d[k])
        _lib.check(self.lib.dmk_mech_restart_response(self.h), "dmk_mech_restart_response")
        for k in RESTART_STIFFNESS:
            self.set_scalar(k, d[k])
        self.set_scalar("C_minMaxAvgLastInc", d["C_minMaxAvg"])
        _lib.check(self.lib.dmk_grid_set_restart(self.h, int(restart_inc)), "dmk_grid_set_restart")

    # --- adjoint ------------------------------------------------------------
    def adjoint_tape(self, max_increments: int, segment: int = 0) -> None:
        """Record the current state and, from now on, every converged increment.
        segment > 0: keep only one segment of that many increments on the tape
        plus a checkpoint every segment increments; the reverse sweep runs the
        earlier segments again (increments must come from run_loadcase)."""
        _lib.check(self.lib.dmk_adjoint_tape_segmented(self.h, int(max_increments), int(segment)),
                   "dmk_adjoint_tape_segmented")
        self.taping = True

    def adjoint_record(self) -> None:
        _lib.check(self.lib.dmk_adjoint_record(self.h), "dmk_adjoint_record")

    def adjoint_info(self) -> dict:
        n = C.c_int32()
        _lib.check(self.lib.dmk_adjoint_info(self.h, C.byref(n), None, None), "dmk_adjoint_info")
        P = np.zeros(9 * n.value)
        dt = np.zeros(n.value)
        _lib.check(self.lib.dmk_adjoint_info(self.h, C.byref(n), _lib.dptr(P), _lib.dptr(dt)), "dmk_adjoint_info")
        return {"n": n.value, "P_av": P.reshape(-1, 9), "dt": dt}

    def adjoint_backward(self, P_av_bar: np.ndarray, rtol: float = 1e-12) -> tuple[np.ndarray, np.ndarray]:
        """Cotangents of the recorded P_av (n, 9 column-major) -> per-phase parameter
        cotangents (n_phases, 176) and per-increment solver statistics (n, 4)."""
        n = int(np.asarray(P_av_bar).size // 9)
        Pb = np.ascontiguousarray(P_av_bar, float).reshape(-1)
        tb = np.zeros((len(self.material.phases), ADJ_NOUT))
        st = np.zeros((n, 4))
        _lib.check(self.lib.dmk_adjoint_backward(self.h, _lib.dptr(Pb), float(rtol), _lib.dptr(tb), _lib.dptr(st)),
                   "dmk_adjoint_backward")
        return tb, st

    def adjoint_release(self) -> None:
        _lib.check(self.lib.dmk_adjoint_release(self.h), "dmk_adjoint_release")
        self.taping = False

    def close(self):
        if getattr(self, "h", None):
            self.lib.dmk_mech_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GridMechSpectralPolarization(GridMechSpectralBasic):
    """One RVE on the GPU behind the spectral-polarization plug-in interface
    (``src/grid/grid_mech_spectral_polarization.f90``): the iterate is the pair
    (F, F_tau), convergence adds the curl RMS of F; numerics ``alpha``, ``beta``,
    ``eps_abs_curl``, ``eps_rel_curl``.  The same methods as the basic scheme;
    restart data adds F_tau and F_tau_lastinc (:430-470)."""
    SCHEME = 1
    RESTART_FIELDS = RESTART_FIELDS + ("F_tau", "F_tau_lastinc")

    def adjoint_tape(self, max_increments: int) -> None:
        raise NotImplementedError("the adjoint differentiates the spectral-basic scheme")


class GridMechSpectralGalerkin(GridMechSpectralBasic):
    """One RVE on the GPU behind the spectral-Galerkin plug-in interface
    (``src/grid/grid_mech_spectral_Galerkin.f90``): Newton steps with a
    backtracking line search, each solving G*(K:dF) = G*P by GMRES with the
    matrix-free Jacobian (:607-660); F_aim follows the stress BC once per Newton
    step (:666-681).  Restart data as the basic scheme's."""
    SCHEME = 2

    def adjoint_tape(self, max_increments: int) -> None:
        raise NotImplementedError("the adjoint differentiates the spectral-basic scheme")


# load.yaml solver:mechanical (DAMASK_grid.f90:171-204)
MECHANICAL_SOLVERS = {"spectral_basic": GridMechSpectralBasic,
                      "spectral_polarization": GridMechSpectralPolarization,
                      "spectral_Galerkin": GridMechSpectralGalerkin}


def grid_mechanical(name: str):
    """Solver class for a load file's ``solver: mechanical`` entry."""
    try:
        return MECHANICAL_SOLVERS[name]
    except KeyError:
        from .config import ConfigError
        if name == "FEM":
            raise ConfigError(601, f"{name} is a valid grid solver type of the reference, not built here") from None
        raise ConfigError(601, f"{name} is not a valid grid solver type") from None


class DamaskGrid:
    """Increment/cutback loop of ``DAMASK_grid`` (``DAMASK_grid.f90:288-460``) driven
    from Python, with a callback at every ``f_out`` increment for result output."""

    def __init__(self, solver: GridMechSpectralBasic):
        self.s = solver
        self.t = 0.0
        self.Delta_t = 1.0
        self.Delta_t_prev = 0.0
        self.cutBack = False
        self.cutBackLevel = 0
        self.totalIncs = 0
        self.sta = []

    @staticmethod
    def geometric_fraction(r: float, i: int, N: int) -> float:
        if abs(r - 1.0) <= 1.0e-9:
            return 1.0 / N
        return (1.0 - r) * r**i / (r - r**(N + 1))

    def run(self, loadcases, on_output=None, N_cutback_max: int | None = None):
        maxc = self.s.numerics.N_cutback_max if N_cutback_max is None else N_cutback_max
        for lc in loadcases:
            t_0 = self.t
            guess = lc.estimate_rate
            for inc in range(1, lc.N + 1):
                self.totalIncs += 1
                self.Delta_t_prev = self.Delta_t
                self.Delta_t = lc.t * self.geometric_fraction(lc.r, inc, lc.N) * 2.0**(-self.cutBackLevel)
                stepFraction = 0
                while stepFraction < 2**self.cutBackLevel:
                    t_remaining = lc.t + t_0 - self.t
                    self.t += self.Delta_t
                    stepFraction += 1
                    self.s.forward(self.cutBack, guess, self.Delta_t, self.Delta_t_prev, t_remaining, lc)
                    if not self.cutBack:
                        self.s.materialpoint_forward()
                    sol = self.s.solution()
                    if sol.converged:
                        self.coords = self.s.updateCoords()
                        if self.s.taping:
                            self.s.adjoint_record()
                        self.Delta_t_prev = self.Delta_t
                        self.cutBack = False
                        guess = True
                        self.sta.append((self.totalIncs, self.t, self.cutBackLevel, sol.converged,
                                         sol.iterationsNeeded, 1))
                    elif self.cutBackLevel < maxc:
                        self.cutBack = True
                        stepFraction = (stepFraction - 1) * 2
                        self.cutBackLevel += 1
                        self.t -= self.Delta_t
                        self.Delta_t /= 2.0
                    else:
                        raise _lib.DamaskHipError("IO_error 950: no convergence")
                self.cutBackLevel = max(0, self.cutBackLevel - 1)
                if on_output is not None and inc % lc.f_out == 0:
                    on_output(self.totalIncs, self.t)
