"""Input side of the grid solver: material.yaml, load-case YAML and numerics YAML.

Restates the parts of the reference's configuration handling that the
spectral-basic phenopowerlaw path consumes:

* material.yaml parsing and the cell -> (phase, entry) index maps,
  ``src/material.f90:95-218``;
* phenopowerlaw + Hooke parameters, ``src/phase_mechanical_plastic_phenopowerlaw.f90:146-285``
  and ``src/phase_mechanical_elastic.f90:58-79``;
* load-case parsing incl. masks ('x') and the load-case checks,
  ``src/grid/DAMASK_grid.f90:472-718`` (checks ``:661-673``);
* numerics defaults, ``src/grid/grid_mech_spectral_basic.f90:130-136``,
  ``src/phase_mechanical.f90:281-295``, ``src/grid/DAMASK_grid.f90:147-150``.

Errors raise ``ConfigError`` with the reference's IO_error number where one exists.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Any

import numpy as np
import yaml


class ConfigError(ValueError):
    """Invalid input; ``code`` is the reference's IO_error number."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"IO_error {code}: {msg}")
        self.code = code


LATTICES = {"cF": 0, "cI": 1}
NSLIP_MAX = {"cF": [12, 6], "cI": [12, 12, 24]}


def _load(text_or_dict) -> dict:
    if isinstance(text_or_dict, dict):
        return text_or_dict
    return yaml.safe_load(text_or_dict)


@dataclasses.dataclass
class Phase:
    """One phenopowerlaw phase with Hooke elasticity (cubic lattices)."""
    name: str
    lattice: str
    N_sl: list[int]
    C_11: float
    C_12: float
    C_44: float
    dot_gamma_0_sl: list[float]
    n_sl: list[float]
    a_sl: list[float]
    h_0_sl_sl: list[float]
    xi_0_sl: list[float]
    xi_inf_sl: list[float]
    h_sl_sl: list[float]
    c_1: list[float]
    c_2: list[float]
    f_sat_sl_tw: list[float]
    a_nS: list[float]
    atol_xi: float = 1.0
    atol_gamma: float = 1.0e-6
    output: list[str] = dataclasses.field(default_factory=list)
    output_plastic: list[str] = dataclasses.field(default_factory=list)

    @property
    def nsl(self) -> int:
        return int(sum(self.N_sl))

    def _expand(self, v: list[float]) -> np.ndarray:
        return np.repeat(np.asarray(v, float), self.N_sl)

    def expanded(self) -> dict[str, np.ndarray]:
        """Per-system parameters (math_expand over N_sl)."""
        return {k: self._expand(getattr(self, k)) for k in
                ("dot_gamma_0_sl", "n_sl", "a_sl", "h_0_sl_sl", "xi_0_sl", "xi_inf_sl",
                 "c_1", "c_2", "f_sat_sl_tw")}


def _as_list(v, n, name, default=None):
    if v is None:
        if default is None:
            raise ConfigError(211, f"missing {name}")
        return [default] * n
    v = list(v) if isinstance(v, (list, tuple)) else [v]
    if len(v) != n:
        raise ConfigError(211, f"{name} needs {n} entries")
    return [float(x) for x in v]


def parse_phase(name: str, d: dict) -> Phase:
    lattice = d.get("lattice")
    if lattice not in LATTICES:
        raise ConfigError(130, f"lattice {lattice!r} of phase {name!r} is not cF or cI")
    mech = d["mechanical"]
    el = mech["elastic"]
    if el.get("type") != "Hooke":
        raise ConfigError(200, f"elasticity {el.get('type')!r}")
    pl = mech["plastic"]
    if pl.get("type") != "phenopowerlaw":
        raise ConfigError(200, f"plasticity {pl.get('type')!r} (only phenopowerlaw on this path)")
    N_sl = [int(x) for x in pl.get("N_sl", [])]
    if sum(N_sl) <= 0:
        raise ConfigError(211, "N_sl")
    if any(n < 0 for n in N_sl):
        raise ConfigError(144, "N_sl")
    if any(a > b for a, b in zip(N_sl, NSLIP_MAX[lattice])) or len(N_sl) > len(NSLIP_MAX[lattice]):
        raise ConfigError(145, "N_sl")
    if pl.get("N_tw"):
        raise ConfigError(200, "twinning is not on this path")
    nf = len(N_sl)
    p = Phase(
        name=name, lattice=lattice, N_sl=N_sl,
        C_11=float(el["C_11"]), C_12=float(el["C_12"]), C_44=float(el["C_44"]),
        dot_gamma_0_sl=_as_list(pl.get("dot_gamma_0_sl"), nf, "dot_gamma_0_sl"),
        n_sl=_as_list(pl.get("n_sl"), nf, "n_sl"),
        a_sl=_as_list(pl.get("a_sl"), nf, "a_sl"),
        h_0_sl_sl=_as_list(pl.get("h_0_sl-sl"), nf, "h_0_sl-sl"),
        xi_0_sl=_as_list(pl.get("xi_0_sl"), nf, "xi_0_sl"),
        xi_inf_sl=_as_list(pl.get("xi_inf_sl"), nf, "xi_inf_sl"),
        h_sl_sl=[float(x) for x in pl["h_sl-sl"]],
        c_1=_as_list(pl.get("c_1"), nf, "c_1", 0.0),
        c_2=_as_list(pl.get("c_2"), nf, "c_2", 1.0),
        f_sat_sl_tw=_as_list(pl.get("f_sat_sl-tw"), nf, "f_sat_sl-tw", 0.0),
        a_nS=[float(x) for x in (pl.get("a_non-Schmid") or [[]])[0]] if pl.get("a_non-Schmid") else [],
        atol_xi=float(pl.get("atol_xi", 1.0)),
        atol_gamma=float(pl.get("atol_gamma", 1.0e-6)),
        output=list(mech.get("output", [])),
        output_plastic=list(pl.get("output", [])),
    )
    bad = [k for k in ("dot_gamma_0_sl", "n_sl", "a_sl", "xi_0_sl", "xi_inf_sl")
           if any(v <= 0.0 for v in getattr(p, k))]
    if bad:
        raise ConfigError(211, " ".join(bad))
    if p.atol_xi < 0.0 or p.atol_gamma < 0.0:
        raise ConfigError(211, "atol")
    if p.a_nS and lattice != "cI":
        raise ConfigError(130, "non-Schmid coefficient not implemented for lattice " + lattice)
    if len(p.a_nS) > 6:
        raise ConfigError(130, "more than 6 non-Schmid coefficients for lattice cI")
    return p


@dataclasses.dataclass
class Material:
    """Parsed material.yaml: phases, homogenizations and per-material constituents."""
    phases: list[Phase]
    homogenizations: list[tuple[str, str, int]]     # (name, mechanical type, N_constituents)
    mat_phase: np.ndarray          # (n_materials,) phase index, one constituent
    mat_homog: np.ndarray          # (n_materials,) homogenization index
    mat_O: np.ndarray              # (n_materials, 4) quaternions
    mat_v: np.ndarray              # (n_materials,)

    def phase_index(self, name: str) -> int:
        return [p.name for p in self.phases].index(name)

    def cell_maps(self, material_of_cell: np.ndarray) -> dict[str, np.ndarray]:
        """Cell -> (homogenization, entry) and (phase, entry) maps, ``material.f90:189-216``.

        Entries are 0-based here (the reference stores them 1-based; DADF5's
        ``/cell_to`` stores them 0-based).
        """
        mc = np.asarray(material_of_cell, np.int64)
        if mc.min() < 0 or mc.max() >= len(self.mat_phase):
            raise ConfigError(155, "More materials requested than found in material.yaml")
        ph = self.mat_phase[mc].astype(np.int32)
        ho = self.mat_homog[mc].astype(np.int32)
        ph_entry = np.zeros_like(ph)
        ho_entry = np.zeros_like(ho)
        for i in range(len(self.phases)):
            sel = ph == i
            ph_entry[sel] = np.arange(int(sel.sum()), dtype=np.int32)
        for i in range(len(self.homogenizations)):
            sel = ho == i
            ho_entry[sel] = np.arange(int(sel.sum()), dtype=np.int32)
        return {"phase": ph, "phase_entry": ph_entry, "homogenization": ho,
                "homogenization_entry": ho_entry, "O": self.mat_O[mc], "v": self.mat_v[mc]}


def parse_material(text_or_dict) -> Material:
    d = _load(text_or_dict)
    phase_names = list(d["phase"].keys())
    phases = [parse_phase(n, d["phase"][n]) for n in phase_names]
    homs = []
    for n, h in d["homogenization"].items():
        mt = (h.get("mechanical") or {}).get("type", "pass")
        if mt not in ("pass", "isostrain"):
            raise ConfigError(200, f"homogenization {mt!r} is not on this path")
        nc = int(h["N_constituents"])
        if nc != 1:
            raise ConfigError(148, "only one constituent per material point on this path")
        homs.append((n, mt, nc))
    hom_names = [h[0] for h in homs]
    n = len(d["material"])
    mat_phase = np.zeros(n, np.int32)
    mat_homog = np.zeros(n, np.int32)
    mat_O = np.zeros((n, 4))
    mat_v = np.zeros(n)
    for m, mat in enumerate(d["material"]):
        mat_homog[m] = hom_names.index(mat["homogenization"])
        cons = mat["constituents"]
        if len(cons) != homs[mat_homog[m]][2]:
            raise ConfigError(148, f"material {m}")
        c = cons[0]
        mat_phase[m] = phase_names.index(c["phase"])
        q = np.asarray(c["O"], float)
        if abs(np.linalg.norm(q) - 1.0) > 1.0e-8:
            raise ConfigError(402, "fromQuaternion")
        mat_O[m] = q
        mat_v[m] = float(c["v"])
        if abs(mat_v[m] - 1.0) > 1.0e-9:
            raise ConfigError(153, "constituent")
        V_e = c.get("V_e")
        if V_e is not None and not np.allclose(np.asarray(V_e, float), np.eye(3), rtol=0, atol=0):
            raise ConfigError(200, "V_e other than identity is not on this path")
    return Material(phases, homs, mat_phase, mat_homog, mat_O, mat_v)


# ---------------------------------------------------------------------------- load cases
DEF_TYPES = {"L": 0, "dot_F": 1, "F_dot": 1, "F": 2}
STRESS_TYPES = {"P": 0, "dot_P": 1, "P_dot": 1}


@dataclasses.dataclass
class LoadCase:
    """One entry of ``loadstep`` (``DAMASK_grid.f90:496-600``).

    Tensors are (3,3) with [i, j] = component ij; masks are True where the
    entry is 'x' (not prescribed)."""
    def_type: int
    def_values: np.ndarray
    def_mask: np.ndarray
    stress_type: int            # -1: none
    stress_values: np.ndarray
    stress_mask: np.ndarray
    q_R: np.ndarray
    t: float
    N: int
    r: float = 1.0
    f_out: int = 1
    f_restart: int = 2**31 - 1
    estimate_rate: bool = True


def _masked(rows) -> tuple[np.ndarray, np.ndarray]:
    v = np.zeros((3, 3))
    m = np.zeros((3, 3), bool)
    for i in range(3):
        for j in range(3):
            e = rows[i][j]
            if isinstance(e, str) and e.strip() == "x":
                m[i, j] = True
            else:
                v[i, j] = float(e)
    return v, m


def ax2qu_degrees(ax) -> np.ndarray:
    """``rotations.f90:218-236,670-686`` fromAxisAngle(degrees, P=-1)."""
    ax = np.asarray(ax, float)
    angle = ax[3] * (2.0 * math.acos(-1.0) / 360.0)
    if abs(angle) <= 10.0 * np.finfo(float).tiny:
        return np.array([1.0, 0.0, 0.0, 0.0])
    if abs(np.linalg.norm(ax[:3]) - 1.0) > np.finfo(float).eps or angle < 0.0 or angle > math.pi:
        raise ConfigError(402, "fromAxisAngle")
    c, s = math.cos(angle * 0.5), math.sin(angle * 0.5)
    return np.array([c, ax[0] * s, ax[1] * s, ax[2] * s])


def parse_load(text_or_dict) -> tuple[str, list[LoadCase]]:
    d = _load(text_or_dict)
    solver = (d.get("solver") or {}).get("mechanical", "spectral_basic")
    cases = []
    for l, step in enumerate(d["loadstep"], start=1):
        bc = step.get("boundary_conditions", step.get("boundary_condition"))
        mech = bc["mechanical"]
        def_type = None
        stress_type = -1
        dv = np.zeros((3, 3)); dm = np.ones((3, 3), bool)
        sv = np.zeros((3, 3)); sm = np.ones((3, 3), bool)
        for key, val in mech.items():
            if key in DEF_TYPES:
                if def_type is not None:
                    raise ConfigError(830, f"load case {l}: only one of L, dot_F, or F allowed in mechanical BC")
                def_type = DEF_TYPES[key]
                dv, dm = _masked(val)
            elif key in STRESS_TYPES:
                if stress_type != -1:
                    raise ConfigError(830, f"load case {l}: only one of P or dot_P allowed in mechanical BC")
                stress_type = STRESS_TYPES[key]
                sv, sm = _masked(val)
        q_R = ax2qu_degrees(mech.get("R", [0.0, 0.0, 1.0, 0.0]))
        if solver == "spectral_Galerkin" and "R" in mech:
            raise ConfigError(830, f"load case {l} contains rotation R, which is not supported when using "
                                   "the Galerkin solver")
        if def_type is None:
            raise ConfigError(830, f"load case {l} is incomplete: L, F_dot/dot_F, or F missing")
        if np.any(sm == dm):
            raise ConfigError(830, f"mask consistency for mechanical violated in load case {l}")
        if np.any(~(sm | sm.T) & (np.eye(3) < 1)):
            raise ConfigError(830, f"mixed boundary conditions allow rotation in load case {l}")
        disc = step["discretization"]
        f_out = step.get("f_out", 1)
        f_out = 2**31 - 1 if f_out == "none" else int(f_out)
        lc = LoadCase(def_type, dv, dm, stress_type, sv, sm, q_R, float(disc["t"]), int(disc["N"]),
                      float(disc.get("r", 1.0)), f_out, int(step.get("f_restart", 2**31 - 1)),
                      bool(step.get("estimate_rate", True)))
        if lc.r <= 0.0:
            raise ConfigError(830, f"non-positive ratio for geometric progression in load case {l}")
        if lc.t <= 0.0:
            raise ConfigError(830, f"negative time increment in load case {l}")
        if lc.N < 1:
            raise ConfigError(830, f"non-positive number of increments in load case {l}")
        if lc.f_out < 1:
            raise ConfigError(830, f"non-positive result frequency in load case {l}")
        if lc.f_restart < 1:
            raise ConfigError(830, f"non-positive restart frequency in load case {l}")
        cases.append(lc)
    return solver, cases


# ---------------------------------------------------------------------------- numerics
@dataclasses.dataclass
class Numerics:
    # solver:grid:mechanical (grid_mech_spectral_basic.f90:130-136)
    N_iter_min: int = 1
    N_iter_max: int = 100
    eps_abs_div: float = 1.0e-4
    eps_rel_div: float = 5.0e-4
    eps_abs_P: float = 1.0e3
    eps_rel_P: float = 1.0e-3
    update_gamma: bool = False
    # spectral_polarization only (grid_mech_spectral_polarization.f90:136-147)
    eps_abs_curl: float = 1.0e-10
    eps_rel_curl: float = 5.0e-4
    alpha: float = 1.0
    beta: float = 1.0
    # solver:grid (DAMASK_grid.f90:147-150, spectral_utilities.f90:176-197)
    N_cutback_max: int = 3
    N_staggered_iter_max: int = 10
    divergence_correction: int = 1        # 0 none, 1 grid+size, 2 size
    memory_efficient: bool = True
    # nonlinear solver: 0 NGMRES (the reference's SNES type), 1 Richardson
    solver: int = 0
    ngmres_m: int = 30
    snes_max_it: int = 50
    # phase:mechanical (phase_mechanical.f90:281-295)
    r_cutback_min: float = 1.0e-3
    r_cutback: float = 0.25
    r_increase: float = 1.5
    eps_rel_state: float = 1.0e-6
    N_iter_state_max: int = 20
    N_iter_Lp_max: int = 40
    r_linesearch_Lp: float = 0.5
    eps_rel_Lp: float = 1.0e-6
    eps_abs_Lp: float = 1.0e-8
    f_update_jacobi_Lp: int = 1
    N_iter_Li_max: int = 40
    r_linesearch_Li: float = 0.5
    eps_rel_Li: float = 1.0e-6
    eps_abs_Li: float = 1.0e-8
    f_update_jacobi_Li: int = 1

    def validate(self) -> None:
        bad = []
        if self.eps_abs_div <= 0: bad.append("eps_abs_div(P)")
        if self.eps_rel_div <= 0: bad.append("eps_rel_div(P)")
        if self.eps_abs_P <= 0: bad.append("eps_abs_P")
        if self.eps_rel_P <= 0: bad.append("eps_rel_P")
        if self.eps_abs_curl <= 0: bad.append("eps_abs_curl(F)")
        if self.eps_rel_curl <= 0: bad.append("eps_rel_curl(F)")
        if self.alpha <= 0 or self.alpha > 2: bad.append("alpha")
        if self.beta < 0 or self.beta > 2: bad.append("beta")
        if self.N_iter_max < 1: bad.append("N_iter_max")
        if self.N_iter_min > self.N_iter_max or self.N_iter_min < 1: bad.append("N_iter_min")
        if self.r_cutback_min <= 0: bad.append("r_cutback_min")
        if self.r_cutback <= 0: bad.append("r_cutback")
        if self.r_increase <= 0: bad.append("r_increase")
        if self.r_linesearch_Lp <= 0: bad.append("r_linesearch_Lp")
        if self.eps_rel_Lp <= 0: bad.append("eps_rel_Lp")
        if self.eps_abs_Lp <= 0: bad.append("eps_abs_Lp")
        if self.f_update_jacobi_Lp < 1: bad.append("f_update_jacobi_Lp")
        if self.N_iter_state_max < 1: bad.append("N_iter_state_max")
        if self.N_iter_Lp_max < 1: bad.append("N_iter_Lp_max")
        if self.N_cutback_max < 0: bad.append("N_cutback_max")
        if not (1 <= self.ngmres_m <= 30): bad.append("ngmres_m")
        if bad:
            raise ConfigError(301, " ".join(bad))


def parse_numerics(text_or_dict=None) -> Numerics:
    n = Numerics()
    d = _load(text_or_dict) if text_or_dict else {}
    d = d or {}
    grid = ((d.get("solver") or {}).get("grid") or {})
    gm = grid.get("mechanical") or {}
    for key, attr, typ in (("N_iter_min", "N_iter_min", int), ("N_iter_max", "N_iter_max", int),
                           ("eps_abs_div(P)", "eps_abs_div", float), ("eps_rel_div(P)", "eps_rel_div", float),
                           ("eps_abs_P", "eps_abs_P", float), ("eps_rel_P", "eps_rel_P", float),
                           ("update_gamma", "update_gamma", bool),
                           ("eps_abs_curl(F)", "eps_abs_curl", float), ("eps_rel_curl(F)", "eps_rel_curl", float),
                           ("alpha", "alpha", float), ("beta", "beta", float)):
        if key in gm:
            setattr(n, attr, typ(gm[key]))
    petsc = str(gm.get("PETSc_options", ""))
    toks = petsc.split()
    for i, t in enumerate(toks):
        if t in ("-snes_type", "-mechanical_snes_type") and i + 1 < len(toks):
            n.solver = {"ngmres": 0, "richardson": 1}.get(toks[i + 1], n.solver)
        if t in ("-snes_max_it", "-mechanical_snes_max_it") and i + 1 < len(toks):
            n.snes_max_it = int(toks[i + 1])
        if t in ("-snes_ngmres_m", "-mechanical_snes_ngmres_m") and i + 1 < len(toks):
            n.ngmres_m = int(toks[i + 1])
    if "N_cutback_max" in grid:
        n.N_cutback_max = int(grid["N_cutback_max"])
    fft = grid.get("FFT") or {}
    dc = fft.get("divergence_correction", "grid+size")
    n.divergence_correction = {"none": 0, "grid+size": 1, "size+grid": 1, "size": 2}.get(dc, -1)
    if n.divergence_correction < 0:
        raise ConfigError(301, str(dc))
    if fft.get("derivative", "continuous") != "continuous":
        raise ConfigError(601, "only the continuous spectral derivative is on this path")
    pm = ((d.get("phase") or {}).get("mechanical") or {})
    for key, typ in (("r_cutback_min", float), ("r_cutback", float), ("r_increase", float),
                     ("eps_rel_state", float), ("N_iter_state_max", int)):
        if key in pm:
            setattr(n, key, typ(pm[key]))
    pp = pm.get("plastic") or {}
    for key, typ in (("N_iter_Lp_max", int), ("r_linesearch_Lp", float), ("eps_rel_Lp", float),
                     ("eps_abs_Lp", float), ("f_update_jacobi_Lp", int)):
        if key in pp:
            setattr(n, key, typ(pp[key]))
    pe = pm.get("eigen") or {}
    n.eps_rel_Li = float(pe.get("eps_rel_Li", n.eps_rel_Lp))
    n.eps_abs_Li = float(pe.get("eps_abs_Li", n.eps_abs_Lp))
    n.validate()
    return n
