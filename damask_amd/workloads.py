"""Synthetic benchmark inputs: seeded Voronoi polycrystals with the FCC
phenopowerlaw material of ``examples/grid/material.yaml:7-24`` and uniaxial
tension along x (``examples/grid/tensionX.yaml:7-14``)."""
from __future__ import annotations

import copy

import numpy as np
import yaml

from . import geom
from .config import parse_load, parse_material

FCC_PHASE = {
    "lattice": "cF",
    "mechanical": {
        "output": ["F", "P", "F_e", "F_p", "L_p", "O"],
        "elastic": {"type": "Hooke", "C_11": 106.75e9, "C_12": 60.41e9, "C_44": 28.34e9},
        "plastic": {
            "type": "phenopowerlaw", "output": ["xi_sl", "gamma_sl"],
            "N_sl": [12], "n_sl": [20], "a_sl": [2.25], "atol_xi": 1.0,
            "dot_gamma_0_sl": [0.001], "h_0_sl-sl": [75.0e6],
            "h_sl-sl": [1, 1, 1.4, 1.4, 1.4, 1.4, 1.4],
            "xi_0_sl": [31.0e6], "xi_inf_sl": [63.0e6],
        },
    },
}

# pure Schmid law on both families ({110}<111> and {112}<111>)
BCC_PHASE = {
    "lattice": "cI",
    "mechanical": {
        "output": ["F", "P", "F_e", "F_p", "L_p", "O"],
        "elastic": {"type": "Hooke", "C_11": 233.3e9, "C_12": 135.5e9, "C_44": 128.0e9},
        "plastic": {
            "type": "phenopowerlaw", "output": ["xi_sl", "gamma_sl"],
            "N_sl": [12, 12], "n_sl": [20, 20], "a_sl": [2.0, 2.0], "atol_xi": 1.0,
            "dot_gamma_0_sl": [0.001, 0.001], "h_0_sl-sl": [1.0e9, 1.0e9],
            # examples/config/phase/mechanical/plastic/phenopowerlaw_DP-steel-ferrite.yaml
            "h_sl-sl": [1, 1.4, 1, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4,
                        1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4, 1.4],
            "xi_0_sl": [95.0e6, 95.0e6], "xi_inf_sl": [222.0e6, 222.0e6],
        },
    },
}


# the same 24 systems with the non-Schmid coefficients of the reference's
# BCC regression material (tests/integration/resources/plastic_detect_changes/
# material.yaml:286-301, six coefficients of 0.1); in a cI lattice they act on
# the {110}<111> family only (src/crystal.f90:1428-1441)
BCC_NS_PHASE = copy.deepcopy(BCC_PHASE)
BCC_NS_PHASE["mechanical"]["plastic"]["a_non-Schmid"] = [[0.1] * 6]

# BCC_NS_PHASE with the critical resolved shear stresses of that regression
# material (xi_0 = 405.8 MPa, xi_inf = 872.9 MPa).  At the 95 MPa of
# BCC_PHASE the non-Schmid terms make the material-point iterations of most
# cells fail near yield, so full runs cut back over and over (what the
# sub-stepping tests want, and what makes tangent-linear CPU runs of a whole
# load case take hours); with these values yield sets in near 0.6 % strain and
# the iterations converge as for the Schmid phase
BCC_NS_STRONG_PHASE = copy.deepcopy(BCC_NS_PHASE)
BCC_NS_STRONG_PHASE["mechanical"]["plastic"].update({"xi_0_sl": [405.8e6, 405.8e6], "xi_inf_sl": [872.9e6, 872.9e6]})


def material_dict(quaternions: np.ndarray, phase: dict | None = None, phase_name: str = "pheno") -> dict:
    return {
        "homogenization": {"SX": {"N_constituents": 1, "mechanical": {"type": "isostrain"}}},
        "phase": {phase_name: phase or FCC_PHASE},
        "material": [{"homogenization": "SX",
                      "constituents": [{"phase": phase_name, "v": 1.0, "O": [float(x) for x in q]}]}
                     for q in quaternions],
    }


def tension_x(t: float, N: int, f_out: int | None = None) -> dict:
    return {"solver": {"mechanical": "spectral_basic"},
            "loadstep": [{"boundary_conditions": {"mechanical": {
                "dot_F": [[1.0e-3, 0, 0], [0, "x", 0], [0, 0, "x"]],
                "P": [["x", "x", "x"], ["x", 0, "x"], ["x", "x", 0]]}},
                "discretization": {"t": t, "N": N}, "f_out": f_out or N}]}


def deformation_rate(dot_F, t: float, N: int, f_out: int | None = None) -> dict:
    """Load case with every component of dot_F prescribed (no stress BC)."""
    return {"solver": {"mechanical": "spectral_basic"},
            "loadstep": [{"boundary_conditions": {"mechanical": {"dot_F": [[float(x) for x in row] for row in dot_F]}},
                          "discretization": {"t": t, "N": N}, "f_out": f_out or N}]}


def random_symmetric_rate(rng, rate: float = 1.0e-3, max_trace: float = 1.0e-4) -> np.ndarray:
    """Seeded mixed deformation rate: a symmetric dot_F whose deviatoric part is
    drawn from a normal distribution and scaled to the equivalent rate
    sqrt(2/3 dev:dev) = rate, plus a trace drawn uniformly in [-max_trace, max_trace]."""
    A = rng.standard_normal((3, 3))
    D = 0.5 * (A + A.T)
    dev = D - np.trace(D) / 3.0 * np.eye(3)
    dev *= rate / np.sqrt(2.0 / 3.0 * np.sum(dev * dev))
    return dev + rng.uniform(-max_trace, max_trace) / 3.0 * np.eye(3)


def polycrystal(cells, n_grains: int, seed: int, size=(1.0, 1.0, 1.0), phase: dict | None = None):
    """Seeded periodic Voronoi RVE: returns (Grid, material dict)."""
    rng = np.random.default_rng(seed)
    seeds = geom.seeds_from_random(size, n_grains, rng)
    grid = geom.voronoi_tessellation(cells, size, seeds)
    q = geom.rotation_from_random(n_grains, rng)
    return grid, material_dict(q, phase)


def headline(cells: int = 64, n_grains: int = 200, seed: int = 0, t: float = 50.0, N: int = 100):
    """64^3, 200-grain FCC phenopowerlaw tension to 5 % (the headline workload)."""
    grid, mat = polycrystal([cells] * 3, n_grains, seed)
    return grid, parse_material(mat), parse_load(tension_x(t, N))[1]


def dump_yaml(d: dict) -> str:
    return yaml.safe_dump(d, sort_keys=False)
