"""Parameter gradients of stress-strain curves through the HIP adjoint.

The forward run records every converged increment on the device
(``dmk_adjoint_tape``); the reverse sweep (``dmk_adjoint_backward``) solves,
per increment, the adjoint of the spectral equilibrium and of the converged
material-point increment (implicit-function theorem) and returns the
cotangents of the phenopowerlaw parameters.  Nothing here exists in the
reference; the forward it differentiates is grid_mech_spectral_basic.

Parameter vector of one phase (``param_names``): xi_0, xi_inf, h_0, n, a per
slip family, then h_sl-sl per interaction type -- 12 entries for the
one-family cF phase.
"""
from __future__ import annotations

import copy

import numpy as np

from .config import Material, Numerics, Phase
from .grid import ADJ_NSYS, GridMechSpectralBasic

# (config attribute, name, block in the per-phase cotangent of dmk_adjoint_backward)
FAMILY_PARAMS = (("xi_0_sl", "xi_0", 0), ("xi_inf_sl", "xi_inf", 2), ("h_0_sl_sl", "h_0", 3), ("n_sl", "n", 1),
                 ("a_sl", "a", 4))
H_BLOCK = 6 * ADJ_NSYS


def param_names(phase: Phase) -> list[str]:
    nf = len(phase.N_sl)
    names = [f"{name}[{f}]" for _, name, _ in FAMILY_PARAMS for f in range(nf)]
    return names + [f"h_sl-sl[{t}]" for t in range(len(phase.h_sl_sl))]


def get_params(phase: Phase) -> np.ndarray:
    v = [x for attr, _, _ in FAMILY_PARAMS for x in getattr(phase, attr)]
    return np.asarray(v + list(phase.h_sl_sl), float)


def set_params(phase: Phase, theta) -> Phase:
    p = copy.deepcopy(phase)
    theta = np.asarray(theta, float).reshape(-1)
    nf = len(p.N_sl)
    if theta.size != 5 * nf + len(p.h_sl_sl):
        raise ValueError(f"expected {5 * nf + len(p.h_sl_sl)} parameters, got {theta.size}")
    k = 0
    for attr, _, _ in FAMILY_PARAMS:
        setattr(p, attr, [float(x) for x in theta[k:k + nf]])
        k += nf
    p.h_sl_sl = [float(x) for x in theta[k:]]
    return p


def reduce_cotangent(phase: Phase, tb: np.ndarray) -> np.ndarray:
    """Per-system cotangents (one row of dmk_adjoint_backward's output) ->
    gradient in the layout of ``get_params`` (sum over the systems of a family)."""
    fam = np.repeat(np.arange(len(phase.N_sl)), phase.N_sl)
    m = len(fam)
    out = []
    for _, _, blk in FAMILY_PARAMS:
        per_sys = tb[blk * ADJ_NSYS: blk * ADJ_NSYS + m]
        out += [per_sys[fam == f].sum() for f in range(len(phase.N_sl))]
    out += list(tb[H_BLOCK: H_BLOCK + len(phase.h_sl_sl)])
    return np.asarray(out, float)


def _P33(P9: np.ndarray) -> np.ndarray:
    """(n, 9) column-major -> (n, 3, 3) [i, j]"""
    return np.transpose(np.asarray(P9, float).reshape(-1, 3, 3), (0, 2, 1)).copy()


def _P9(P33: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.transpose(np.asarray(P33, float), (0, 2, 1)).reshape(-1, 9))


def run_curve(grid, material: Material, loadcases, numerics: Numerics | None = None,
              max_increments: int | None = None, segment: int = 0) -> tuple[GridMechSpectralBasic, np.ndarray]:
    """Forward run with a tape; returns the solver (holding the tape) and the
    volume-averaged P of every recorded increment, (n, 3, 3).  segment > 0:
    checkpoints every segment increments instead of the full tape."""
    s = GridMechSpectralBasic(grid, material, numerics)
    cap = max_increments or 2 * sum(lc.N for lc in loadcases) + 16
    s.adjoint_tape(cap, segment)
    for lc in loadcases:
        r = s.run_loadcase(lc)
        if r["rc"] != 0:
            s.close()
            raise RuntimeError("forward run did not converge")
    return s, _P33(s.adjoint_info()["P_av"])


def curve_gradient(solver: GridMechSpectralBasic, phase_index: int, P_bar: np.ndarray,
                   rtol: float = 1e-12) -> tuple[np.ndarray, np.ndarray]:
    """Gradient of sum(P_bar * P_av) with respect to the parameters of one phase,
    and the per-increment adjoint statistics."""
    tb, st = solver.adjoint_backward(_P9(P_bar), rtol)
    return reduce_cotangent(solver.material.phases[phase_index], tb[phase_index]), st


try:
    import torch

    class StressStrainCurve(torch.autograd.Function):
        """theta (parameters of one phase) -> P_av of every recorded increment, with
        the backward pass through dmk_adjoint_backward."""

        @staticmethod
        def forward(ctx, theta, grid, material, loadcases, numerics, phase_index, rtol):
            mat = copy.deepcopy(material)
            mat.phases[phase_index] = set_params(mat.phases[phase_index], theta.detach().cpu().numpy())
            s, P = run_curve(grid, mat, loadcases, numerics)
            ctx.solver, ctx.phase_index, ctx.rtol = s, phase_index, rtol
            return torch.from_numpy(P).to(theta.dtype)

        @staticmethod
        def backward(ctx, gP):
            g, _ = curve_gradient(ctx.solver, ctx.phase_index, gP.detach().cpu().numpy(), ctx.rtol)
            ctx.solver.close()
            return torch.from_numpy(g), None, None, None, None, None, None

    def stress_strain_curve(theta, grid, material, loadcases, numerics=None, phase_index: int = 0,
                            rtol: float = 1e-12):
        return StressStrainCurve.apply(theta, grid, material, loadcases, numerics, phase_index, rtol)
except ImportError:      # torch is optional for the numpy entry points
    pass
