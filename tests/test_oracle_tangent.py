"""Pin the oracle's tangent-linear (dual-number) build, the FP64 forward-mode
reference gradient of the adjoint check:

* its values are the plain oracle's run of the same load case (the dual build
  compiles the same C sources with each double carrying a derivative), to the
  few extra iterations the derivative may take to settle;
* its derivatives of the volume-averaged stress history along single
  parameters (phenopowerlaw xi_0, n, an h_sl-sl entry, and the elastic C_11)
  match central differences of the plain oracle, to the differences' own
  truncation and solver-tolerance noise;
* runs in concurrent threads give the same bits as a run on its own (the
  adjoint check evaluates the parameters concurrently).
"""
import copy
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest


def _case():
    from damask_amd import config, workloads
    grid, mat = workloads.polycrystal([4] * 3, 3, 3)
    mat = config.parse_material(mat)
    mat.phases[0].atol_xi, mat.phases[0].atol_gamma = 1e-8, 1e-14
    _, lcs = config.parse_load(workloads.tension_x(6.0, 3))
    num = config.Numerics(eps_abs_div=1e-8, eps_rel_div=1e-11, eps_abs_P=1e-3, eps_rel_P=1e-11,
                          N_iter_max=400, snes_max_it=400, eps_rel_state=1e-13, N_iter_state_max=200,
                          eps_rel_Lp=1e-11, eps_abs_Lp=1e-16, N_iter_Lp_max=200)
    return grid, mat, lcs[0], num


# (descriptor field, index) -> how to perturb the parsed phase
_DIRS = [("xi_0", 0, "xi_0_sl"), ("n", 0, "n_sl"), ("h_sl_sl", 2, "h_sl_sl"), ("C_11", 0, "C_11")]


def _perturbed(mat, attr, idx, rel):
    m2 = copy.deepcopy(mat)
    p = m2.phases[0]
    v = getattr(p, attr)
    if isinstance(v, (list, tuple, np.ndarray)):
        v = list(v)
        h = rel * v[idx]
        v[idx] += h
    else:
        h = rel * v
        v = v + h
    setattr(p, attr, v)
    return m2, h


def test_dual_build_values_and_derivatives(oracle):
    grid, mat, lc, num = _case()
    pr = oracle.Problem(grid, mat, num, threads=1)
    r0 = pr.run_loadcase(lc)
    pr.close()
    assert r0["rc"] == 0
    with ThreadPoolExecutor(len(_DIRS)) as ex:
        res = list(ex.map(lambda d: oracle.tangent_loadcase(grid, mat, num, lc, [oracle.seed_desc(d[0], d[1])],
                                                            threads=1), _DIRS))
    for (field, idx, attr), r in zip(_DIRS, res):
        assert r["rc"] == 0
        # values: the plain oracle's, to the solver tolerance (eps_rel_div 1e-11)
        assert np.abs(r["P_av"] - r0["P_av"]).max() <= 1e-10 * np.abs(r0["P_av"]).max()
        vals = []
        for sgn in (1.0, -1.0):
            m2, h = _perturbed(mat, attr, idx, sgn * 1e-4)
            p2 = oracle.Problem(grid, m2, num, threads=1)
            r2 = p2.run_loadcase(lc)
            p2.close()
            assert r2["rc"] == 0
            vals.append(r2["P_av"])
        fd = (vals[0] - vals[1]) / (2.0 * abs(h))
        dP = r["dP_av"]
        assert np.isfinite(dP).all()
        err = np.abs(dP - fd).max() / np.abs(fd).max()
        print(f"{field}[{idx}]: max|dP_av| {np.abs(dP).max():.4e}, rel. deviation from central differences {err:.2e}")
        # central differences at h = 1e-4 theta: truncation ~1e-8, solver noise
        # (P_av to ~1e-4 Pa) / 2h up to ~1e-5 of the largest entry
        assert err < 2e-4, (field, idx, err)
    # a run on its own gives the same bits as the concurrent one
    alone = oracle.tangent_loadcase(grid, mat, num, lc, [oracle.seed_desc("xi_0", 0)], threads=1)
    assert np.array_equal(alone["dP_av"], res[0]["dP_av"])
    assert np.array_equal(alone["P_av"], res[0]["P_av"])


def test_dual_build_zero_seed_is_the_plain_run(oracle):
    """with no parameter seeded every derivative stays exactly zero and the run
    takes the plain oracle's iterations"""
    grid, mat, lc, num = _case()
    pr = oracle.Problem(grid, mat, num, threads=1)
    r0 = pr.run_loadcase(lc)
    pr.close()
    r = oracle.tangent_loadcase(grid, mat, num, lc, [oracle.PhaseDesc()], threads=1)
    assert r["rc"] == 0
    assert not np.any(r["dP_av"])
    assert np.array_equal(r["stats"], r0["stats"])
    assert np.array_equal(r["P_av"], r0["P_av"])
