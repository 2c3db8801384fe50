"""Adjoint parameter gradients of the spectral-basic run (HIP path through the
C-ABI) against central finite differences of the same HIP forward run, and
the torch.autograd front end against the numpy entry point."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(cells=8, grains=4, seed=3, t=8.0, N=5, lattice="cF"):
    from damask_amd import config, workloads
    phase = workloads.BCC_NS_STRONG_PHASE if lattice == "cI" else None
    grid, mat = workloads.polycrystal([cells] * 3, grains, seed, phase=phase)
    mat = config.parse_material(mat)
    mat.phases[0].atol_xi, mat.phases[0].atol_gamma = 1e-8, 1e-14
    _, lcs = config.parse_load(workloads.tension_x(t, N))
    num = config.Numerics(eps_abs_div=1e-8, eps_rel_div=1e-11, eps_abs_P=1e-3, eps_rel_P=1e-11,
                          N_iter_max=400, snes_max_it=400, eps_rel_state=1e-13, N_iter_state_max=200,
                          eps_rel_Lp=1e-11, eps_abs_Lp=1e-16, N_iter_Lp_max=200)
    return grid, mat, lcs, num


def _loss(P, w):
    # weighted sum over the curve of P_xx plus a small P_xy share
    return float(np.sum(w * P[:, 0, 0]) + 0.1 * np.sum(w * P[:, 0, 1]))


def _Pbar(n, w):
    Pb = np.zeros((n, 3, 3))
    Pb[:, 0, 0] = w
    Pb[:, 0, 1] = 0.1 * w
    return Pb


def test_adjoint_gradient_matches_central_differences(gpu_lib):
    from damask_amd import adjoint
    grid, mat, lcs, num = _case()
    s, P = adjoint.run_curve(grid, mat, lcs, num)
    n = P.shape[0]
    assert n == lcs[0].N
    w = np.linspace(1.0, 2.0, n) / 1e6
    g, st = adjoint.curve_gradient(s, 0, _Pbar(n, w))
    s.close()
    print("adjoint stats (its, rel, consistency, substeps):\n", st)
    assert np.all(st[:, 1] < 1e-10), st
    theta0 = adjoint.get_params(mat.phases[0])
    names = adjoint.param_names(mat.phases[0])
    fd = np.zeros_like(theta0)
    for k in range(theta0.size):
        vals = []
        for sgn in (1.0, -1.0):
            th = theta0.copy()
            th[k] *= 1.0 + sgn * 1e-4
            m2 = copy.deepcopy(mat)
            m2.phases[0] = adjoint.set_params(mat.phases[0], th)
            s2, P2 = adjoint.run_curve(grid, m2, lcs, num)
            s2.close()
            vals.append(_loss(P2, w))
        fd[k] = (vals[0] - vals[1]) / (2e-4 * theta0[k])
    for k in range(theta0.size):
        print(f"{names[k]:12s} adjoint {g[k]: .10e}  fd {fd[k]: .10e}  rel {abs(g[k] - fd[k]) / abs(fd[k]):.2e}")
    scale = np.abs(fd * theta0).max()
    assert np.all(np.abs(g - fd) * np.abs(theta0) <= 1e-5 * np.abs(fd * theta0) + 1e-7 * scale)


def test_autograd_curve_loss(gpu_lib):
    torch = pytest.importorskip("torch")
    from damask_amd import adjoint
    grid, mat, lcs, num = _case(N=3)
    theta = torch.tensor(adjoint.get_params(mat.phases[0]), dtype=torch.float64, requires_grad=True)
    P = adjoint.stress_strain_curve(theta, grid, mat, lcs, num)
    target = P.detach() * 1.01
    loss = ((P[:, 0, 0] - target[:, 0, 0]) ** 2).mean()
    loss.backward()
    s, P2 = adjoint.run_curve(grid, mat, lcs, num)
    assert np.allclose(P2, P.detach().numpy(), rtol=0, atol=1e-6 * np.abs(P2).max())
    Pb = np.zeros_like(P2)
    Pb[:, 0, 0] = 2.0 * (P2[:, 0, 0] - target[:, 0, 0].numpy()) / P2.shape[0]
    g, _ = adjoint.curve_gradient(s, 0, Pb)
    s.close()
    assert np.allclose(theta.grad.numpy(), g, rtol=1e-8, atol=1e-12 * np.abs(g).max())


def test_adjoint_gradient_matches_oracle_differences(gpu_lib):
    """The GPU adjoint against central differences of the CPU oracle's own run
    (its restated NGMRES solver and material point), an implementation that
    shares no code with the GPU path"""
    from damask_amd import adjoint
    from oracle import pyoracle
    grid, mat, lcs, num = _case()
    s, P = adjoint.run_curve(grid, mat, lcs, num)
    n = P.shape[0]
    w = np.linspace(1.0, 2.0, n) / 1e6
    g, _ = adjoint.curve_gradient(s, 0, _Pbar(n, w), rtol=1e-12)
    s.close()
    theta0 = adjoint.get_params(mat.phases[0])

    def oracle_loss(th):
        m2 = copy.deepcopy(mat)
        m2.phases[0] = adjoint.set_params(mat.phases[0], th)
        pr = pyoracle.Problem(grid, m2, num, threads=8)
        r = pr.run_loadcase(lcs[0])
        pr.close()
        assert r["rc"] == 0
        return _loss(adjoint._P33(r["P_av"]), w)

    fd = np.zeros_like(theta0)
    for k in range(theta0.size):
        tp, tm = theta0.copy(), theta0.copy()
        tp[k] *= 1.0 + 1e-4
        tm[k] *= 1.0 - 1e-4
        fd[k] = (oracle_loss(tp) - oracle_loss(tm)) / (2e-4 * theta0[k])
    names = adjoint.param_names(mat.phases[0])
    for k in range(theta0.size):
        print(f"{names[k]:12s} adjoint {g[k]: .10e}  oracle fd {fd[k]: .10e}  rel {abs(g[k] - fd[k]) / abs(fd[k]):.2e}")
    scale = np.abs(fd * theta0).max()
    assert np.all(np.abs(g - fd) * np.abs(theta0) <= 1e-5 * np.abs(fd * theta0) + 1e-7 * scale)


@pytest.mark.parametrize("lattice", ["cF", "cI"])
def test_adjoint_gradient_matches_forward_mode(gpu_lib, lattice):
    """The GPU adjoint against the FP64 forward-mode reference gradient: the
    CPU oracle's tangent-linear (dual-number) build, one parameter direction
    per run, exact up to the solvers' tolerances (no difference quotient).
    cF: 12 systems; cI: two families of 12, non-Schmid terms on the {110}<111>
    family (the regression material's coefficients and critical stresses)."""
    from concurrent.futures import ThreadPoolExecutor

    from damask_amd import adjoint
    from oracle import pyoracle
    # the cI phase yields near 0.6 %: four increments to 1.6 % reach well into plastic flow
    grid, mat, lcs, num = _case(cells=4, grains=3, N=4, lattice=lattice, t=8.0 if lattice == "cF" else 16.0)
    s, P = adjoint.run_curve(grid, mat, lcs, num)
    n = P.shape[0]
    w = np.linspace(1.0, 2.0, n) / 1e6
    g, _ = adjoint.curve_gradient(s, 0, _Pbar(n, w), rtol=1e-12)
    s.close()
    ph = mat.phases[0]
    field = {"xi_0_sl": "xi_0", "xi_inf_sl": "xi_inf", "h_0_sl_sl": "h_0", "n_sl": "n", "a_sl": "a"}
    seeds = [(field[attr], f) for attr, _, _ in adjoint.FAMILY_PARAMS for f in range(len(ph.N_sl))]
    seeds += [("h_sl_sl", t) for t in range(len(ph.h_sl_sl))]
    assert len(seeds) == g.size

    def tangent(sd):
        r = pyoracle.tangent_loadcase(grid, mat, num, lcs[0], [pyoracle.seed_desc(*sd)], threads=1)
        assert r["rc"] == 0
        return _loss(adjoint._P33(r["dP_av"]), w)      # the loss is linear in P_av

    with ThreadPoolExecutor(min(len(seeds), 16)) as ex:
        fm = np.array(list(ex.map(tangent, seeds)))
    theta0 = adjoint.get_params(ph)
    names = adjoint.param_names(ph)
    for k in range(g.size):
        print(f"{names[k]:12s} adjoint {g[k]: .10e}  forward mode {fm[k]: .10e}  "
              f"rel {abs(g[k] - fm[k]) / abs(fm[k]):.2e}")
    scale = np.abs(fm * theta0).max()
    assert np.all(np.abs(g - fm) * np.abs(theta0) <= 1e-5 * np.abs(fm * theta0) + 1e-8 * scale)


@pytest.mark.parametrize("segment", [3, 4])
def test_checkpointed_adjoint_equals_full_tape(gpu_lib, segment):
    """Checkpointing with recomputation (dmk_adjoint_tape_segmented): one
    segment on the tape, a checkpoint every `segment` increments, the earlier
    segments run again in the reverse sweep.  11 increments (segments of 3 and
    4 leave a short last one): gradients and adjoint statistics equal the full
    tape's bit for bit, a second reverse sweep gives them again, and the handle
    holds the end-of-run state afterwards."""
    from damask_amd import adjoint
    grid, mat, lcs, num = _case(N=11, t=16.0)
    s, P = adjoint.run_curve(grid, mat, lcs, num)
    n = P.shape[0]
    w = np.linspace(1.0, 2.0, n) / 1e6
    g_full, st_full = adjoint.curve_gradient(s, 0, _Pbar(n, w), rtol=1e-10)
    F_end = s.field("F")
    s.close()
    s2, P2 = adjoint.run_curve(grid, mat, lcs, num, max_increments=n, segment=segment)
    assert np.array_equal(P2, P)
    g_seg, st_seg = adjoint.curve_gradient(s2, 0, _Pbar(n, w), rtol=1e-10)
    assert np.array_equal(g_seg, g_full)
    assert np.array_equal(st_seg, st_full)
    assert np.array_equal(s2.field("F"), F_end)
    g_again, _ = adjoint.curve_gradient(s2, 0, _Pbar(n, w), rtol=1e-10)
    assert np.array_equal(g_again, g_full)
    s2.close()
