"""The benchmark configurations at their full sizes on the GPU.

* 16^3, 20 grains, 20 increments to 1 %: the whole run against the oracle.
* 64^3, 200 grains, 100 increments to 5 %: increments 1-3 against the oracle
  (all host threads), then the whole run with the convergence criteria of
  grid_mech_spectral_basic.f90:422-458 checked at every increment.
* 256 RVEs (edge 8-32, 10-50 grains, xi_0 and n per RVE, tension along a
  random direction, 40 increments): the batched forward and adjoint; a seeded
  sample of 16 RVEs reproduces its single runs bit for bit.
* 32^3, 50 grains, 60 increments to 3 %: the adjoint gradient of the curve
  loss against central differences at a relative step of 1e-6.
"""
import copy
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _workload(cells, grains, t, N, seed=0, phase=None):
    from damask_amd import config, workloads
    grid, mat = workloads.polycrystal([cells] * 3, grains, seed, phase=phase)
    _, lcs = config.parse_load(workloads.tension_x(t, N))
    return grid, config.parse_material(mat), lcs[0]


def _sub(lc, n):
    s = copy.copy(lc)
    s.t, s.N = lc.t / lc.N * n, n
    return s


def _per_point(g, o, N):
    """F, P, xi and O per point, GPU against oracle"""
    F_g = g.field("F")
    F_o = o.solver_get(0).reshape(-1, 3, 3).transpose(0, 2, 1)
    assert np.abs(F_g - F_o).max() < 1e-8
    P_g = g.field("P")
    P_o = o.get(1).reshape(-1, 3, 3).transpose(0, 2, 1)
    rel = np.abs(P_g - P_o).max(axis=(1, 2)) / np.abs(P_o).max(axis=(1, 2))
    assert rel.max() < 1e-6, rel.max()
    xg = g.field("xi_sl")
    xo = o.get(8, xg.shape[1]).reshape(N, -1)
    assert (np.abs(xg - xo) / xo).max() < 1e-6
    Og = g.field("O")
    Oo = o.get(7, 4).reshape(-1, 4)
    assert np.abs(Og - Oo).max() < 1e-6


def test_16cube_20_grains_full_run(gpu_lib, oracle):
    """16^3, 20 grains, tension 20 increments to 1 % (default numerics:
    eps_abs_div 1e-4, eps_rel_P 1e-3): same iteration counts as the oracle, P_av
    and F_aim within 1e-8, F, P, xi and O per point"""
    from damask_amd.grid import GridMechSpectralBasic
    grid, mat, lc = _workload(16, 20, 10.0, 20)
    g = GridMechSpectralBasic(grid, mat)
    o = oracle.Problem(grid, mat, threads=min(16, os.cpu_count() or 1))
    rg = g.run_loadcase(lc)
    ro = o.run_loadcase(lc)
    assert rg["rc"] == ro["rc"] == 0
    print("iterations gpu", rg["stats"][:, 2].tolist(), "oracle", ro["stats"][:, 2].tolist())
    assert np.array_equal(rg["stats"][:, 1:], ro["stats"][:, 1:])
    assert _relmax(rg["P_av"], ro["P_av"]) < 1e-8
    assert np.abs(rg["F_aim"] - ro["F_aim"]).max() < 1e-8 * np.abs(ro["F_aim"]).max()
    _per_point(g, o, grid.n_cells)
    g.close()
    o.close()


def test_64cube_200_grains(gpu_lib, oracle):
    """64^3, 200 grains, 100 increments to 5 %: increments 1-3 against the
    oracle; then the whole run with the reference's convergence test
    (err_div < max(max|P_av| eps_rel_div, eps_abs_div), err_BC likewise with
    eps_rel_P / eps_abs_P, grid_mech_spectral_basic.f90:436-441) holding at
    every increment, no cutback, and a monotone tensile P_xx"""
    from damask_amd.config import Numerics
    from damask_amd.grid import DamaskGrid, GridMechSpectralBasic
    grid, mat, lc = _workload(64, 200, 50.0, 100)
    num = Numerics()
    g = GridMechSpectralBasic(grid, mat, num)
    o = oracle.Problem(grid, mat, num, threads=min(16, os.cpu_count() or 1))
    first = _sub(lc, 3)
    rg = g.run_loadcase(first)
    ro = o.run_loadcase(first)
    o_fields = {"F": o.solver_get(0), "P": o.get(1), "xi": o.get(8, g.nsl)}
    o.close()
    assert rg["rc"] == ro["rc"] == 0
    print("increments 1-3, iterations gpu", rg["stats"][:, 2].tolist(), "oracle", ro["stats"][:, 2].tolist())
    assert np.array_equal(rg["stats"][:, 1:], ro["stats"][:, 1:])
    assert _relmax(rg["P_av"], ro["P_av"]) < 1e-8
    assert np.abs(g.field("F") - o_fields["F"].reshape(-1, 3, 3).transpose(0, 2, 1)).max() < 1e-8
    P_o = o_fields["P"].reshape(-1, 3, 3).transpose(0, 2, 1)
    assert (np.abs(g.field("P") - P_o).max(axis=(1, 2)) / np.abs(P_o).max(axis=(1, 2))).max() < 1e-6
    xo = o_fields["xi"].reshape(grid.n_cells, -1)
    assert (np.abs(g.field("xi_sl") - xo) / xo).max() < 1e-6
    g.close()

    s = GridMechSpectralBasic(grid, mat, num)
    drv = DamaskGrid(s)
    rec = []

    def check(inc, t):
        err_div, err_BC = s.scalar("err")
        P = s.scalar("P_av")
        pmax = np.abs(P).max()
        tol_div = max(pmax * num.eps_rel_div, num.eps_abs_div)
        tol_BC = max(pmax * num.eps_rel_P, num.eps_abs_P)
        rec.append((inc, err_div / tol_div, err_BC / tol_BC, P[0, 0]))
        assert err_div < tol_div and err_BC < tol_BC, (inc, err_div, tol_div, err_BC, tol_BC)
        assert np.all(np.isfinite(P))

    lc1 = copy.copy(lc)
    lc1.f_out = 1
    drv.run([lc1], on_output=check)
    s.close()
    rec = np.array(rec)
    print("worst err_div/tol", rec[:, 1].max(), "worst err_BC/tol", rec[:, 2].max(), "final P_xx", rec[-1, 3])
    assert len(rec) == 100 and abs(drv.t - 50.0) < 1e-9
    assert all(s_[2] == 0 for s_ in drv.sta)                 # no cutback
    assert np.all(np.diff(rec[:, 3]) > 0.0)                  # hardening: P_xx rises


@pytest.mark.parametrize("lockstep", [False, True])
def test_batch_256_rves(gpu_lib, monkeypatch, lockstep):
    """the 256-RVE batched forward + adjoint, as 16 concurrent handles or in
    lockstep rounds with all 256 RVEs in flight (one host thread each, the
    bench's path); 16 RVEs drawn with a seeded generator rerun alone give the
    same curves and gradients bit for bit"""
    from damask_amd import adjoint, batch
    from damask_amd.grid import GridMechSpectralBasic
    monkeypatch.setenv("DMK_BATCH_PROGRESS", "1")
    for v in ("DMK_LOCKSTEP", "DMK_LOCKSTEP_GROUPS", "DMK_LOCKSTEP_WAIT_US"):
        monkeypatch.delenv(v, raising=False)
    threads = 256 if lockstep else 16
    items = batch.random_batch(256, 0)
    edges = sorted({int(g.cells[0]) for g, _, _ in items})
    assert edges == [8, 12, 16, 20, 24, 32]
    b = batch.Batch(items)
    b.tape()
    r = b.run(threads=threads, lockstep=lockstep)
    curves = b.curves()
    assert all(c.shape == (40, 3, 3) for c in curves)
    Pbar = []
    for c in curves:
        p = np.zeros_like(c)
        p[:, 0, 0] = 2.0 * (c[:, 0, 0] - 1.01 * c[:, 0, 0]) / c.shape[0]
        Pbar.append(p)
    grads, rb = b.gradients(Pbar, rtol=1e-6, threads=threads, lockstep=lockstep)
    b.close()
    print(f"batch forward with tape {r['wall_s']:.2f} s, adjoint {rb['wall_s']:.2f} s")
    assert all(np.all(np.isfinite(gr)) for gr in grads)
    pick = np.random.default_rng(2024).choice(256, 16, replace=False)
    for i in pick:
        grid, mat, lc = items[i]
        s = GridMechSpectralBasic(grid, mat)
        s.adjoint_tape(2 * lc.N + 16)
        assert s.run_loadcase(lc)["rc"] == 0
        c1 = adjoint._P33(s.adjoint_info()["P_av"])
        assert np.array_equal(c1, curves[i]), i
        g1, _ = adjoint.curve_gradient(s, 0, Pbar[i], rtol=1e-6)
        assert np.array_equal(g1, grads[i]), i
        s.close()


def _tight():
    """Solver settings for difference quotients at a relative step of 1e-6.
    Measured at 32^3: the divergence of the FP64 transforms levels off near
    2e-12 |P_av| after about 20 NGMRES iterations, so a tolerance near it
    stops the two runs of a quotient after different iteration counts, and
    the jump of P_av between two counts, amplified by the 1e-6 step, swamps the
    derivative.  Every increment therefore runs N_iter_min = 40 iterations
    (grid_mech_spectral_basic.f90:436-441 requires totalIter >= N_iter_min
    before the tolerances count), by which both runs sit at the rounding floor,
    and the tolerances (1e-9 of |P_av|) are then met; the material-point
    iterations converge to 1e-11 of L_p and 1e-13 of the state."""
    from damask_amd.config import Numerics
    return Numerics(N_iter_min=40, eps_abs_div=1e-8, eps_rel_div=1e-9, eps_abs_P=1e-3, eps_rel_P=1e-9,
                    N_iter_max=400, snes_max_it=400, eps_rel_state=1e-13, N_iter_state_max=200,
                    eps_rel_Lp=1e-11, eps_abs_Lp=1e-16, N_iter_Lp_max=200)


def test_32cube_adjoint_against_central_differences(gpu_lib):
    """32^3, 50 grains, tension 60 increments to 3 %; loss = mean squared
    deviation of the P_xx curve from the curve at seeded parameters perturbed
    within +-10 %.  The adjoint gradient with respect to the 12 parameters
    (xi_0, xi_inf, h_0, n, a, h_sl-sl[7]) against central differences of the
    GPU forward run, both runs of a quotient with the same spectral iteration
    counts in every increment (40, see _tight):

    * at a relative step of 1e-4 every parameter within 1e-5 relative;
    * at a relative step of 1e-6 every parameter within 1e-5 relative, or,
      where the quotient's own noise is larger, within three times that noise.
      The noise is measured, not assumed: P_av after 40 and after 41
      iterations of every increment differs by the FP64 floor of the spectral
      solve, and that difference, carried through the loss and divided by the
      step, bounds what a quotient can resolve for the parameter.  The
      parameters the noise limits are printed."""
    from damask_amd import adjoint
    from damask_amd.grid import GridMechSpectralBasic
    grid, mat, lc = _workload(32, 50, 30.0, 60)
    mat.phases[0].atol_xi, mat.phases[0].atol_gamma = 1e-8, 1e-14
    num = _tight()
    theta0 = adjoint.get_params(mat.phases[0])
    rng = np.random.default_rng(1)
    theta_t = theta0 * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, theta0.size))

    def curve(theta, nm=num):
        m2 = copy.deepcopy(mat)
        m2.phases[0] = adjoint.set_params(mat.phases[0], theta)
        s = GridMechSpectralBasic(grid, m2, nm)
        r = s.run_loadcase(lc)
        s.close()
        assert r["rc"] == 0, f"forward run did not converge (increments done: {int((r['stats'][:, 0] > 0).sum())})"
        return adjoint._P33(r["P_av"])[:, 0, 0], r["stats"][:, 2]

    target, _ = curve(theta_t)
    s, P = adjoint.run_curve(grid, mat, [lc], num)
    n = P.shape[0]
    assert n == 60
    Pb = np.zeros_like(P)
    Pb[:, 0, 0] = 2.0 * (P[:, 0, 0] - target) / n
    g, st = adjoint.curve_gradient(s, 0, Pb, rtol=1e-12)
    s.close()
    # the floor: the same run with one more iteration per increment
    num41 = _tight()
    num41.N_iter_min = 41
    c40, _ = curve(theta0)
    c41, i41 = curve(theta0, num41)
    assert np.all(i41 == 41)
    dP = np.abs(c41 - c40)
    dL = 2.0 * np.mean(np.abs(c40 - target) * dP)            # loss change the floor can cause
    print(f"target and adjoint done; spectral floor max |dP_xx|/|P_xx| = {(dP / np.abs(c40)).max():.2e}", flush=True)
    names = adjoint.param_names(mat.phases[0])
    for h in (1e-4, 1e-6):
        fd = np.zeros_like(theta0)
        for k in range(theta0.size):
            tp, tm = theta0.copy(), theta0.copy()
            tp[k] *= 1.0 + h
            tm[k] *= 1.0 - h
            cp, ip = curve(tp)
            cm, im = curve(tm)
            assert np.array_equal(ip, im) and np.all(ip == 40)
            fd[k] = (np.mean((cp - target) ** 2) - np.mean((cm - target) ** 2)) / (2.0 * h * theta0[k])
        rel = np.abs(g - fd) / np.abs(fd)
        noise = dL / (h * np.abs(theta0) * np.abs(fd))
        print(f"step {h:g}:", flush=True)
        for k in range(theta0.size):
            print(f"  {names[k]:12s} adjoint {g[k]: .12e}  fd {fd[k]: .12e}  rel {rel[k]:.2e}  "
                  f"quotient noise {noise[k]:.1e}")
        if h == 1e-4:
            assert np.all(rel < 1e-5), dict(zip(names, rel))
        else:
            limited = [names[k] for k in range(theta0.size) if rel[k] >= 1e-5]
            print("  limited by the quotient's noise:", limited or "none")
            assert np.all((rel < 1e-5) | (rel < 3.0 * noise)), dict(zip(names, rel))


def test_mixed_load_batch_64x48_checkpointed(gpu_lib, monkeypatch):
    """64 RVEs of 48^3, 300 grains, the 24-system BCC phase, seeded random
    symmetric deformation rates, 50 increments: batched forward on tapes with a
    checkpoint every 8 increments, batched gradient of the curve loss (the
    reverse sweep runs the earlier segments again).  Two RVEs drawn with a
    seeded generator, rerun alone on a full tape, give the same curves and
    gradients bit for bit."""
    from damask_amd import adjoint, batch
    from damask_amd.grid import GridMechSpectralBasic
    monkeypatch.setenv("DMK_BATCH_PROGRESS", "1")
    items = batch.mixed_load_batch(64, 0, edge=48, grains=300, increments=50)
    b = batch.Batch(items)
    b.tape(segment=8)
    r = b.run(threads=16)
    curves = b.curves()
    assert all(c.shape == (50, 3, 3) for c in curves)
    Pbar = [2.0 * (c - 1.01 * c) / c.shape[0] for c in curves]
    grads, rb = b.gradients(Pbar, rtol=1e-6, threads=16)
    b.close()
    print(f"forward with checkpointed tapes {r['wall_s']:.1f} s, reverse sweeps {rb['wall_s']:.1f} s")
    assert all(np.all(np.isfinite(gr)) and np.abs(gr).max() > 0 for gr in grads)
    for i in np.random.default_rng(7).choice(64, 2, replace=False):
        grid, mat, lc = items[i]
        s = GridMechSpectralBasic(grid, mat)
        s.adjoint_tape(2 * lc.N + 16)
        assert s.run_loadcase(lc)["rc"] == 0
        assert np.array_equal(adjoint._P33(s.adjoint_info()["P_av"]), curves[i]), i
        g1, _ = adjoint.curve_gradient(s, 0, Pbar[i], rtol=1e-6)
        assert np.array_equal(g1, grads[i]), i
        s.close()
