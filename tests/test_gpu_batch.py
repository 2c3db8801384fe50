"""Batched RVEs (dmk_batch_run / dmk_batch_adjoint through the C-ABI):
concurrent runs reproduce the same RVEs run one at a time, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lockstep", [False, True])
def test_batch_matches_single_runs(gpu_lib, monkeypatch, lockstep):
    """Taped forward runs and gradients of six RVEs of two sizes, as
    concurrent handles or in lockstep (the forward's residual evaluations and
    the reverse sweeps' adjoint operators and dot products in shared rounds,
    the sweeps leaving the lockstep at different times), equal the single
    runs bit for bit."""
    from damask_amd import adjoint, batch, config
    from damask_amd.grid import GridMechSpectralBasic
    monkeypatch.delenv("DMK_LOCKSTEP", raising=False)
    monkeypatch.setenv("DMK_LOCKSTEP_GROUPS", "1")
    items = batch.random_batch(6, master_seed=11, edges=(8, 12), grains=(5, 12), increments=4, strain=0.004)
    b = batch.Batch(items)
    b.tape()
    b.run(threads=4, lockstep=lockstep)
    curves = b.curves()
    Pbar = [np.broadcast_to(np.eye(3), c.shape) * 1e-6 for c in curves]
    grads, _ = b.gradients(Pbar, rtol=1e-8, threads=4, lockstep=lockstep)
    b.close()
    for (grid, mat, lc), c, g, pb in zip(items, curves, grads, Pbar):
        s = GridMechSpectralBasic(grid, mat)
        s.adjoint_tape(2 * lc.N + 16)
        r = s.run_loadcase(lc)
        assert r["rc"] == 0
        c1 = adjoint._P33(s.adjoint_info()["P_av"])
        assert np.array_equal(c1, c)
        g1, _ = adjoint.curve_gradient(s, 0, pb, rtol=1e-8)
        assert np.array_equal(g1, g)
        s.close()


@pytest.mark.parametrize("mode", ["residual", "residual, 2 groups", "residual, partial rounds", "stress"])
def test_lockstep_batch_matches_single_runs(gpu_lib, monkeypatch, mode):
    """dmk_batch_run_lockstep: the RVEs in flight run each residual
    evaluation as one round (mode "residual": stress integration, response
    reductions, transforms with Gamma, dot products and reports, one launch
    per kernel over the round, the transforms per FFT configuration, with the
    odd 7x6x5 grid's configuration among them; with 2 groups,
    DMK_LOCKSTEP_GROUPS=2, the host threads form two lockstep groups whose
    rounds alternate; with partial rounds, DMK_LOCKSTEP_WAIT_US=0, a round
    starts as soon as a waiting RVE's wait runs out, without the RVEs not yet
    submitted; mode "stress", DMK_LOCKSTEP=stress: only the stress
    integrations are shared).  12-system
    FCC RVEs of two sizes and a 24-system BCC RVE with non-Schmid terms on its
    {110}<111> family; RVEs finish at different times and leave the lockstep.
    Curves equal the single runs bit for bit."""
    from damask_amd import adjoint, batch, config, workloads
    if mode == "stress":
        monkeypatch.setenv("DMK_LOCKSTEP", "stress")
    else:
        monkeypatch.delenv("DMK_LOCKSTEP", raising=False)
    monkeypatch.setenv("DMK_LOCKSTEP_GROUPS", "2" if "2 groups" in mode else "1")
    monkeypatch.setenv("DMK_LOCKSTEP_WAIT_US", "0" if "partial" in mode else "")
    from damask_amd.grid import GridMechSpectralBasic
    items = batch.random_batch(7, master_seed=5, edges=(8, 12), grains=(5, 12), increments=4, strain=0.004)
    grid, mat = workloads.polycrystal([7, 6, 5], 6, 21, size=(1.0, 0.9, 0.8), phase=workloads.BCC_NS_STRONG_PHASE)
    _, lcs = config.parse_load(workloads.tension_x(12.0, 3))
    items.append((grid, config.parse_material(mat), lcs[0]))
    b = batch.Batch(items)
    b.tape()
    b.run(threads=5, lockstep=True)
    curves = b.curves()
    b.close()
    for (grid, mat, lc), c in zip(items, curves):
        s = GridMechSpectralBasic(grid, mat)
        s.adjoint_tape(2 * lc.N + 16)
        r = s.run_loadcase(lc)
        assert r["rc"] == 0
        c1 = adjoint._P33(s.adjoint_info()["P_av"])
        s.close()
        assert c1.shape == c.shape
        assert np.array_equal(c1, c), np.abs(c1 - c).max() / np.abs(c1).max()


def test_lockstep_checkpointed_gradients_match(gpu_lib, monkeypatch):
    """Checkpointed tapes (a checkpoint every 2 increments) recorded in
    lockstep rounds; the lockstep reverse sweeps run the earlier segments
    again, so their rounds mix residual evaluations of the recomputed
    increments with adjoint-operator and dot-product requests.  Curves and
    gradients equal those of concurrent handles, bit for bit."""
    from damask_amd import batch
    monkeypatch.delenv("DMK_LOCKSTEP", raising=False)
    monkeypatch.setenv("DMK_LOCKSTEP_GROUPS", "1")
    items = batch.random_batch(5, master_seed=13, edges=(8, 12), grains=(5, 12), increments=5, strain=0.005)
    out = []
    for lockstep in (False, True):
        b = batch.Batch(items)
        b.tape(segment=2)
        b.run(threads=5, lockstep=lockstep)
        curves = b.curves()
        Pbar = [np.broadcast_to(np.eye(3), c.shape) * 1e-6 for c in curves]
        grads, _ = b.gradients(Pbar, rtol=1e-8, threads=5, lockstep=lockstep)
        b.close()
        out.append((curves, grads))
    for c0, c1 in zip(out[0][0], out[1][0]):
        assert np.array_equal(c0, c1)
    for g0, g1 in zip(out[0][1], out[1][1]):
        assert np.all(np.isfinite(g0)) and np.array_equal(g0, g1)
