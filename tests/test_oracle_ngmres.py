"""Pin the oracle's NGMRES (the restatement of PETSc SNES NGMRES that both the
oracle's spectral solver and the affine entry run) on affine residuals
r(x) = A x - b, where the mathematics fixes what each step must produce:

* the combined candidate x^A = (1 - sum a) x^M + sum a_i x_i must minimise the
  residual of the same combination of residuals over a (numpy's lstsq gives the
  minimiser; for an affine r the residual of x^A is exactly that combination);
* the accepted iterate is x^A when the candidate is selected, x^M otherwise;
* with a history at least as long as the dimension and no restart, the
  iteration is GMRES on A and ends within n + 1 iterations;
* the seeded problems drive both outcomes of the difference-based selection
  and the difference-based restart."""
import numpy as np
import pytest


def _problem(seed, n, scale):
    rng = np.random.default_rng(seed)
    return np.eye(n) + scale * rng.standard_normal((n, n)), rng.standard_normal(n)


# (seed, n, M, scale): seeded problems covering every branch combination
CASES = [(2, 7, 30, 1.0), (27, 3, 3, 3.0), (41, 5, 2, 0.3), (0, 8, 3, 2.0), (117, 3, 3, 2.0),
         (5, 6, 30, 0.3), (9, 4, 1, 0.5), (13, 8, 30, 1.5)]


def _check_records(A, b, rec):
    for k, r in enumerate(rec):
        FM, XM, XA, FA, X = r["FM"], r["XM"], r["XA"], r["FA"], r["X"]
        a, Fh, Xh = r["alpha"], r["Fh"], r["Xh"]
        scale = max(np.abs(FM).max(), np.abs(Fh).max(), 1e-300)
        # residuals are those of the affine map, up to the rounding of A x - b
        ulp = 1e-14 * (np.abs(A).sum(1).max() * max(np.abs(XM).max(), np.abs(XA).max()) + np.abs(b).max())
        assert np.allclose(FM, A @ XM - b, rtol=0, atol=ulp)
        assert np.allclose(FA, A @ XA - b, rtol=0, atol=ulp)
        # x^A is the affine combination with coefficients alpha
        comb_x = (1.0 - a.sum()) * XM + a @ Xh
        mag = abs(1.0 - a.sum()) * np.abs(XM).max() + np.abs(a) @ np.abs(Xh).max(axis=1)
        assert np.allclose(XA, comb_x, rtol=0, atol=1e-12 * (mag + np.abs(XA).max()))
        # alpha minimises |F^M + sum a_i (F_i - F^M)|
        D = (Fh - FM).T
        a_ls, *_ = np.linalg.lstsq(D, -FM, rcond=None)
        best = np.linalg.norm(FM + D @ a_ls)
        # the method solves the normal equations (D^T D) a = -D^T F^M, whose
        # rounding grows with the square of D's (effective) condition number
        sv = np.linalg.svd(D, compute_uv=False)
        sv = sv[sv > sv[0] * 1e-8 * D.shape[1]] if sv[0] > 0 else sv[:1]
        c = sv[0] / sv[-1] if sv[-1] > 0 else 1.0
        slack = 1e-15 * c * c * np.linalg.norm(FM)
        assert np.linalg.norm(FA) <= best * (1.0 + 1e-7) + 10 * ulp + slack, (k, np.linalg.norm(FA), best, c)
        # where alpha is unique, it is the minimiser itself
        if len(sv) == D.shape[1]:
            tol = max(1e-12, 1e-15 * c * c)
            if tol < 1e-4:
                assert np.abs(a - a_ls).max() <= tol * max(np.abs(a_ls).max(), 1.0), (k, c, a, a_ls)
        # the accepted iterate
        assert np.array_equal(X, XA if r["selectA"] else XM)
        assert np.isclose(r["fAnorm"], np.linalg.norm(FA), rtol=1e-12)
        assert np.isclose(r["fMnorm"], np.linalg.norm(FM), rtol=1e-12)


@pytest.mark.parametrize("seed,n,M,scale", CASES)
def test_ngmres_steps_on_affine_residual(oracle, seed, n, M, scale):
    A, b = _problem(seed, n, scale)
    reason, x, rec = oracle.ngmres_affine(A, b, np.zeros(n), M=M, max_it=40, atol=1e-10)
    assert len(rec) > 0
    _check_records(A, b, rec)
    if reason > 0:
        assert np.linalg.norm(A @ x - b) <= 1e-10


def test_ngmres_branches_are_driven(oracle):
    sel_false = restart = converged = 0
    for seed, n, M, scale in CASES:
        A, b = _problem(seed, n, scale)
        reason, _, rec = oracle.ngmres_affine(A, b, np.zeros(n), M=M, max_it=40, atol=1e-10)
        sel_false += sum(not r["selectA"] for r in rec)
        restart += sum(r["restarted"] for r in rec)
        converged += reason > 0
    assert sel_false > 0 and restart > 0 and converged > 0


@pytest.mark.parametrize("seed,n", [(1, 4), (3, 6), (8, 9), (21, 12)])
def test_ngmres_is_gmres_on_linear_problems(oracle, seed, n):
    """full history, well-conditioned A: every candidate is selected, nothing
    restarts, and the residual vanishes within n + 1 iterations (GMRES's
    finite termination)"""
    A, b = _problem(seed, n, 0.3)
    reason, x, rec = oracle.ngmres_affine(A, b, np.zeros(n), M=30, max_it=3 * n, atol=1e-11 * np.linalg.norm(b))
    assert reason > 0
    assert all(r["selectA"] and not r["restarted"] for r in rec)
    assert len(rec) <= n + 1
    assert np.linalg.norm(A @ x - b) <= 1e-11 * np.linalg.norm(b)
    # each candidate's residual is at most that of the best earlier iterate
    for r in rec:
        prev = np.linalg.norm(r["Fh"], axis=1).min()
        assert r["fAnorm"] <= prev * (1.0 + 1e-12)
