"""Host side of the spectral-polarization and -Galerkin schemes (no GPU): the library's
invSym3333 (math.f90:511-532) against the oracle on tensors without minor
symmetry (as dPdF + C_scale is), the numerics keys and checks of
grid_mech_spectral_polarization.f90:136-153, and the solver selection of
DAMASK_grid.f90:171-204."""
import numpy as np
import pytest


def test_library_inv_sym3333_matches_oracle(oracle):
    from damask_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(7)
    for _ in range(20):
        # stiffness-like: a cubic Hooke tensor plus a non-symmetric perturbation
        A = 1e9 * rng.standard_normal(81)
        for i in range(3):
            A[(i + 3 * i) + 9 * (i + 3 * i)] += 2e11
        got, ref = np.zeros(81), np.zeros(81)
        assert lib.dmk_inv_sym3333(_lib.dptr(A), _lib.dptr(got)) == 0
        assert oracle.lib().orc_inv_sym3333(oracle.ptr(A), oracle.ptr(ref)) == 0
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    sing = np.zeros(81)
    out = np.zeros(81)
    assert lib.dmk_inv_sym3333(_lib.dptr(sing), _lib.dptr(out)) != 0


def test_polarization_numerics():
    from damask_amd import config
    n = config.parse_numerics({"solver": {"grid": {"mechanical": {
        "eps_abs_curl(F)": 1e-9, "eps_rel_curl(F)": 1e-4, "alpha": 1.5, "beta": 0.5}}}})
    assert (n.eps_abs_curl, n.eps_rel_curl, n.alpha, n.beta) == (1e-9, 1e-4, 1.5, 0.5)
    d = config.Numerics()
    assert (d.eps_abs_curl, d.eps_rel_curl, d.alpha, d.beta) == (1.0e-10, 5.0e-4, 1.0, 1.0)
    for key, v in (("alpha", 0.0), ("alpha", 2.5), ("beta", -0.1), ("beta", 2.1), ("eps_abs_curl(F)", 0.0),
                   ("eps_rel_curl(F)", -1.0)):
        with pytest.raises(config.ConfigError) as e:
            config.parse_numerics({"solver": {"grid": {"mechanical": {key: v}}}})
        assert e.value.code == 301


def test_solver_selection():
    from damask_amd import config, grid, workloads
    d = workloads.tension_x(10.0, 5)
    d["solver"]["mechanical"] = "spectral_polarization"
    name, _ = config.parse_load(d)
    assert grid.grid_mechanical(name) is grid.GridMechSpectralPolarization
    assert grid.grid_mechanical("spectral_basic") is grid.GridMechSpectralBasic
    assert grid.grid_mechanical("spectral_Galerkin") is grid.GridMechSpectralGalerkin
    for bad in ("FEM", "n/a"):
        with pytest.raises(config.ConfigError) as e:
            grid.grid_mechanical(bad)
        assert e.value.code == 601
    s = grid.numerics_struct(config.Numerics(alpha=1.25), scheme=1)
    assert s.scheme == 1 and s.alpha == 1.25 and s.beta == 1.0 and s.eps_rel_curl == 5.0e-4


def test_load_case_errors():
    """the load-case checks of DAMASK_grid.f90:661-673 raise IO_error 830:
    a rotation with the Galerkin solver, non-positive r, t, N, f_out and f_restart"""
    import copy as _copy

    from damask_amd import config, workloads
    base = workloads.tension_x(10.0, 5)
    config.parse_load(base)
    rot = _copy.deepcopy(base)
    rot["loadstep"][0]["boundary_conditions"]["mechanical"]["R"] = [1.0, 0.0, 0.0, 30.0]
    config.parse_load(rot)                                   # fine for the basic scheme
    rot["solver"]["mechanical"] = "spectral_Galerkin"
    with pytest.raises(config.ConfigError) as e:
        config.parse_load(rot)
    assert e.value.code == 830
    for where, key, v in (("discretization", "r", 0.0), ("discretization", "t", -1.0),
                          ("discretization", "N", 0), (None, "f_out", 0), (None, "f_restart", 0)):
        bad = _copy.deepcopy(base)
        step = bad["loadstep"][0]
        (step[where] if where else step)[key] = v
        with pytest.raises(config.ConfigError) as e:
            config.parse_load(bad)
        assert e.value.code == 830, (key, v)
