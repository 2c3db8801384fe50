"""The power laws' x^y as the stress kernel evaluates it on the device
(dmk_pow_law through the C-ABI): (tau/xi)^n of the kinetics and
|1 - xi/xi_inf|^a of the hardening (phase_mechanical_plastic_phenopowerlaw.f90
:474-476, :386-387), against a long-double reference (numpy.longdouble, a
64-bit mantissa, so its own rounding is 2^-11 of a double ulp).

Quarter-power path (phases whose exponents are all positive multiples of 1/4
up to 64): y = k/4 is taken as b^p by binary powering, with b = x, p = k/4
(k a multiple of 4), b = sqrt(x), p = k/2 (k even) or b = sqrt(sqrt(x)),
p = k (k odd).  In units u = 2^-53 of relative error, the root carries up to
c_b = 1 (sqrt) or 1.5 (fourth root: half the inner root's error plus its own
rounding), which the power amplifies p times, and the products of the
powering add at most p - 1 more: p c_b + p - 1 units, i.e. at most that many
ulp of the result (an ulp is 1 to 2 u relative), the bound asserted for every
k = 1 ... 256.  The library pow of the other phases is asserted within 2 ulp."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eval(lib, x, y, quarter):
    from damask_amd import _lib
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(x)
    _lib.check(lib.dmk_pow_law(_lib.dptr(x), _lib.dptr(y), _lib.dptr(out), len(x), int(quarter)), "dmk_pow_law")
    return out


def _ulps(out, x, y):
    ref = np.power(x.astype(np.longdouble), y.astype(np.longdouble))
    spacing = np.spacing(np.abs(ref.astype(np.float64))).astype(np.longdouble)
    return (np.abs(out.astype(np.longdouble) - ref) / spacing).astype(np.float64)


def _bases(rng):
    # the ranges the power laws see: tau/xi up to about 1.5, |1 - xi/xi_inf| up to 1
    return np.concatenate([np.logspace(-4.0, np.log10(1.5), 1500), rng.uniform(1e-3, 1.5, 1500)])


def test_quarter_power_against_long_double(gpu_lib):
    assert np.finfo(np.longdouble).nmant >= 63
    rng = np.random.default_rng(3)
    xb = _bases(rng)
    worst = {}
    for k in range(1, 257):
        y = np.full_like(xb, k / 4.0)
        e = _ulps(_eval(gpu_lib, xb, y, True), xb, y)
        if k % 4 == 0:
            p, cb = k // 4, 0.0
        elif k % 2 == 0:
            p, cb = k // 2, 1.0
        else:
            p, cb = k, 1.5
        bound = p * cb + (p - 1) + 1.0
        worst[k] = e.max()
        assert e.max() <= bound, (k / 4.0, e.max(), bound)
    # the regression and benchmark phases' exponents: n = 20, a = 2.25 (FCC), a = 2 (BCC)
    print("max ulp: n=20 %.1f, a=2.25 %.1f, a=2 %.1f, y=63.75 %.1f, over all k %.1f"
          % (worst[80], worst[9], worst[8], worst[255], max(worst.values())), flush=True)


def test_library_pow_against_long_double(gpu_lib):
    rng = np.random.default_rng(5)
    xb = _bases(rng)
    y = rng.uniform(0.5, 40.0, xb.size)
    e = _ulps(_eval(gpu_lib, xb, y, False), xb, y)
    assert e.max() <= 2.0, e.max()


def test_power_law_special_cases(gpu_lib):
    """zero, negative, infinite and NaN bases follow pow's rules on both paths"""
    x = np.array([0.0, -2.0, -2.0, -2.0, np.inf, np.nan, 1.0])
    y = np.array([20.0, 3.0, 20.0, 2.25, 2.25, 20.0, 63.75])
    with np.errstate(invalid="ignore"):
        want = np.power(x, y)
    for quarter in (True, False):
        got = _eval(gpu_lib, x, y, quarter)
        assert np.array_equal(np.isnan(got), np.isnan(want)), (quarter, got)
        fin = ~np.isnan(want)
        assert np.array_equal(got[fin], want[fin]), (quarter, got)


def test_quarter_path_rejects_other_exponents(gpu_lib):
    from damask_amd import _lib
    with pytest.raises(_lib.DamaskHipError):
        _eval(gpu_lib, np.array([0.5]), np.array([20.1]), True)
