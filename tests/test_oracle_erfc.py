"""Pins the erfc series (src/ewald.jl:41-130): the coefficient literals are regenerated from the
construction the reference documents (Chebyshev discrete cosine transform of
log(erfc(z)/t) + z^2, t = 2/(2+z), in high precision) and the series are checked against
scipy's erfc to the reference's own tolerances (test/ewald_tests.jl:8-29)."""
from __future__ import annotations

import mpmath
import numpy as np
import pytest
from scipy.special import erfc

import oracle as O

mpmath.mp.dps = 50


def _h(z):
    t = 2 / (2 + z)
    return mpmath.log(mpmath.erfc(z) / t) + z * z


def _cheb(f, n):
    """c_j = (2/n) sum_k f(x_k) cos(j pi (k - 1/2)/n), x_k = cos(pi (k - 1/2)/n) (Numerical Recipes chebft)."""
    xs = [mpmath.cos(mpmath.pi * (k - mpmath.mpf(1) / 2) / n) for k in range(1, n + 1)]
    fx = [f(x) for x in xs]
    return [2 * mpmath.fsum(fx[k - 1] * mpmath.cos(j * mpmath.pi * (k - mpmath.mpf(1) / 2) / n) for k in range(1, n + 1)) / n
            for j in range(n)]


def test_erfc_dev_coefficients_regenerate():
    """28-term series on x = 2t - 1 in [-1, 1], z = 4/(x+1) - 2 (src/ewald.jl:41-56)."""
    c = _cheb(lambda x: _h(4 / (x + 1) - 2), 28)
    lit = O.erfc_coefficients("erfc_dev")
    for j in range(28):
        assert abs(float(c[j]) - lit[j]) <= 4 * np.spacing(abs(lit[j])) + 1e-19, (j, float(c[j]), lit[j])


@pytest.mark.parametrize("n,which,tol", [(17, "pair64", 4.0), (8, "pair32", 1.0)])
def test_pair_erfc_coefficients_regenerate(n, which, tol):
    """Series on [0, 4] in the rescaled variable xi (src/ewald.jl:76-113): nodes mapped back by
    x_sub = xlo + (xi + 1)(1 - xlo)/2, z = 4/(x_sub + 1) - 2. The Float32 set is its own n = 8 fit
    rounded to Float32."""
    xmax = mpmath.mpf(4)
    xlo = 2 * (2 / (2 + xmax)) - 1
    # the evaluation code fixes z = 4/(x_sub + 1) - 2 (t = 2/(2+z), x_sub = 2t - 1)
    c = _cheb(lambda xi: _h(4 / (xlo + (xi + 1) * (1 - xlo) / 2 + 1) - 2), n)
    lit = O.erfc_coefficients(which)
    for j in range(n):
        ref = float(c[j]) if which == "pair64" else float(np.float32(float(c[j])))
        sp = np.spacing(abs(lit[j])) if which == "pair64" else float(np.spacing(np.float32(abs(lit[j]))))
        assert abs(ref - lit[j]) <= tol * sp + 1e-20, (j, ref, lit[j])


def test_erfc_dev_matches_scipy():
    for x in np.arange(0, 6.0001, 0.01):
        assert O.erfc_dev(x) == pytest.approx(erfc(x), rel=1e-11)
    assert O.erfc_dev(-1.0) == pytest.approx(erfc(-1.0), rel=1e-11)
    assert O.erfc_dev(8.0) == pytest.approx(erfc(8.0), rel=1e-9)


def test_pair_erfc_matches_scipy_over_fitted_range():
    xs = list(np.linspace(0.0, 4.0, 4001)) + list(np.random.default_rng(2026).random(2000) * 4.0)
    for x in xs:
        assert O.pair_erfc_f64(x) == pytest.approx(erfc(x), rel=1e-12)
        x32 = float(np.float32(x))
        assert O.pair_erfc_f32(x32) == pytest.approx(erfc(x32), rel=2e-6)
