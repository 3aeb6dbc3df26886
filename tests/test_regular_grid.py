"""Grids, checked through the C library on the host (no GPU): the regular grid without poles
(sht_reg_fast / sht_reg_dct, nodes theta_k = (k + 1/2) pi / n, Fejer's first rule), the
regular grid with both poles (sht_reg_poles, Clenshaw-Curtis) and odd Gauss grids.  The
weights must integrate powers of cos(theta) exactly (int_{-1}^{1} x^p dx = 2/(p+1) for even
p, 0 for odd p), a property the rules' formulas do not restate; odd Gauss nodes against
numpy's Gauss-Legendre rule."""
import numpy as np
import pytest

import shtns_amd


@pytest.mark.parametrize("lmax,n", [(31, 64), (63, 128), (100, 250), (255, 512)])
def test_fejer_nodes_and_moments(lmax, n):
    s = shtns_amd.sht(lmax)
    s.set_grid(n, 2 * max(n, 128), flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_reg_fast, polar_opt=0.0)
    x = np.asarray(s.cos_theta)
    assert np.allclose(np.arccos(x), (np.arange(n) + 0.5) * np.pi / n, rtol=0, atol=1e-14)
    w = np.asarray(s.gauss_wts())
    assert w.shape == (n // 2,)
    wf = np.concatenate([w, w[::-1]])
    for j in range(n // 2):
        assert abs((wf * x ** (2 * j)).sum() - 2.0 / (2 * j + 1)) <= 1e-15, j


def test_reg_dct_is_the_same_grid():
    a, b = shtns_amd.sht(31), shtns_amd.sht(31)
    a.set_grid(64, 128, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_reg_fast, polar_opt=0.0)
    b.set_grid(64, 128, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_reg_dct, polar_opt=0.0)
    assert np.array_equal(np.asarray(a.cos_theta), np.asarray(b.cos_theta))
    assert np.array_equal(np.asarray(a.gauss_wts()), np.asarray(b.gauss_wts()))


@pytest.mark.parametrize("n", [9, 17, 64, 65, 257])
def test_clenshaw_curtis_nodes_and_moments(n):
    """sht_reg_poles: theta_k = k pi / (n-1) including both poles; the Clenshaw-Curtis weights
    integrate every power of cos(theta) of degree < n exactly (odd powers to zero)."""
    s = shtns_amd.sht(7)
    s.set_grid(n, 32, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_reg_poles, polar_opt=0.0)
    x = np.asarray(s.cos_theta)
    assert np.allclose(np.arccos(np.clip(x, -1, 1)), np.arange(n) * np.pi / (n - 1), rtol=0, atol=1e-14)
    assert x[0] == 1.0 and x[-1] == -1.0
    w = np.asarray(s.gauss_wts())
    assert w.shape == ((n + 1) // 2,)
    if n & 1:
        w = w.copy()
        w[-1] *= 2          # the equator row's stored pair weight is halved
        wf = np.concatenate([w, w[-2::-1]])
    else:
        wf = np.concatenate([w, w[::-1]])
    for p in range(n):
        exact = 2.0 / (p + 1) if p % 2 == 0 else 0.0
        assert abs((wf * x ** p).sum() - exact) <= 2e-15, p


@pytest.mark.parametrize("n", [3, 33, 101])
def test_odd_gauss_nodes(n):
    """Odd Gauss grids: nodes against numpy's rule, weights against a 40-digit Newton solve
    (numpy's polar weights are off by ~4e-12 at n = 101), exact moments."""
    mpmath = pytest.importorskip("mpmath")
    s = shtns_amd.sht(2)
    s.set_grid(n, 8, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    xr, _ = np.polynomial.legendre.leggauss(n)
    assert np.allclose(np.asarray(s.cos_theta), xr[::-1], rtol=0, atol=2e-15)
    w = np.asarray(s.gauss_wts()).copy()
    w[-1] *= 2
    x = np.asarray(s.cos_theta)
    with mpmath.workdps(40):
        for k in sorted({0, 1, n // 4, n // 2}):
            xk = mpmath.mpf(x[k])
            for _ in range(6):
                dp = n * (xk * mpmath.legendre(n, xk) - mpmath.legendre(n - 1, xk)) / (xk * xk - 1)
                xk -= mpmath.legendre(n, xk) / dp
            dp = n * (xk * mpmath.legendre(n, xk) - mpmath.legendre(n - 1, xk)) / (xk * xk - 1)
            wk = 2 / ((1 - xk * xk) * dp * dp)
            assert abs(w[k] / float(wk) - 1) <= 1e-13, k
    wf = np.concatenate([w, w[-2::-1]])
    for p in range(0, 2 * n, 2):     # exact for degree < 2n
        assert abs((wf * x ** p).sum() - 2.0 / (p + 1)) <= 4e-15, p


def test_allow_padding_adds_none():
    """SHT_ALLOW_PADDING permits padding; none is added (nspat = nlat * nphi)."""
    s = shtns_amd.sht(31)
    s.set_grid(64, 128, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.SHT_ALLOW_PADDING | shtns_amd.sht_gauss,
               polar_opt=0.0)
    assert s.nlat == 64 and s.nphi == 128 and s.nspat == 64 * 128
