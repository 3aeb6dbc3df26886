"""Regular grid without poles (sht_reg_fast / sht_reg_dct): nodes theta_k = (k + 1/2) pi / n
and Fejer's first quadrature rule, checked through the C library on the host (no GPU):
the weights must integrate every even power of cos(theta) below degree n exactly
(int_{-1}^{1} x^{2j} dx = 2 / (2j + 1)), a property the rule's formula does not restate."""
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


def test_reg_poles_rejected():
    s = shtns_amd.sht(31)
    with pytest.raises(RuntimeError):
        s.set_grid(64, 128, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_reg_poles, polar_opt=0.0)


def test_allow_padding_adds_none():
    """SHT_ALLOW_PADDING permits padding; none is added (nspat = nlat * nphi)."""
    s = shtns_amd.sht(31)
    s.set_grid(64, 128, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.SHT_ALLOW_PADDING | shtns_amd.sht_gauss,
               polar_opt=0.0)
    assert s.nlat == 64 and s.nphi == 128 and s.nspat == 64 * 128
