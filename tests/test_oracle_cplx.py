"""CPU: the oracle's complex-field transforms (SH_to_spat_cplx / spat_cplx_to_SH,
sht_func.c:943-1052) pinned by scipy's complex spherical harmonics (orthonormal, with the
Condon-Shortley phase, Y_l^{-m} = (-1)^m conj(Y_l^m)) summed directly, by round trips, and
by the reduction to the real-field transforms for a real field."""
import numpy as np
import pytest
import scipy.special as sp

from oracle.oracle import Oracle
from shtns_amd.synthetic import spectral_field


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def random_alm(o, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(o.nlm_cplx) + 1j * rng.standard_normal(o.nlm_cplx)


@pytest.mark.parametrize("case", [(12, 12, 16, 32), (12, 5, 16, 16), (20, 20, 24, 48)])
def test_synth_cplx_matches_scipy(case):
    L, M, nlat, nphi = case
    o = Oracle(L, M, 1, nlat, nphi)
    a = random_alm(o, 1)
    th, ph = np.arccos(o.cos_theta), 2 * np.pi * np.arange(nphi) / nphi
    TH, PH = np.meshgrid(th, ph)          # [nphi, nlat]
    zd = np.zeros((nphi, nlat), dtype=np.complex128)
    for l in range(L + 1):
        for m in range(-min(l, M), min(l, M) + 1):
            zd += a[o.lm_cplx(l, m)] * sp.sph_harm_y(l, m, TH, PH)
    assert rel_l2(o.synth_cplx(a), zd) <= 1e-13
    assert rel_l2(o.analys_cplx(zd), a) <= 1e-13


def test_layout_size():
    # nlm_cplx_calc (shtns.h:136): (L+1)^2 for mmax = lmax, and LM_cplx of the last entry
    assert Oracle(12, 12, 1, 16, 32).nlm_cplx == 13 * 13
    o = Oracle(12, 5, 1, 16, 16)
    assert o.nlm_cplx == 6 * 6 + 7 * 11 and o.lm_cplx(12, 5) == o.nlm_cplx - 1 and o.lm_cplx(6, -5) == 36


def test_real_field_reduces_to_real_transform():
    """For a real field the complex coefficients are the real transform's for m >= 0 and
    (-1)^m conj of them for m < 0."""
    o = Oracle(31, 31, 1, 32, 64)
    q = spectral_field(31, 31, seed=3)
    v = o.synth(q)
    assert rel_l2(o.synth_cplx(np.zeros(o.nlm_cplx)) + v, v) == 0
    a = o.analys_cplx(v.astype(np.complex128))
    for l in range(32):
        for m in range(l + 1):
            assert abs(a[o.lm_cplx(l, m)] - q[o.idx(l, m)]) <= 1e-13
            assert abs(a[o.lm_cplx(l, -m)] - (-1) ** m * np.conj(q[o.idx(l, m)])) <= 1e-13
    z = o.synth_cplx(a)
    assert np.max(np.abs(z.imag)) <= 1e-13 and rel_l2(z.real, v) <= 1e-13
