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


def _dtheta_Y(l, m, th, ph):
    """dY_l^m/dtheta from Y_l^{m+-1} (CS-phased orthonormal Y, scipy)."""
    out = np.zeros(np.broadcast(th, ph).shape, dtype=complex)
    if m + 1 <= l:
        out += 0.5 * np.sqrt((l - m) * (l + m + 1)) * np.exp(-1j * ph) * sp.sph_harm_y(l, m + 1, th, ph)
    if m - 1 >= -l:
        out -= 0.5 * np.sqrt((l + m) * (l - m + 1)) * np.exp(1j * ph) * sp.sph_harm_y(l, m - 1, th, ph)
    return out


@pytest.mark.parametrize("case", [(12, 12, 16, 32), (14, 6, 16, 16)])
def test_sphtor_cplx_matches_analytic(case):
    """Vt = dS/dtheta + (1/sin) dT/dphi, Vp = (1/sin) dS/dphi - dT/dtheta over all -l <= m <= l."""
    L, M, nlat, nphi = case
    o = Oracle(L, M, 1, nlat, nphi)
    S, T = random_alm(o, 5), random_alm(o, 6)
    S[0] = T[0] = 0
    TH, PH = np.meshgrid(np.arccos(o.cos_theta), 2 * np.pi * np.arange(nphi) / nphi)
    zt = np.zeros(TH.shape, dtype=complex)
    zp = np.zeros_like(zt)
    for l in range(1, L + 1):
        for m in range(-min(l, M), min(l, M) + 1):
            Y, dY = sp.sph_harm_y(l, m, TH, PH), _dtheta_Y(l, m, TH, PH)
            cs, ct = S[o.lm_cplx(l, m)], T[o.lm_cplx(l, m)]
            zt += cs * dY + ct * 1j * m * Y / np.sin(TH)
            zp += cs * 1j * m * Y / np.sin(TH) - ct * dY
    a, b = o.synth_sphtor_cplx(S, T)
    assert rel_l2(a, zt) <= 1e-13 and rel_l2(b, zp) <= 1e-13
    s2, t2 = o.analys_sphtor_cplx(zt, zp)
    assert rel_l2(s2, S) <= 1e-13 and rel_l2(t2, T) <= 1e-13


def test_sphtor_cplx_real_field_reduces_to_real_transform():
    """Random (not band-limited) real vector field: m >= 0 coefficients equal the real vector
    analysis, m < 0 ones (-1)^m conj; synthesis of a real field's coefficients is real."""
    o = Oracle(31, 31, 1, 32, 64)
    rng = np.random.default_rng(9)
    vt, vp = rng.standard_normal((2, 64, 32))
    sr, tr = o.analys_sphtor(vt, vp)
    sc, tc = o.analys_sphtor_cplx(vt + 0j, vp + 0j)
    for l in range(32):
        for m in range(l + 1):
            for c, r in ((sc, sr), (tc, tr)):
                assert abs(c[o.lm_cplx(l, m)] - r[o.idx(l, m)]) <= 1e-14 * (1 + abs(r[o.idx(l, m)]))
                assert abs(c[o.lm_cplx(l, -m)] - (-1) ** m * np.conj(r[o.idx(l, m)])) <= 1e-14 * (1 + abs(r[o.idx(l, m)]))
    s = spectral_field(31, 31, seed=10, power=3.0)
    t = spectral_field(31, 31, seed=11, power=3.0)
    s[0] = t[0] = 0
    wt, wp = o.synth_sphtor(s, t)
    S = np.zeros(o.nlm_cplx, dtype=complex)
    T = np.zeros_like(S)
    for l in range(32):
        for m in range(l + 1):
            for C, r in ((S, s), (T, t)):
                C[o.lm_cplx(l, m)] = r[o.idx(l, m)]
                C[o.lm_cplx(l, -m)] = (-1) ** m * np.conj(r[o.idx(l, m)])
    a, b = o.synth_sphtor_cplx(S, T)
    assert rel_l2(a, wt) <= 1e-13 and rel_l2(b, wp) <= 1e-13
    assert np.max(np.abs(a.imag)) <= 1e-13 * np.max(np.abs(wt))
