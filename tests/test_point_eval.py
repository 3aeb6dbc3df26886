"""CPU: point and latitude evaluation (SH_to_point_cplx, SH_to_grad_point, SHqst_to_point,
SH_to_lat, SHqst_to_lat; sht_func.c:516-885).  These are host functions (no grid and no GPU
needed).  Checked against direct sums of scipy's spherical harmonics and their analytic
theta derivative, against the oracle's grid synthesis on a Gauss latitude, and at the poles
against the limit from nearby points."""
import numpy as np
import pytest
import scipy.special as sp

import shtns_amd
from shtns_amd.synthetic import spectral_field
from oracle.oracle import Oracle


def dtheta_Y(l, m, th, ph):
    out = 0j
    if m + 1 <= l:
        out += 0.5 * np.sqrt((l - m) * (l + m + 1)) * np.exp(-1j * ph) * sp.sph_harm_y(l, m + 1, th, ph)
    if m - 1 >= -l:
        out -= 0.5 * np.sqrt((l + m) * (l - m + 1)) * np.exp(1j * ph) * sp.sph_harm_y(l, m - 1, th, ph)
    return out


def analytic_qst(s, q, sl, tl, th, ph):
    vr = vt = vp = 0.0
    for l in range(s.lmax + 1):
        for m in range(0, min(l, s.mmax * s.mres) + 1, s.mres):
            f = 1.0 if m == 0 else 2.0
            Y = sp.sph_harm_y(l, m, th, ph)
            dY = dtheta_Y(l, m, th, ph)
            i = s.idx(l, m)
            vr += f * np.real(q[i] * Y)
            vt += f * np.real(sl[i] * dY + tl[i] * 1j * m * Y / np.sin(th))
            vp += f * np.real(sl[i] * 1j * m * Y / np.sin(th) - tl[i] * dY)
    return vr, vt, vp


POINTS = [(0.3, 1.1), (-0.77, 4.0), (0.999, 2.5), (0.0, 0.0)]


def test_point_cplx_matches_scipy():
    s = shtns_amd.sht(12)
    rng = np.random.default_rng(1)
    a = rng.standard_normal(s.nlm_cplx) + 1j * rng.standard_normal(s.nlm_cplx)
    for cost, phi in POINTS + [(1.0, 0.4), (-1.0, 2.0)]:
        th = np.arccos(cost)
        z = sum(a[s.zidx(l, m)] * sp.sph_harm_y(l, m, th, phi) for l in range(13) for m in range(-l, l + 1))
        assert abs(s.SH_to_point_cplx(a, cost, phi) - z) <= 1e-13 * np.linalg.norm(a)


@pytest.mark.parametrize("sizes", [(15, 15, 1), (20, 6, 3)])
def test_qst_point_matches_analytic(sizes):
    s = shtns_amd.sht(*sizes)
    q = spectral_field(sizes[0], s.mmax, seed=2, mres=sizes[2])
    sl = spectral_field(sizes[0], s.mmax, seed=3, mres=sizes[2])
    tl = spectral_field(sizes[0], s.mmax, seed=4, mres=sizes[2])
    for cost, phi in POINTS:
        ref = analytic_qst(s, q, sl, tl, np.arccos(cost), phi)
        got = s.SHqst_to_point(q, sl, tl, cost, phi)
        assert np.allclose(got, ref, rtol=0, atol=1e-13 * (1 + np.abs(ref).max())), (cost, got, ref)
        assert abs(got[0] - s.SH_to_point(q, cost, phi)) <= 1e-14
        g = s.SH_to_grad_point(q, sl, cost, phi)
        z = np.zeros_like(tl)
        h = s.SHqst_to_point(q, sl, z, cost, phi)
        assert np.allclose(g, h, rtol=0, atol=1e-14)


def test_qst_point_at_poles_is_the_limit():
    s = shtns_amd.sht(20)
    q, sl, tl = (spectral_field(20, 20, seed=k) for k in (5, 6, 7))
    for cost, phi in [(1.0, 0.7), (-1.0, 2.2)]:
        near = cost - np.sign(cost) * 5e-15      # theta ~ 1e-7 from the pole
        got = s.SHqst_to_point(q, sl, tl, cost, phi)
        lim = s.SHqst_to_point(q, sl, tl, near, phi)
        assert np.all(np.isfinite(got))
        assert np.allclose(got, lim, rtol=0, atol=1e-5), (got, lim)


@pytest.mark.parametrize("case", [(31, 31, 1, 32, 64), (40, 13, 3, 44, 80)])
def test_lat_matches_oracle_grid(case):
    L, M, mres, nlat, nphi = case
    s = shtns_amd.sht(L, M, mres)
    o = Oracle(L, M, mres, nlat, nphi)
    q = spectral_field(L, M, seed=8, mres=mres)
    sl = spectral_field(L, M, seed=9, mres=mres, power=3.0)
    tl = spectral_field(L, M, seed=10, mres=mres, power=3.0)
    sl[0] = tl[0] = 0
    v = o.synth(q)
    vt, vp = o.synth_sphtor(sl, tl)
    # the grid spans 2 pi / mres in longitude; SH_to_lat's nphi points span 2 pi (order m in
    # Fourier bin m, sht_func.c:790-805), so nphi * mres points start with the grid's
    n = nphi * mres
    for k in (0, 5, nlat // 2):
        x = o.cos_theta[k]
        vr = np.empty(n)
        s.SH_to_lat(q, x, vr)
        assert np.max(np.abs(vr[:nphi] - v[:, k])) <= 1e-13 * np.max(np.abs(v))
        a, b, c = (np.empty(n) for _ in range(3))
        s.SHqst_to_lat(q, sl, tl, x, a, b, c)
        assert np.array_equal(a, vr)
        assert np.max(np.abs(b[:nphi] - vt[:, k])) <= 1e-13 * np.max(np.abs(vt))
        assert np.max(np.abs(c[:nphi] - vp[:, k])) <= 1e-13 * np.max(np.abs(vp))


def test_lat_truncation_and_nphi_clamp():
    """ltr and mtr drop degrees and orders; orders with 2 m >= nphi are dropped."""
    from shtns_amd import _lib
    s = shtns_amd.sht(31)
    o = Oracle(31, 31, 1, 32, 64)
    q = spectral_field(31, 31, seed=11)
    x = o.cos_theta[3]
    lib = _lib.load()
    vr = np.empty(64)
    lib.SH_to_lat(s._cfg, q.ctypes.data, x, vr.ctypes.data, 64, 20, 9)
    qc = q.copy()
    qc[(s.l > 20) | (s.m > 9)] = 0
    assert np.max(np.abs(vr - o.synth(qc)[:, 3])) <= 1e-13 * np.max(np.abs(vr))
    vr8 = np.empty(8)
    s.SH_to_lat(q, x, vr8)       # nphi = 8 keeps m <= 3
    qc = q.copy()
    qc[s.m > 3] = 0
    ref = np.array([s.SH_to_point(qc, x, 2 * np.pi * j / 8) for j in range(8)])
    assert np.max(np.abs(vr8 - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("sizes", [(20, 20, 1), (24, 6, 4)])
def test_zrotate(sizes):
    """The rotated field at longitude phi is the original at phi - alpha; m = 0 is unchanged;
    in place gives the same result."""
    s = shtns_amd.sht(*sizes)
    q = spectral_field(sizes[0], s.mmax, seed=12, mres=sizes[2])
    alpha = 0.8
    r = s.Zrotate(q, alpha)
    assert np.array_equal(r[s.m == 0], q[s.m == 0])
    for cost, phi in POINTS:
        assert abs(s.SH_to_point(r, cost, phi) - s.SH_to_point(q, cost, phi - alpha)) <= 1e-13
    from shtns_amd import _lib
    q2 = q.copy()
    _lib.load().SH_Zrotate(s._cfg, q2.ctypes.data, alpha, q2.ctypes.data)
    assert np.array_equal(q2, r)


def test_legendre_deriv_array_matches_scipy():
    s = shtns_amd.sht(30, 30)
    for im in (0, 1, 7):
        for x in (0.4, -0.9, 1.0):
            th = np.arccos(x)
            yl, dyl = np.empty(31 - im), np.empty(31 - im)
            from shtns_amd import _lib
            lnz = _lib.load().legendre_sphPlm_deriv_array(s._cfg, 30, im, x, np.sqrt(1 - x * x),
                                                          yl.ctypes.data, dyl.ctypes.data)
            assert lnz >= im
            for l in range(im, 31):
                P = np.real(sp.sph_harm_y(l, im, th, 0.0))
                dP = np.real(dtheta_Y(l, im, th, 0.0))
                if im == 0:
                    assert abs(yl[l] - P) <= 1e-13
                elif x != 1.0:
                    assert abs(yl[l - im] * np.sin(th) - P) <= 1e-13
                if x != 1.0 or im <= 1:
                    assert abs(dyl[l - im] - dP) <= 1e-12 * (1 + l)


@pytest.mark.parametrize("sizes", [(20, 20, 1), (24, 6, 2)])
def test_mul_ct_and_st_dt(sizes):
    """cos(theta) f and sin(theta) df/dtheta through SH_mul_mx, checked at points (degree
    lmax is left empty so the products stay within lmax)."""
    s = shtns_amd.sht(*sizes)
    q = spectral_field(sizes[0], s.mmax, seed=13, mres=sizes[2])
    q[s.l == s.lmax] = 0
    rc = s.SH_mul_mx(s.mul_ct_matrix(), q)
    rd = s.SH_mul_mx(s.st_dt_matrix(), q)
    for cost, phi in POINTS:
        f = s.SH_to_point(q, cost, phi)
        assert abs(s.SH_to_point(rc, cost, phi) - cost * f) <= 1e-13
        gt = s.SH_to_grad_point(q, q, cost, phi)[1]
        assert abs(s.SH_to_point(rd, cost, phi) - np.sqrt(1 - cost * cost) * gt) <= 1e-12
