"""Pins the CPU oracle (oracle/sht_oracle.c) before it is trusted as the parity reference.

Checks, each chosen so a plausible mistake (wrong sign, dropped term, swapped index or
component, missing factor 2 for m>0) fails one of them:
  * SHTns's own known-answer test (examples/test_suite.c:96-152), coefficients at 1e-14;
  * Gauss nodes/weights against scipy.special.roots_legendre and the quadrature sums of
    sht_init.c:607-624;
  * normalized P_lm against the mpmath golden fixture (tests/golden, up to l = 4095) and
    against scipy.special.sph_harm_y;
  * synthesis against a brute-force sum of scipy spherical harmonics at every grid point;
  * vector synthesis against the analytic gradient/curl of scipy spherical harmonics;
  * round trips, Parseval's identity, truncation and equatorial parity;
  * the oracle's FFT against numpy.fft.
"""
import json
import os

import numpy as np
import pytest
from scipy.special import roots_legendre, sph_harm_y

from oracle.oracle import Oracle, fft
from shtns_amd.synthetic import spectral_field

HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 30, 49, 64, 96, 97, 192, 256, 384, 3072])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-13 * max(1, np.sqrt(n)))
    assert np.allclose(fft(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-13 * max(1, np.sqrt(n)))


# ---------------------------------------------------------------- Gauss grid
@pytest.mark.parametrize("n", [4, 32, 128, 1536])
def test_gauss_nodes(n):
    """Nodes against scipy; weights against a 40-digit Newton solve at a few nodes (scipy's
    own weights lose ~1e-8 relative accuracy next to the poles at n=1536)."""
    import mpmath as mp
    o = Oracle(min(n - 1, 63), min(n - 1, 63), 1, n, 128)
    x, w = roots_legendre(n)
    assert np.all(np.diff(o.cos_theta) < 0)             # north first
    assert np.max(np.abs(np.sort(o.cos_theta) - x)) < 1e-15
    assert np.max(np.abs(np.sort(o.gauss_w) - np.sort(w))) < 1e-13
    assert abs(o.gauss_w.sum() - 2.0) < 1e-14
    assert abs((o.gauss_w * o.cos_theta ** 2).sum() - 2.0 / 3) < 1e-14
    assert abs((o.gauss_w * o.sin_theta ** 2).sum() - 4.0 / 3) < 1e-14
    with mp.workdps(40):
        for i in sorted({0, 1, n // 4, n // 2 - 1}):
            z = mp.mpf(o.cos_theta[i])
            for _ in range(4):
                p1, p2 = z, mp.mpf(1)
                for l in range(2, n + 1):
                    p1, p2 = ((2 * l - 1) * z * p1 - (l - 1) * p2) / l, p1
                pp = n * (p2 - z * p1)
                z = z - p1 * (1 - z * z) / pp
            p1, p2 = z, mp.mpf(1)
            for l in range(2, n + 1):
                p1, p2 = ((2 * l - 1) * z * p1 - (l - 1) * p2) / l, p1
            wex = 2 * (1 - z * z) / (n * (p2 - z * p1)) ** 2
            assert abs(o.cos_theta[i] - float(z)) < 2e-16
            assert abs(o.gauss_w[i] - float(wex)) < 1e-10 * float(wex)


# ---------------------------------------------------------------- Legendre functions
def test_legendre_golden_high_degree():
    data = json.load(open(os.path.join(HERE, "golden", "legendre_golden.json")))["points"]
    cache = {}
    worst = 0.0
    for p in data:
        l, m, x = p["l"], p["m"], float(p["x"])
        ref = float(p["value"])
        key = (l, m)
        if key not in cache:
            L = max(l, 2)
            cache[key] = Oracle(L, m, 1, L + 1 + ((L + 1) & 1), 2 * m + 2)
        o = cache[key]
        y = o.legendre(m, x, ltop=max(l, m))[l - m]
        # FP64 three-term recurrences lose accuracy like l^2 eps next to the poles
        # (measured: 4.5e-10 relative at l=4095, x=0.9999999); values below the double
        # range must come out as (nearly) zero rather than garbage.
        if abs(ref) > 1e-200:
            err = abs(y - ref) / abs(ref)
            assert err <= 3e-17 * (l + 1) ** 2 + 1e-15, (l, m, x, y, ref)
            worst = max(worst, err / (3e-17 * (l + 1) ** 2 + 1e-15))
        else:
            assert abs(y) <= 1e-190, (l, m, x, y, ref)
    assert worst <= 1.0


def test_legendre_vs_scipy():
    L = 200
    o = Oracle(L, L, 1, L + 2, 2 * L + 2)
    for m in [0, 1, 3, 50, 150, 200]:
        for x in [0.95, 0.3, -0.6, 0.0]:
            y = o.legendre(m, x)
            ref = np.real(sph_harm_y(np.arange(m, L + 1), m, np.arccos(x), 0.0))
            assert np.max(np.abs(y - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


# ---------------------------------------------------------------- known-answer test
def _spat_func_test(theta, phi):   # examples/test_suite.c:90-94
    return (1.0 + 0.01 * np.cos(theta) + 0.1 * (3. * np.cos(theta) ** 2 - 1.0)
            + (np.cos(phi) + 0.3 * np.sin(phi)) * np.sin(theta)
            + (np.cos(2. * phi) + 0.1 * np.sin(2. * phi)) * np.sin(theta) ** 2 * (7.0 * np.cos(theta) ** 2 - 1.0) * 3. / 8.)


def test_known_answer_test_suite():
    """examples/test_suite.c:96-152: lmax=7, mmax=3, nlat=64, nphi=8, tolerance 1e-14."""
    o = Oracle(7, 3, 1, 64, 8)
    th = np.arccos(o.cos_theta)
    ph = 2 * np.pi * np.arange(8) / 8
    v = _spat_func_test(th[None, :], ph[:, None])
    q = o.analys(v)
    ftol = 1e-14
    assert q[0].imag == 0.0
    assert o.idx(1, 1) == 8
    assert abs(q[0].real - np.sqrt(4 * np.pi)) < ftol
    assert abs(q[1].real - 0.01 * np.sqrt(4 * np.pi / 3)) < ftol
    assert abs(q[2].real - 0.1 * np.sqrt(16 * np.pi / 5)) < ftol
    assert abs(q[8].real + np.sqrt(2 * np.pi / 3)) < ftol
    assert abs(q[8].imag - 0.3 * np.sqrt(2 * np.pi / 3)) < ftol
    i42 = o.idx(4, 2)
    assert abs(q[i42].real - 0.5 * np.sqrt(2 * np.pi / 5)) < ftol
    assert abs(q[i42].imag + 0.05 * np.sqrt(2 * np.pi / 5)) < ftol
    others = [i for i in range(3, o.nlm) if i not in (8, i42)]
    assert np.max(np.abs(q[others])) < ftol
    v2 = o.synth(q)
    assert np.max(np.abs(v2 - v)) < ftol


# ---------------------------------------------------------------- brute force synthesis
def _brute_synth(o, q, th, ph):
    TH, PH = np.meshgrid(th, ph)
    out = np.zeros_like(TH)
    for l in range(o.lmax + 1):
        for m in range(0, min(l, o.mmax * o.mres) + 1, o.mres):
            c = q[o.idx(l, m)]
            Y = sph_harm_y(l, m, TH, PH)
            out += (1.0 if m == 0 else 2.0) * np.real(c * Y)
    return out


@pytest.mark.parametrize("case", [(31, 31, 1, 32, 64), (20, 10, 1, 24, 22), (24, 6, 4, 26, 14)])
def test_synth_vs_bruteforce(case):
    lmax, mmax, mres, nlat, nphi = case
    o = Oracle(*case)
    q = spectral_field(lmax, mmax, seed=11, mres=mres)
    th = np.arccos(o.cos_theta)
    ph = 2 * np.pi * np.arange(nphi) / (nphi * mres)
    ref = _brute_synth(o, q, th, ph)
    assert rel(o.synth(q), ref) < 1e-14


# ---------------------------------------------------------------- vector transforms
def _dtheta_Y(l, m, th, ph):
    """dY_l^m/dtheta for the CS-phased orthonormal Y (scipy), from Y_l^{m+-1}:
    dY/dtheta = 1/2 [ sqrt((l-m)(l+m+1)) e^{-i phi} Y^{m+1} - sqrt((l+m)(l-m+1)) e^{i phi} Y^{m-1} ]."""
    out = np.zeros(np.broadcast(th, ph).shape, dtype=complex)
    if m + 1 <= l:
        out += 0.5 * np.sqrt((l - m) * (l + m + 1)) * np.exp(-1j * ph) * sph_harm_y(l, m + 1, th, ph)
    if m - 1 >= -l:
        out -= 0.5 * np.sqrt((l + m) * (l - m + 1)) * np.exp(1j * ph) * sph_harm_y(l, m - 1, th, ph)
    return out


def test_dtheta_identity_vs_finite_difference():
    th, ph, h = 0.7, 0.3, 1e-6
    for (l, m) in [(1, 0), (3, 2), (5, 1), (6, 6)]:
        fd = (sph_harm_y(l, m, th + h, ph) - sph_harm_y(l, m, th - h, ph)) / (2 * h)
        assert abs(_dtheta_Y(l, m, th, ph) - fd) < 1e-8


def test_vector_synth_vs_analytic():
    """Vt = dS/dtheta + (1/sin) dT/dphi ;  Vp = (1/sin) dS/dphi - dT/dtheta  (doc/main.dox:344-352)."""
    L = 15
    o = Oracle(L, L, 1, 16, 32)
    s = spectral_field(L, L, seed=40, power=3.0)
    t = spectral_field(L, L, seed=41, power=3.0)
    s[0] = t[0] = 0
    th = np.arccos(o.cos_theta)
    ph = 2 * np.pi * np.arange(32) / 32
    TH, PH = np.meshgrid(th, ph)
    vt = np.zeros_like(TH)
    vp = np.zeros_like(TH)
    for l in range(1, L + 1):
        for m in range(0, l + 1):
            f = 1.0 if m == 0 else 2.0
            dY = _dtheta_Y(l, m, TH, PH)
            Y = sph_harm_y(l, m, TH, PH)
            cs, ct = s[o.idx(l, m)], t[o.idx(l, m)]
            vt += f * np.real(cs * dY + ct * 1j * m * Y / np.sin(TH))
            vp += f * np.real(cs * 1j * m * Y / np.sin(TH) - ct * dY)
    vt_o, vp_o = o.synth_sphtor(s, t)
    assert rel(vt_o, vt) < 1e-13
    assert rel(vp_o, vp) < 1e-13


@pytest.mark.parametrize("case", [(63, 63, 1, 64, 128), (40, 13, 3, 44, 80)])
def test_vector_round_trip(case):
    o = Oracle(*case)
    s = spectral_field(case[0], case[1], seed=50, mres=case[2], power=3.0)
    t = spectral_field(case[0], case[1], seed=51, mres=case[2], power=3.0)
    s[0] = t[0] = 0
    vt, vp = o.synth_sphtor(s, t)
    s2, t2 = o.analys_sphtor(vt, vp)
    assert rel(s2, s) < 1e-13 and rel(t2, t) < 1e-13


# ---------------------------------------------------------------- invariants
@pytest.mark.parametrize("case", [(127, 127, 1, 128, 256), (63, 20, 2, 64, 96)])
def test_round_trip_and_parseval(case):
    o = Oracle(*case)
    q = spectral_field(case[0], case[1], seed=60, mres=case[2])
    v = o.synth(q)
    assert rel(o.analys(v), q) < 1e-13
    # Parseval (doc/main.dox:312-318): int f^2 = sum C_m |Q|^2, C_0 = 1, C_{m>0} = 2
    m = np.concatenate([np.full(case[0] + 1 - im * case[2], im * case[2]) for im in range(case[1] + 1)])
    energy_spec = np.sum(np.where(m == 0, 1.0, 2.0) * np.abs(q) ** 2)
    dphi = 2 * np.pi / case[4]
    energy_grid = np.sum(v ** 2 * o.gauss_w[None, :]) * dphi
    assert abs(energy_grid - energy_spec) < 1e-12 * energy_spec


def test_truncation_and_parity():
    o = Oracle(63, 63, 1, 64, 128)
    q = spectral_field(63, 63, seed=70)
    ltr = 20
    qt = q.copy()
    l = np.concatenate([np.arange(im, 64) for im in range(64)])
    qt[l > ltr] = 0
    assert rel(o.synth(q, ltr), o.synth(qt)) < 1e-15
    qa = o.analys(o.synth(q), ltr)
    assert np.all(qa[l > ltr] == 0)
    # equatorial parity: keeping only l+m even gives a north/south symmetric field
    mm = np.concatenate([np.full(64 - im, im) for im in range(64)])
    qe = np.where(((l + mm) % 2) == 0, q, 0)
    v = o.synth(qe)
    assert np.max(np.abs(v - v[:, ::-1])) < 1e-13 * np.max(np.abs(v))


# ---------------------------------------------------------------- sensitivity of the checks
def test_checks_reject_planted_mistakes():
    """Plausible mistakes, planted in the oracle's output, must be caught by the checks above."""
    o = Oracle(31, 31, 1, 32, 64)
    q = spectral_field(31, 31, seed=80)
    th = np.arccos(o.cos_theta)
    ph = 2 * np.pi * np.arange(64) / 64
    ref = _brute_synth(o, q, th, ph)
    m = np.concatenate([np.full(32 - im, im) for im in range(32)])
    bad = {
        "no CS phase": np.where(m % 2 == 1, -q, q),
        "conjugated coefficients": np.conj(q),
        "missing factor 2 for m>0": np.where(m > 0, 0.5 * q, q),
        "re/im swapped": q.imag + 1j * q.real,
    }
    for name, qb in bad.items():
        assert rel(o.synth(qb), ref) > 1e-3, name


# ---- polar optimization (sht_init.c:522-557; applied as kernel_SH_to_spat.gen.c:235-257,
#      kernel_spat_to_SH.gen.c:198-201) ----
def _brute_tm(o, eps):
    """tm[im] from its definition: first latitude (from the pole) where max_l |Y_l^m| >= eps,
    never below tm[im-1], rounded down to even -- computed with scipy's harmonics."""
    L, nl2 = o.lmax, o.nlat // 2
    th = np.arccos(o.cos_theta[:nl2])
    tm, it = [0], 0
    for im in range(1, o.mmax + 1):
        m = im * o.mres
        ls = np.arange(m, L + 1)
        while it < nl2 and np.abs(np.real(sph_harm_y(ls, m, th[it], 0.0))).max() < eps:
            it += 1
        tm.append(it)
    return np.array([(t // 2) * 2 for t in tm])


@pytest.mark.parametrize("eps", [1e-10, 1e-4])
def test_polar_tm_vs_definition(eps):
    o = Oracle(63, 63, 1, 96, 128)
    o.set_polar_opt(eps)
    ref = _brute_tm(o, eps)
    assert np.array_equal(o.polar_tm(), ref)
    assert ref.max() > 0   # the threshold skips something at these sizes


def test_polar_skip_semantics():
    """With eps > 0, order m's rows at latitudes k < tm[m] (both hemispheres) are zero in
    synthesis and left out of the analysis; every other value equals the eps = 0 result."""
    L, nlat, nphi, eps = 63, 96, 128, 1e-4
    o0, oe = Oracle(L, L, 1, nlat, nphi), Oracle(L, L, 1, nlat, nphi)
    oe.set_polar_opt(eps)
    tm = oe.polar_tm()
    rng = np.random.default_rng(5)
    for m in (12, 30, 55):
        assert tm[m] > 0
        q = np.zeros(o0.nlm, dtype=complex)
        i0 = o0.idx(m, m)
        q[i0:i0 + L + 1 - m] = rng.standard_normal(L + 1 - m) + 1j * rng.standard_normal(L + 1 - m)
        v0, ve = o0.synth(q), oe.synth(q)
        rows = np.r_[0:tm[m], nlat - tm[m]:nlat]
        keep = np.setdiff1d(np.arange(nlat), rows)
        assert np.abs(ve[:, rows]).max() == 0.0
        assert rel(ve[:, keep], v0[:, keep]) <= 1e-13
    # analysis: order m of A_eps(V) equals order m of A_0(V with those rows zeroed)
    v = rng.standard_normal((nphi, nlat))
    qe = oe.analys(v)
    for m in (0, 12, 30, 55):
        vm = v.copy()
        vm[:, np.r_[0:tm[m], nlat - tm[m]:nlat]] = 0.0
        qm = o0.analys(vm)
        sl = slice(o0.idx(m, m), o0.idx(m, m) + L + 1 - m)
        assert rel(qe[sl], qm[sl]) <= 1e-13


# ---------------------------------------------------------------- other grids
# sht_reg_poles (Clenshaw-Curtis with both poles), Fejer without poles, odd nlat (an equator
# row), pinned the same way: brute-force scipy sums at the grid points, round trips, and for
# the vector functions at the poles the limit of the analytic field at theta -> 0, pi.
@pytest.mark.parametrize("case", [(20, 20, 1, 41, 48, Oracle.REGULAR_POLES), (20, 20, 1, 42, 48, Oracle.REGULAR_POLES),
                                  (24, 12, 2, 25, 30, Oracle.GAUSS), (20, 20, 1, 45, 44, Oracle.REGULAR)])
def test_synth_other_grids_vs_bruteforce(case):
    lmax, mmax, mres, nlat, nphi, grid = case
    o = Oracle(lmax, mmax, mres, nlat, nphi, grid=grid)
    q = spectral_field(lmax, mmax, seed=12, mres=mres)
    th = np.arccos(np.clip(o.cos_theta, -1, 1))
    ph = 2 * np.pi * np.arange(nphi) / (nphi * mres)
    assert rel(o.synth(q), _brute_synth(o, q, th, ph)) < 1e-14
    if grid == Oracle.REGULAR_POLES:
        assert o.cos_theta[0] == 1.0 and o.cos_theta[-1] == -1.0
    # exact quadrature for band-limited data: nlat > 2 lmax for the regular grids
    assert rel(o.analys(o.synth(q)), q) < 1e-13


def _vector_analytic(o, s, t, th, ph):
    TH, PH = np.meshgrid(th, ph)
    vt = np.zeros_like(TH)
    vp = np.zeros_like(TH)
    for l in range(1, o.lmax + 1):
        for m in range(0, l + 1):
            f = 1.0 if m == 0 else 2.0
            dY = _dtheta_Y(l, m, TH, PH)
            Y = sph_harm_y(l, m, TH, PH)
            cs, ct = s[o.idx(l, m)], t[o.idx(l, m)]
            vt += f * np.real(cs * dY + ct * 1j * m * Y / np.sin(TH))
            vp += f * np.real(cs * 1j * m * Y / np.sin(TH) - ct * dY)
    return vt, vp


@pytest.mark.parametrize("nlat", [41, 40])
def test_vector_reg_poles(nlat):
    L, nphi = 15, 32
    o = Oracle(L, L, 1, nlat, nphi, grid=Oracle.REGULAR_POLES)
    s = spectral_field(L, L, seed=42, power=3.0)
    t = spectral_field(L, L, seed=43, power=3.0)
    s[0] = t[0] = 0
    th = np.arccos(np.clip(o.cos_theta, -1, 1))
    ph = 2 * np.pi * np.arange(nphi) / nphi
    vt_o, vp_o = o.synth_sphtor(s, t)
    inner = slice(1, nlat - 1)
    vt, vp = _vector_analytic(o, s, t, th[inner], ph)
    assert rel(vt_o[:, inner], vt) < 1e-13 and rel(vp_o[:, inner], vp) < 1e-13
    # pole rows: the limit of the field (theta = 1e-7 from each pole; the field is smooth)
    for row, tp in ((0, 1e-7), (nlat - 1, np.pi - 1e-7)):
        lt, lp = _vector_analytic(o, s, t, np.array([tp]), ph)
        assert np.max(np.abs(vt_o[:, row] - lt[:, 0])) < 1e-5 * np.max(np.abs(vt))
        assert np.max(np.abs(vp_o[:, row] - lp[:, 0])) < 1e-5 * np.max(np.abs(vp))
    s2, t2 = o.analys_sphtor(vt_o, vp_o)
    assert rel(s2, s) < 1e-13 and rel(t2, t) < 1e-13
