"""GPU: Legendre transforms at one order, no FFT: SH_to_spat_ml / spat_to_SH_ml (shtns.h:329-334,
sht_com.c:96-102) and the Python SH_to_spat_m / spat_to_SH_m (shtns_numpy.i:662-671).
The scales are those the reference's own uses fix: the round trip of time_SHT.c:929-930,
and the complex transforms (sht_func.c:943-1052), where SH_to_spat_ml gives Fourier row m of
a synthesis before the FFT and spat_to_SH_ml / nphi gives the analysis from the FFT's row m.
Checked here against the oracle's full transforms of single-order fields."""
import ctypes

import numpy as np
import pytest

import shtns_amd
from shtns_amd import _lib
from shtns_amd.synthetic import spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def make(lmax, mmax, mres, nlat, nphi):
    s = shtns_amd.sht(lmax, mmax, mres)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s, Oracle(lmax, mmax, mres, nlat, nphi)


@pytest.mark.parametrize("case", [(63, 63, 1, 96, 192), (60, 15, 4, 64, 48), (255, 255, 1, 384, 768)])
@pytest.mark.parametrize("which", ["zero", "low", "high"])
def test_ml_against_oracle(case, which, device):
    L, M, mres, nlat, nphi = case
    s, o = make(L, M, mres, nlat, nphi)
    im = {"zero": 0, "low": 1, "high": M}[which]
    m = im * mres
    q = spectral_field(L, M, seed=90 + im, mres=mres)
    sel = s.m == m
    ql = np.ascontiguousarray(q[sel])
    assert ql.size == L + 1 - m
    # synthesis: Fourier row m of the oracle's synthesis of the single-order field
    qm = np.zeros_like(q)
    qm[sel] = q[sel]
    v = o.synth(qm)                               # [nphi, nlat]
    # phi spans 2 pi / mres, so FFT bin im holds order m; v = G_0 at m = 0 and
    # v = 2 Re(G_m e^{i m phi}) = G_m e^{i m phi} + conj: bin im = nphi G_m
    row = np.fft.fft(v, axis=0)[im] / nphi
    G = np.empty(nlat, dtype=np.complex128)
    s.SH_to_spat_m(ql, G, im)
    ref = row.real + 0j if im == 0 else row
    assert rel_l2(G, ref) <= TOL, rel_l2(G, ref)
    # analysis: (1/nphi) spat_to_SH_ml of the FFT's row m gives the oracle's analysis at order m
    rng = np.random.default_rng(91)
    w = rng.standard_normal((nphi, nlat))
    F = np.ascontiguousarray(np.fft.fft(w, axis=0)[im])
    Q = np.empty(L + 1 - m, dtype=np.complex128)
    s.spat_to_SH_m(F, Q, im)
    assert rel_l2(Q / nphi, o.analys(w)[sel]) <= TOL
    # round trip (time_SHT.c:929-930)
    Q2 = np.empty_like(Q)
    s.spat_to_SH_m(G, Q2, im)
    assert rel_l2(Q2, ql) <= 1e-13


def test_ml_truncation_and_errors(device):
    s, o = make(63, 63, 1, 96, 192)
    lib = _lib.load()
    im, ltr = 5, 40
    ql = spectral_field(63, 63, seed=92)[s.m == 5].copy()
    G = np.empty(96, dtype=np.complex128)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.SH_to_spat_ml(s._cfg, im, p(ql), p(G), ltr)
    qt = ql.copy()
    qt[ltr - 5 + 1:] = 0
    G2 = np.empty_like(G)
    s.SH_to_spat_m(qt, G2, im)
    assert rel_l2(G, G2) <= 1e-14
    Q = np.full(59, 7.0 + 7.0j)
    lib.spat_to_SH_ml(s._cfg, im, p(G), p(Q), ltr)
    assert np.all(Q[ltr - 5 + 1:] == 0)
    assert rel_l2(Q[:ltr - 5 + 1], qt[:ltr - 5 + 1]) <= 1e-13
    lib.SH_to_spat_ml(s._cfg, 64, p(ql), p(G), 63)     # im > mmax
    assert lib.shtns_last_error() != 0


@pytest.mark.parametrize("case", [(63, 63, 1, 96, 192), (60, 15, 4, 64, 48), (255, 255, 1, 384, 768)])
@pytest.mark.parametrize("which", ["zero", "low", "high"])
def test_ml_vector_against_oracle(case, which, device):
    L, M, mres, nlat, nphi = case
    s, o = make(L, M, mres, nlat, nphi)
    im = {"zero": 0, "low": 1, "high": M}[which]
    m = im * mres
    sel = s.m == m
    sl = spectral_field(L, M, seed=93 + im, mres=mres, power=3.0)
    tl = spectral_field(L, M, seed=94 + im, mres=mres, power=3.0)
    sl[0] = tl[0] = 0
    S, T = (np.ascontiguousarray(a[sel]) for a in (sl, tl))
    sm, tm = np.zeros_like(sl), np.zeros_like(tl)
    sm[sel], tm[sel] = sl[sel], tl[sel]
    vt, vp = o.synth_sphtor(sm, tm)
    rt, rp = (np.fft.fft(a, axis=0)[im] / nphi for a in (vt, vp))
    Gt, Gp = (np.empty(nlat, dtype=np.complex128) for _ in range(2))
    s.SHsphtor_to_spat_m(S, T, Gt, Gp, im)
    if im == 0:
        rt, rp = rt.real + 0j, rp.real + 0j
    assert rel_l2(Gt, rt) <= TOL and rel_l2(Gp, rp) <= TOL, (rel_l2(Gt, rt), rel_l2(Gp, rp))
    rng = np.random.default_rng(95)
    a, b = rng.standard_normal((2, nphi, nlat))
    Ft, Fp = (np.ascontiguousarray(np.fft.fft(x, axis=0)[im]) for x in (a, b))
    S2, T2 = (np.empty(L + 1 - m, dtype=np.complex128) for _ in range(2))
    s.spat_to_SHsphtor_m(Ft, Fp, S2, T2, im)
    s_ref, t_ref = o.analys_sphtor(a, b)
    assert rel_l2(S2 / nphi, s_ref[sel]) <= TOL and rel_l2(T2 / nphi, t_ref[sel]) <= TOL
    S3, T3 = np.empty_like(S2), np.empty_like(T2)
    s.spat_to_SHsphtor_m(Gt, Gp, S3, T3, im)
    assert rel_l2(S3, S) <= 1e-13 and rel_l2(T3, T) <= 1e-13


def test_ml_shorthands_equal_components(device):
    s, o = make(63, 63, 1, 96, 192)
    lib = _lib.load()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    im = 3
    n = 64 - im
    rng = np.random.default_rng(96)
    Q, S, T = (rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3))
    z = np.zeros(n, dtype=np.complex128)
    out = lambda: np.empty(96, dtype=np.complex128)
    a, b, c, d = out(), out(), out(), out()
    lib.SHsph_to_spat_ml(s._cfg, im, p(S), p(a), p(b), 63)
    s.SHsphtor_to_spat_m(S, z, c, d, im)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    lib.SHtor_to_spat_ml(s._cfg, im, p(T), p(a), p(b), 63)
    s.SHsphtor_to_spat_m(z, T, c, d, im)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    vr, vt, vp = out(), out(), out()
    lib.SHqst_to_spat_ml(s._cfg, im, p(Q), p(S), p(T), p(vr), p(vt), p(vp), 63)
    s.SH_to_spat_m(Q, a, im)
    s.SHsphtor_to_spat_m(S, T, c, d, im)
    assert np.array_equal(vr, a) and np.array_equal(vt, c) and np.array_equal(vp, d)
    q2, s2, t2 = (np.empty(n, dtype=np.complex128) for _ in range(3))
    lib.spat_to_SHqst_ml(s._cfg, im, p(vr), p(vt), p(vp), p(q2), p(s2), p(t2), 63)
    assert lib.shtns_last_error() == 0
    assert rel_l2(q2, Q) <= 1e-13 and rel_l2(s2, S) <= 1e-13 and rel_l2(t2, T) <= 1e-13
