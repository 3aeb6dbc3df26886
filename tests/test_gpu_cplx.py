"""GPU: complex-valued transforms SH_to_spat_cplx / spat_cplx_to_SH (shtns.h:272, 277;
sht_func.c:943-1052), SHsphtor_to_spat_cplx / spat_cplx_to_SHsphtor (shtns.h:292-294;
sht_func.c:1060-1195), the 3-component forms (shtns.h:308-309), their device forms and the
Python synth_cplx / analys_cplx (shtns_numpy.i:469-535), against the oracle's complex
transforms (pinned against scipy's complex spherical harmonics on the CPU)."""
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


def make(lmax, mmax, nlat, nphi, layout=shtns_amd.SHT_THETA_CONTIGUOUS):
    s = shtns_amd.sht(lmax, mmax)
    s.set_grid(nlat, nphi, flags=layout | shtns_amd.sht_gauss, polar_opt=0.0)
    return s, Oracle(lmax, mmax, 1, nlat, nphi)


def random_alm(n, seed, lmax, o):
    """complex coefficients with the spectra's usual decay: |a_lm| ~ 1/(l+1)"""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for l in range(lmax + 1):
        for m in range(-min(l, o.mmax), min(l, o.mmax) + 1):
            a[o.lm_cplx(l, m)] /= (l + 1)
    return a


@pytest.mark.parametrize("case", [(31, 32, 64), (63, 96, 192), (127, 128, 256)])
def test_cplx_matches_oracle(case, device):
    L, nlat, nphi = case
    s, o = make(L, L, nlat, nphi)
    assert s.nlm_cplx == (L + 1) ** 2 == o.nlm_cplx
    a = random_alm(s.nlm_cplx, 70, L, o)
    z = s.synth_cplx(a)
    assert z.shape == (nphi, nlat) and z.dtype == np.complex128
    assert rel_l2(z, o.synth_cplx(a)) <= TOL
    # analysis of a field that is not band-limited
    rng = np.random.default_rng(71)
    w = rng.standard_normal((nphi, nlat)) + 1j * rng.standard_normal((nphi, nlat))
    assert rel_l2(s.analys_cplx(w), o.analys_cplx(w)) <= TOL
    # band-limited round trip
    assert rel_l2(s.analys_cplx(z), a) <= 1e-13


def test_cplx_phi_contiguous(device):
    s, o = make(47, 47, 48, 96, layout=shtns_amd.SHT_PHI_CONTIGUOUS)
    assert s.spat_shape == (48, 96)
    a = random_alm(s.nlm_cplx, 72, 47, o)
    z = s.synth_cplx(a)
    assert rel_l2(z, o.synth_cplx(a).T) <= TOL
    assert rel_l2(s.analys_cplx(z), a) <= 1e-13


def test_cplx_mmax_below_lmax_c_api(device):
    """The C functions take mmax < lmax (nlm_cplx_calc layout); the Python methods, like
    SHTns's, require lmax = mmax."""
    s, o = make(40, 20, 64, 48)
    assert s.nlm_cplx == o.nlm_cplx
    a = random_alm(s.nlm_cplx, 73, 40, o)
    z = np.empty((48, 64), dtype=np.complex128)
    lib = _lib.load()
    lib.SH_to_spat_cplx(s._cfg, a.ctypes.data_as(ctypes.c_void_p), z.ctypes.data_as(ctypes.c_void_p))
    assert lib.shtns_last_error() == 0
    assert rel_l2(z, o.synth_cplx(a)) <= TOL
    a2 = np.empty_like(a)
    lib.spat_cplx_to_SH(s._cfg, z.ctypes.data_as(ctypes.c_void_p), a2.ctypes.data_as(ctypes.c_void_p))
    assert rel_l2(a2, a) <= 1e-13
    with pytest.raises(RuntimeError):
        s.synth_cplx(a)


def test_cplx_real_field_equals_real_transform(device):
    """For a real field the m >= 0 coefficients are the real transform's, bit for bit (the
    real part is analysed in the same kernels as one field of a batch), and the m < 0 ones
    (-1)^m conj of them."""
    s, o = make(63, 63, 96, 192)
    q = spectral_field(63, 63, seed=74)
    v = s.synth(q)
    a = s.analys_cplx(v.astype(np.complex128))
    q2 = s.analys(v)
    for l in range(64):
        for m in range(l + 1):
            assert a[s.zidx(l, m)] == q2[s.idx(l, m)], (l, m)
            assert abs(a[s.zidx(l, -m)] - (-1) ** m * np.conj(q2[s.idx(l, m)])) <= 1e-15 * (1 + abs(q2[s.idx(l, m)]))
    z = s.synth_cplx(a)
    assert np.max(np.abs(z.imag)) <= 1e-12 * np.max(np.abs(v)) and rel_l2(z.real, v) <= 1e-13


def test_cplx_torch_equals_host(device):
    import torch
    s, o = make(95, 95, 96, 192)
    a = random_alm(s.nlm_cplx, 75, 95, o)
    z = s.synth_cplx(a)
    zt = s.synth_cplx(torch.from_numpy(a).to(device))
    assert zt.dtype == torch.complex128 and tuple(zt.shape) == (192, 96)
    assert np.array_equal(zt.cpu().numpy(), z)
    at = s.analys_cplx(zt)
    assert np.array_equal(at.cpu().numpy(), s.analys_cplx(z))


def test_cplx_rejects(device):
    s, _ = make(31, 31, 32, 64)
    a = np.zeros(s.nlm_cplx, dtype=np.complex128)
    with pytest.raises(RuntimeError):
        s.synth_cplx(a, a, a, a)
    with pytest.raises(RuntimeError):
        s.synth_cplx(np.zeros(s.nlm, dtype=np.complex128))
    with pytest.raises(RuntimeError):
        s.analys_cplx(np.zeros(s.spat_shape))    # float64, not complex


@pytest.mark.parametrize("case", [(31, 32, 64), (63, 96, 192)])
def test_sphtor_cplx_matches_oracle(case, device):
    L, nlat, nphi = case
    s, o = make(L, L, nlat, nphi)
    S, T = random_alm(s.nlm_cplx, 80, L, o), random_alm(s.nlm_cplx, 81, L, o)
    S[0] = T[0] = 0
    zt, zp = s.synth_cplx(S, T)
    rt, rp = o.synth_sphtor_cplx(S, T)
    assert rel_l2(zt, rt) <= TOL and rel_l2(zp, rp) <= TOL
    rng = np.random.default_rng(82)
    wt, wp = rng.standard_normal((2, nphi, nlat)) + 1j * rng.standard_normal((2, nphi, nlat))
    s2, t2 = s.analys_cplx(wt, wp)
    rs, rt2 = o.analys_sphtor_cplx(wt, wp)
    assert rel_l2(s2, rs) <= TOL and rel_l2(t2, rt2) <= TOL
    s3, t3 = s.analys_cplx(zt, zp)
    assert rel_l2(s3, S) <= 1e-13 and rel_l2(t3, T) <= 1e-13


def test_sphtor_cplx_mmax_below_lmax_c_api(device):
    s, o = make(40, 20, 64, 48)
    S, T = random_alm(s.nlm_cplx, 83, 40, o), random_alm(s.nlm_cplx, 84, 40, o)
    S[0] = T[0] = 0
    zt, zp = (np.empty((48, 64), dtype=np.complex128) for _ in range(2))
    lib = _lib.load()
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    lib.SHsphtor_to_spat_cplx(s._cfg, p(S), p(T), p(zt), p(zp))
    assert lib.shtns_last_error() == 0
    rt, rp = o.synth_sphtor_cplx(S, T)
    assert rel_l2(zt, rt) <= TOL and rel_l2(zp, rp) <= TOL
    S2, T2 = np.empty_like(S), np.empty_like(T)
    lib.spat_cplx_to_SHsphtor(s._cfg, p(zt), p(zp), p(S2), p(T2))
    assert rel_l2(S2, S) <= 1e-13 and rel_l2(T2, T) <= 1e-13


def test_sphtor_cplx_real_field(device):
    """A real vector field: m >= 0 coefficients as the real vector analysis, m < 0 ones
    (-1)^m conj."""
    s, o = make(63, 63, 96, 192)
    rng = np.random.default_rng(85)
    vt, vp = rng.standard_normal((2, 192, 96))
    sr, tr = s.analys(vt, vp)
    sc, tc = s.analys_cplx(vt + 0j, vp + 0j)
    scale = np.linalg.norm(sr)
    for l in range(64):
        for m in range(l + 1):
            for c, r in ((sc, sr), (tc, tr)):
                assert abs(c[s.zidx(l, m)] - r[s.idx(l, m)]) <= 1e-14 * scale
                assert abs(c[s.zidx(l, -m)] - (-1) ** m * np.conj(r[s.idx(l, m)])) <= 1e-14 * scale


def test_qst_cplx_torch_equals_host_and_components(device):
    import torch
    s, o = make(63, 63, 64, 128)
    Q, S, T = (random_alm(s.nlm_cplx, 86 + i, 63, o) for i in range(3))
    S[0] = T[0] = 0
    zr, zt, zp = s.synth_cplx(Q, S, T)
    assert np.array_equal(zr, s.synth_cplx(Q))
    ut, up = s.synth_cplx(S, T)
    assert np.array_equal(zt, ut) and np.array_equal(zp, up)
    assert rel_l2(zr, o.synth_cplx(Q)) <= TOL
    d = [torch.from_numpy(x).to(device) for x in (Q, S, T)]
    wr, wt, wp = s.synth_cplx(*d)
    for a, b in ((wr, zr), (wt, zt), (wp, zp)):
        assert np.array_equal(a.cpu().numpy(), b)
    q2, s2, t2 = s.analys_cplx(wr, wt, wp)
    h = s.analys_cplx(zr, zt, zp)
    for a, b in zip((q2, s2, t2), h):
        assert np.array_equal(a.cpu().numpy(), b)
    for a, b in zip(h, (Q, S, T)):
        assert rel_l2(a, b) <= 1e-13


def test_cplx_batch_equals_single(device):
    """Batched device calls (leading dimensions of torch tensors, cu_*_cplx_batch) give each
    field's single-call result bit for bit; padded distances through the C ABI."""
    import torch
    s, o = make(47, 47, 48, 96)
    B = 5
    A = np.stack([random_alm(s.nlm_cplx, 100 + f, 47, o) for f in range(B)]).reshape(B, s.nlm_cplx)
    S = np.stack([random_alm(s.nlm_cplx, 110 + f, 47, o) for f in range(B)])
    T = np.stack([random_alm(s.nlm_cplx, 120 + f, 47, o) for f in range(B)])
    S[:, 0] = T[:, 0] = 0
    dA, dS, dT = (torch.from_numpy(x).to(device) for x in (A, S, T))
    z = s.synth_cplx(dA.reshape(B, 1, -1))
    assert tuple(z.shape) == (B, 1, 96, 48)
    zt, zp = s.synth_cplx(dS, dT)
    for f in range(B):
        assert torch.equal(z[f, 0], s.synth_cplx(dA[f]))
        a1, b1 = s.synth_cplx(dS[f], dT[f])
        assert torch.equal(zt[f], a1) and torch.equal(zp[f], b1)
    a2 = s.analys_cplx(z)
    s2, t2 = s.analys_cplx(zt, zp)
    for f in range(B):
        assert torch.equal(a2[f, 0], s.analys_cplx(z[f, 0]))
        c1, d1 = s.analys_cplx(zt[f], zp[f])
        assert torch.equal(s2[f], c1) and torch.equal(t2[f], d1)
        assert rel_l2(a2[f, 0].cpu().numpy(), A[f]) <= 1e-13
    assert rel_l2(zt[2].cpu().numpy(), o.synth_sphtor_cplx(S[2], T[2])[0]) <= TOL
    # padded field distances through the C ABI
    lib = _lib.load()
    sd, vd = s.nlm_cplx + 7, s.nspat + 5
    ap = torch.zeros((B, sd), dtype=torch.complex128, device=device)
    ap[:, :s.nlm_cplx] = dA
    zpd = torch.zeros((B, vd), dtype=torch.complex128, device=device)
    st = torch.cuda.current_stream()
    lib.cushtns_set_streams(s._cfg, ctypes.c_void_p(st.cuda_stream), ctypes.c_void_p(st.cuda_stream))
    lib.cu_SH_to_spat_cplx_batch(s._cfg, ctypes.c_void_p(ap.data_ptr()), ctypes.c_void_p(zpd.data_ptr()), B, sd, vd)
    torch.cuda.synchronize()
    assert lib.shtns_last_error() == 0
    assert torch.equal(zpd[:, :s.nspat].reshape(B, 96, 48), z[:, 0])
    assert torch.all(zpd[:, s.nspat:] == 0)
    lib.cu_SH_to_spat_cplx_batch(s._cfg, ctypes.c_void_p(ap.data_ptr()), ctypes.c_void_p(zpd.data_ptr()), B,
                                 s.nlm_cplx - 1, vd)
    assert lib.shtns_last_error() != 0
