"""GPU: complex-field transforms (shtns.h:272, 277, 292-294; sht_func.c:943-1195) at sizes
where the cooperative FFT kernels read and write the interleaved complex arrays directly:
one launch covers both parts of every field, the part taken from bit 3 of the block index,
which needs the latitude groups to come in whole eights.

lmax=170 on Gauss 256 x 512 (sixteen groups of 16 latitudes with 256 threads, eight of 32
with 512) and 248 x 512 (the last group ragged).  FFT kinds forced through cushtns_set_tuning:
5 and 6 (8-byte staged, 256 / 512 threads) take the direct interleaved path in both
directions, 3 (register kernel, 256 threads) the staging batch and interleave pass of the
synthesis and two strided launches of the analysis.  Each kind against the oracle's complex
transforms (rel-L2 <= 1e-12; round trips <= 1e-13); the kinds against each other; the
C ABI's padded field distances; a real field against the real transform."""
import ctypes

import numpy as np
import pytest
import torch

import shtns_amd
from shtns_amd import _lib
from shtns_amd.batch import set_tuning
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
L, NPHI = 170, 512


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def random_alm(n, seed, o):
    """complex coefficients with |a_lm| ~ 1/(l+1)"""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            a[o.lm_cplx(l, m)] /= (l + 1)
    return a


@pytest.fixture(scope="module", params=[256, 248])
def grid(request):
    nlat = request.param
    s = shtns_amd.sht(L, L)
    s.set_grid(nlat, NPHI, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    yield s, Oracle(L, L, 1, nlat, NPHI), nlat
    set_tuning(s, [0, 0, 0, 0, 0])


@pytest.mark.parametrize("kind", [5, 6, 3])
def test_scalar_cplx_kinds(grid, device, kind):
    s, o, nlat = grid
    set_tuning(s, [0, 0, 0, kind, kind])
    a = random_alm(s.nlm_cplx, 200 + nlat, o)
    z = s.synth_cplx(a)
    assert z.shape == (NPHI, nlat)
    ez = rel_l2(z, o.synth_cplx(a))
    rng = np.random.default_rng(201 + nlat)
    w = rng.standard_normal((NPHI, nlat)) + 1j * rng.standard_normal((NPHI, nlat))
    q = s.analys_cplx(w)
    ew = rel_l2(q, o.analys_cplx(w))
    er = rel_l2(s.analys_cplx(z), a)
    print(f"\nnlat {nlat} kind {kind}: synth {ez:.2e} anal {ew:.2e} round trip {er:.2e}")
    assert ez <= TOL and ew <= TOL and er <= 1e-13
    # the direct and the staged path against each other
    set_tuning(s, [0, 0, 0, 3, 3])
    assert rel_l2(z, s.synth_cplx(a)) <= 1e-14 and rel_l2(q, s.analys_cplx(w)) <= 1e-14


@pytest.mark.parametrize("kind", [5, 3])
def test_vector_cplx_kinds(grid, device, kind):
    s, o, nlat = grid
    set_tuning(s, [0, 0, 0, kind, kind])
    S, T = random_alm(s.nlm_cplx, 210 + nlat, o), random_alm(s.nlm_cplx, 211 + nlat, o)
    S[0] = T[0] = 0
    zt, zp = s.synth_cplx(S, T)
    rt, rp = o.synth_sphtor_cplx(S, T)
    assert rel_l2(zt, rt) <= TOL and rel_l2(zp, rp) <= TOL
    rng = np.random.default_rng(212 + nlat)
    wt, wp = rng.standard_normal((2, NPHI, nlat)) + 1j * rng.standard_normal((2, NPHI, nlat))
    s2, t2 = s.analys_cplx(wt, wp)
    rs, rt2 = o.analys_sphtor_cplx(wt, wp)
    assert rel_l2(s2, rs) <= TOL and rel_l2(t2, rt2) <= TOL
    s3, t3 = s.analys_cplx(zt, zp)
    assert rel_l2(s3, S) <= 1e-13 and rel_l2(t3, T) <= 1e-13


def test_real_field_equals_real_transform(grid, device):
    """A real field analysed through the interleaved path: m >= 0 coefficients as the real
    analysis (the real part goes through the same kernels, at twice the stride)."""
    s, o, nlat = grid
    set_tuning(s, [0, 0, 0, 5, 5])
    rng = np.random.default_rng(220 + nlat)
    v = rng.standard_normal((NPHI, nlat))
    a = s.analys_cplx(v.astype(np.complex128))
    q = s.analys(v)
    scale = np.linalg.norm(q)
    for m in range(L + 1):
        for l in range(m, L + 1):
            assert abs(a[s.zidx(l, m)] - q[s.idx(l, m)]) <= 1e-15 * scale, (l, m)
            assert abs(a[s.zidx(l, -m)] - (-1) ** m * np.conj(q[s.idx(l, m)])) <= 1e-15 * scale, (l, -m)


def test_batch_padded_distances(grid, device):
    """Batches of 3 fields through the C ABI with padded (odd) field distances on the direct
    path: each field as the single-field call, padding untouched, in both directions."""
    s, o, nlat = grid
    set_tuning(s, [0, 0, 0, 5, 5])
    B, nsp, nlm = 3, s.nspat, s.nlm_cplx
    A = np.stack([random_alm(nlm, 230 + f, o) for f in range(B)])
    dA = torch.from_numpy(A).to(device)
    single = [s.synth_cplx(dA[f]) for f in range(B)]
    lib = _lib.load()
    st = torch.cuda.current_stream()
    lib.cushtns_set_streams(s._cfg, ctypes.c_void_p(st.cuda_stream), ctypes.c_void_p(st.cuda_stream))
    sd, vd = nlm + 3, nsp + 5
    ap = torch.zeros((B, sd), dtype=torch.complex128, device=device)
    ap[:, :nlm] = dA
    zpd = torch.full((B, vd), 7.0 + 7.0j, dtype=torch.complex128, device=device)
    lib.cu_SH_to_spat_cplx_batch(s._cfg, ctypes.c_void_p(ap.data_ptr()), ctypes.c_void_p(zpd.data_ptr()), B, sd, vd)
    torch.cuda.synchronize()
    assert lib.shtns_last_error() == 0
    for f in range(B):
        assert torch.equal(zpd[f, :nsp].reshape(NPHI, nlat), single[f])
    assert torch.all(zpd[:, nsp:] == 7.0 + 7.0j)
    back = torch.full((B, sd), 5.0 + 5.0j, dtype=torch.complex128, device=device)
    lib.cu_spat_cplx_to_SH_batch(s._cfg, ctypes.c_void_p(zpd.data_ptr()), ctypes.c_void_p(back.data_ptr()), B, vd, sd)
    torch.cuda.synchronize()
    assert lib.shtns_last_error() == 0
    for f in range(B):
        assert torch.equal(back[f, :nlm], s.analys_cplx(single[f]))
        assert rel_l2(back[f, :nlm].cpu().numpy(), A[f]) <= 1e-13
    assert torch.all(back[:, nlm:] == 5.0 + 5.0j)
