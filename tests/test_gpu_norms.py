"""GPU: transforms in the normalizations other than orthonormal (shtns.h:48-55), against the
oracle in the same normalization (oracle/oracle.py NormOracle, pinned to closed forms in
tests/test_norms.py): scalar synthesis and analysis of unfiltered data with 1, 3 and 16 fields
(the 4x4x4 and 16x16x4 kernels), vector transforms, the one-order transforms, the adjoint
identities, complex fields, FP32 data and the sliced contraction."""
import ctypes

import numpy as np
import pytest
import torch

import shtns_amd
from shtns_amd import _lib
from shtns_amd.synthetic import spectral_batch
from oracle.oracle import NormOracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
NORMS = [shtns_amd.sht_schmidt, shtns_amd.sht_fourpi | shtns_amd.SHT_REAL_NORM,
         shtns_amd.sht_schmidt | shtns_amd.SHT_NO_CS_PHASE | shtns_amd.SHT_REAL_NORM]
L, NLAT, NPHI = 47, 72, 128


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def make(norm):
    s = shtns_amd.sht(L, norm=norm)
    s.set_grid(NLAT, NPHI, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s, NormOracle(L, L, 1, NLAT, NPHI, norm=norm)


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("nf", [1, 3, 16])
def test_scalar(norm, nf, device):
    s, o = make(norm)
    q = spectral_batch(L, L, nf, seed=60)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    rng = np.random.default_rng(61)
    w = rng.standard_normal((nf, NPHI, NLAT))
    qa = s.analys(torch.from_numpy(w).to(device)).cpu().numpy()
    for f in {0, nf - 1}:
        assert rel_l2(v[f], o.synth(q[f])) <= TOL
        assert rel_l2(qa[f], o.analys(w[f])) <= TOL
    q2 = s.analys(torch.from_numpy(v).to(device)).cpu().numpy()
    assert rel_l2(q2, q) <= TOL


@pytest.mark.parametrize("norm", NORMS)
def test_vector(norm, device):
    s, o = make(norm)
    S = spectral_batch(L, L, 2, seed=62, power=3.0)
    T = spectral_batch(L, L, 2, seed=64, power=3.0)
    S[:, 0] = T[:, 0] = 0
    vt, vp = (x.cpu().numpy() for x in s.synth(torch.from_numpy(S).to(device), torch.from_numpy(T).to(device)))
    for f in range(2):
        rt, rp = o.synth_sphtor(S[f], T[f])
        assert rel_l2(vt[f], rt) <= TOL and rel_l2(vp[f], rp) <= TOL
    s2, t2 = (x.cpu().numpy() for x in s.analys(torch.from_numpy(vt).to(device), torch.from_numpy(vp).to(device)))
    assert rel_l2(s2, S) <= TOL and rel_l2(t2, T) <= TOL


@pytest.mark.parametrize("norm", NORMS)
def test_one_order(norm, device):
    """SH_to_spat_m / spat_to_SH_m in a normalization equal the orthonormal ones on Q nrm."""
    s, o = make(norm)
    r = shtns_amd.sht(L)
    r.set_grid(NLAT, NPHI, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    q = spectral_batch(L, L, 1, seed=66)[0]
    for im in (0, 3):
        ql = q[s.idx(im, im):s.idx(im, im) + L + 1 - im].copy()
        c = o.c[s.idx(im, im):s.idx(im, im) + L + 1 - im]
        v = np.zeros(NLAT, dtype=complex)
        vr = np.zeros(NLAT, dtype=complex)
        s.SH_to_spat_m(ql, v, im)
        r.SH_to_spat_m(ql * c, vr, im)
        assert rel_l2(v, vr) <= TOL
        qb = np.zeros(L + 1 - im, dtype=complex)
        s.spat_to_SH_m(v, qb, im)
        assert rel_l2(qb, ql) <= 1e-11


@pytest.mark.parametrize("norm", NORMS)
def test_adjoints_cplx_f32_sliced(norm, device):
    s, o = make(norm)
    rng = np.random.default_rng(67)
    a = torch.from_numpy(spectral_batch(L, L, 1, seed=68)[0]).to(device)
    b = torch.from_numpy(spectral_batch(L, L, 1, seed=69)[0]).to(device)
    g = torch.from_numpy(rng.standard_normal((NPHI, NLAT))).to(device)
    lhs = float(torch.sum(s.synth(a) * g))
    rhs = float(torch.sum((torch.conj(a) * s.synth_adjoint(g)).real))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    lhs = float(torch.sum((torch.conj(s.analys(g)) * b).real))
    rhs = float(torch.sum(g * s.analys_adjoint(b)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # complex fields: round trip
    z = rng.standard_normal((NPHI, NLAT)) + 1j * rng.standard_normal((NPHI, NLAT))
    alm = s.analys_cplx(torch.from_numpy(z).to(device))
    z2 = s.synth_cplx(s.analys_cplx(s.synth_cplx(alm))).cpu().numpy()
    assert rel_l2(z2, s.synth_cplx(alm).cpu().numpy()) <= 1e-12
    # FP32 data and the sliced contraction against the FP64 transform
    q = torch.from_numpy(spectral_batch(L, L, 4, seed=70)).to(device)
    v64 = s.synth(q).cpu().numpy()
    v32 = s.synth(q.to(torch.complex64)).cpu().numpy()
    assert rel_l2(v32, v64) <= 2e-6
    s.set_precision(shtns_amd.SHT_PREC_SLICED)
    try:
        v8 = s.synth(q).cpu().numpy()
    finally:
        s.set_precision(shtns_amd.SHT_PREC_AUTO)
    assert rel_l2(v8, v64) <= 1.5e-10
    assert rel_l2(v64[2], o.synth(q[2].cpu().numpy())) <= TOL


def test_noalloc_refuses_other_norms(device):
    s, _ = make(shtns_amd.sht_schmidt)
    lib = _lib.load()
    q = torch.zeros((1, s.nlm), dtype=torch.complex128, device=device)
    v = torch.zeros((1,) + tuple(s.spat_shape), dtype=torch.float64, device=device)
    lib.cu_SH_to_spat_batch_noalloc(s._cfg, ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(v.data_ptr()), s.lmax, 1,
                                    s.nlm, s.nspat)
    assert lib.shtns_last_error() == 1
    lib.shtns_last_error()
