"""GPU: 3-component transforms (radial scalar Q + tangential spheroidal/toroidal S, T):
SHqst_to_spat / spat_to_SHqst (shtns.h:303, 306, 323-324), cu_SHqst_to_spat /
cu_spat_to_SHqst (shtns_cuda.h:50-52) and the 3-array forms of sht.synth / sht.analys
(shtns_numpy.i:396-457), against the oracle's scalar and vector transforms."""
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


def make(lmax, mmax, nlat, nphi):
    s = shtns_amd.sht(lmax, mmax)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s, Oracle(lmax, mmax, 1, nlat, nphi)


def fields(lmax, mmax, seed):
    q = spectral_field(lmax, mmax, seed=seed)
    sl = spectral_field(lmax, mmax, seed=seed + 1, power=3.0)
    tl = spectral_field(lmax, mmax, seed=seed + 2, power=3.0)
    sl[0] = tl[0] = 0   # degree 0 has no tangential part
    return q, sl, tl


@pytest.mark.parametrize("case", [(63, 63, 96, 192), (120, 50, 128, 120)])
def test_qst_numpy_matches_oracle(case, device):
    L, mm, nlat, nphi = case
    s, o = make(L, mm, nlat, nphi)
    q, sl, tl = fields(L, mm, 30)
    vr, vt, vp = s.synth(q, sl, tl)
    vt_ref, vp_ref = o.synth_sphtor(sl, tl)
    assert rel_l2(vr, o.synth(q)) <= TOL
    assert rel_l2(vt, vt_ref) <= TOL and rel_l2(vp, vp_ref) <= TOL
    q2, s2, t2 = s.analys(vr, vt, vp)
    assert rel_l2(q2, o.analys(vr)) <= TOL
    s_ref, t_ref = o.analys_sphtor(vt, vp)
    assert rel_l2(s2, s_ref) <= TOL and rel_l2(t2, t_ref) <= TOL
    # band-limited round trip
    assert rel_l2(q2, q) <= TOL and rel_l2(s2, sl) <= TOL and rel_l2(t2, tl) <= TOL


def test_qst_truncated(device):
    s, o = make(63, 63, 96, 192)
    q, sl, tl = fields(63, 63, 40)
    ltr = 40
    vr, vt, vp = s.synth(q, sl, tl, ltr=ltr)
    cut = s.l > ltr
    qc, sc, tc = q.copy(), sl.copy(), tl.copy()
    qc[cut] = sc[cut] = tc[cut] = 0
    vt_ref, vp_ref = o.synth_sphtor(sc, tc)
    assert rel_l2(vr, o.synth(qc)) <= TOL
    assert rel_l2(vt, vt_ref) <= TOL and rel_l2(vp, vp_ref) <= TOL
    q2, s2, t2 = s.analys(vr, vt, vp, ltr=ltr)
    for a in (q2, s2, t2):
        assert np.all(a[cut] == 0)
    assert rel_l2(q2, qc) <= TOL and rel_l2(s2, sc) <= TOL and rel_l2(t2, tc) <= TOL


def test_qst_device_entry_points_equal_components(device):
    """cu_SHqst_to_spat / cu_spat_to_SHqst through the C ABI give the scalar and vector
    device transforms' results bit for bit, and the torch 3-array forms agree."""
    import torch
    s, o = make(127, 127, 128, 256)
    q, sl, tl = fields(127, 127, 50)
    lib = _lib.load()
    torch.cuda.set_device(device)
    st = torch.cuda.current_stream()
    lib.cushtns_set_streams(s._cfg, ctypes.c_void_p(st.cuda_stream), ctypes.c_void_p(st.cuda_stream))
    dq, ds, dt = (torch.from_numpy(a).to(device) for a in (q, sl, tl))
    vr, vt, vp = (torch.empty(tuple(s.spat_shape), dtype=torch.float64, device=device) for _ in range(3))
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    lib.cu_SHqst_to_spat(s._cfg, p(dq), p(ds), p(dt), p(vr), p(vt), p(vp), 127)
    ur, ut, up = (torch.empty_like(vr) for _ in range(3))
    lib.cu_SH_to_spat(s._cfg, p(dq), p(ur), 127)
    lib.cu_SHsphtor_to_spat(s._cfg, p(ds), p(dt), p(ut), p(up), 127)
    torch.cuda.synchronize()
    assert _lib.load().shtns_last_error() == 0
    assert torch.equal(vr, ur) and torch.equal(vt, ut) and torch.equal(vp, up)
    vt_ref, vp_ref = o.synth_sphtor(sl, tl)
    assert rel_l2(vr.cpu().numpy(), o.synth(q)) <= TOL
    assert rel_l2(vt.cpu().numpy(), vt_ref) <= TOL and rel_l2(vp.cpu().numpy(), vp_ref) <= TOL
    q2, s2, t2 = (torch.empty_like(dq) for _ in range(3))
    lib.cu_spat_to_SHqst(s._cfg, p(vr), p(vt), p(vp), p(q2), p(s2), p(t2), 127)
    torch.cuda.synchronize()
    assert rel_l2(q2.cpu().numpy(), q) <= TOL
    assert rel_l2(s2.cpu().numpy(), sl) <= TOL and rel_l2(t2.cpu().numpy(), tl) <= TOL
    # torch 3-array forms
    wr, wt, wp = s.synth(dq, ds, dt)
    assert torch.equal(wr, vr) and torch.equal(wt, vt) and torch.equal(wp, vp)
    a, b, c = s.analys(wr, wt, wp)
    assert torch.equal(a, q2) and torch.equal(b, s2) and torch.equal(c, t2)


def test_qst_rejects_four_arrays(device):
    s, _ = make(31, 31, 32, 64)
    q = s.spec_array()
    with pytest.raises(RuntimeError):
        s.synth(q, q, q, q)
