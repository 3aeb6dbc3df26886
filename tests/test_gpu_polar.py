"""GPU: the polar optimization threshold eps of shtns_set_grid (SHTns's PolarOptimize,
sht_init.c:522-557, applied as in kernel_SH_to_spat.gen.c:235-257 and
kernel_spat_to_SH.gen.c:198-201): order m skips the latitudes nearer a pole than tm[m], the
first latitude where max_l |Y_l^m| >= eps.  The HIP path is compared with the oracle run with
the same eps (whose skip rule is pinned against its definition in tests/test_oracle.py), and
with its own eps = 0 result, from which it may differ only by the skipped values (all below
eps): rel-L2 <= 1e-11 for synthesis and <= 1e-10 for the analysis of unfiltered data at
eps = 1e-10.  The Python default polar_opt is 1e-10, as SHTns's."""
import numpy as np
import pytest
import torch

import shtns_amd
from shtns_amd.synthetic import spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def make(lmax, nlat, nphi, eps=None, mmax=None):
    s = shtns_amd.sht(lmax, lmax if mmax is None else mmax)
    flags = shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss
    if eps is None:
        s.set_grid(nlat, nphi, flags=flags)   # the default threshold
    else:
        s.set_grid(nlat, nphi, flags=flags, polar_opt=eps)
    return s


@pytest.mark.parametrize("case", [(127, 128, 256, None), (63, 96, 128, 1e-4), (120, 128, 250, 1e-6)])
def test_scalar_polar_matches_oracle(case, device):
    L, nlat, nphi, eps = case
    s = make(L, nlat, nphi, eps)
    o = Oracle(L, L, 1, nlat, nphi)
    o.set_polar_opt(1e-10 if eps is None else eps)
    assert o.polar_tm().max() > 0
    q = spectral_field(L, L, seed=21)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    assert rel_l2(v, o.synth(q)) <= TOL
    w = np.random.default_rng(22).standard_normal((nphi, nlat))   # not band-limited
    qa = s.analys(torch.from_numpy(w).to(device)).cpu().numpy()
    assert rel_l2(qa, o.analys(w)) <= TOL


def test_polar_default_close_to_full_transform(device):
    L, nlat, nphi = 127, 128, 256
    s, s0 = make(L, nlat, nphi), make(L, nlat, nphi, 0.0)
    q = torch.from_numpy(spectral_field(L, L, seed=23)).to(device)
    w = torch.from_numpy(np.random.default_rng(24).standard_normal((nphi, nlat))).to(device)
    v, v0 = s.synth(q).cpu().numpy(), s0.synth(q).cpu().numpy()
    qa, qa0 = s.analys(w).cpu().numpy(), s0.analys(w).cpu().numpy()
    es, ea = rel_l2(v, v0), rel_l2(qa, qa0)
    print(f"eps=1e-10 vs eps=0: synthesis {es:.2e}, analysis {ea:.2e}")
    assert 0.0 < es <= 1e-11 and ea <= 1e-10


def test_vector_polar_matches_oracle(device):
    L, nlat, nphi, eps = 63, 96, 128, 1e-4
    s = make(L, nlat, nphi, eps)
    o = Oracle(L, L, 1, nlat, nphi)
    o.set_polar_opt(eps)
    sl = spectral_field(L, L, seed=25, power=3.0)
    tl = spectral_field(L, L, seed=26, power=3.0)
    sl[0] = tl[0] = 0
    vt, vp = s.synth(torch.from_numpy(sl).to(device), torch.from_numpy(tl).to(device))
    vt_ref, vp_ref = o.synth_sphtor(sl, tl)
    assert rel_l2(vt.cpu().numpy(), vt_ref) <= TOL and rel_l2(vp.cpu().numpy(), vp_ref) <= TOL
    rng = np.random.default_rng(27)
    wt, wp = rng.standard_normal((nphi, nlat)), rng.standard_normal((nphi, nlat))
    s2, t2 = s.analys(torch.from_numpy(wt).to(device), torch.from_numpy(wp).to(device))
    s_ref, t_ref = o.analys_sphtor(wt, wp)
    assert rel_l2(s2.cpu().numpy(), s_ref) <= TOL and rel_l2(t2.cpu().numpy(), t_ref) <= TOL


def test_polar_fixed_order_matches_full_transform(device):
    """SH_to_spat_m / spat_to_SH_m skip the same polar rows (fly_SH_to_spat.gen.c:335-345):
    SH_to_spat_m gives Fourier row m of the full synthesis of the single-order field (bin m
    of the FFT over phi, / nphi) and spat_to_SH_m of the FFT's row m, / nphi, the analysis."""
    L, nlat, nphi, eps = 63, 96, 128, 1e-4
    s = make(L, nlat, nphi, eps)
    o = Oracle(L, L, 1, nlat, nphi)
    o.set_polar_opt(eps)
    tm = o.polar_tm()
    q = spectral_field(L, L, seed=28)
    w = np.random.default_rng(30).standard_normal((nphi, nlat))
    qa = o.analys(w)
    for m in (0, 20, 50):
        sel = s.m == m
        ql = np.ascontiguousarray(q[sel])
        only = np.zeros_like(q)
        only[sel] = ql
        row = np.fft.fft(o.synth(only), axis=0)[m] / nphi
        ref = row.real + 0j if m == 0 else row
        G = np.empty(nlat, dtype=np.complex128)
        s.SH_to_spat_m(ql, G, m)
        assert rel_l2(G, ref) <= TOL
        if tm[m] > 0:
            assert np.abs(G[:tm[m]]).max() == 0.0 and np.abs(G[nlat - tm[m]:]).max() == 0.0
        F = np.ascontiguousarray(np.fft.fft(w, axis=0)[m])
        Q = np.empty(L + 1 - m, dtype=np.complex128)
        s.spat_to_SH_m(F, Q, m)
        assert rel_l2(Q / nphi, qa[sel]) <= TOL


def test_polar_lmax1023_one_field(device):
    L, nlat, nphi = 1023, 1536, 3072
    s = make(L, nlat, nphi)                 # default eps = 1e-10
    o = Oracle(L, L, 1, nlat, nphi)
    o.set_polar_opt(1e-10)
    q = spectral_field(L, L, seed=29)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    v_ref = o.synth(q)
    assert rel_l2(v, v_ref) <= TOL
    qa = s.analys(torch.from_numpy(v_ref).to(device)).cpu().numpy()
    assert rel_l2(qa, o.analys(v_ref)) <= TOL
