"""GPU: grids beyond the even Gauss grid -- the regular grid with both poles (sht_reg_poles,
Clenshaw-Curtis; sht_legendre.c:868-935, sht_init.c:580-600), odd nlat (an equator row that is
its own mirror, weight halved in the pair sums as sht_init.c:598; SHTns's separate odd-nlat
code, sht_odd_nlat.c, computes the same transforms), and the regular grid without poles at
odd nlat.  Scalar and vector transforms against the oracle on the same grid (pinned against
scipy sums and pole limits in tests/test_oracle.py), round trips, the one-order transforms
and the adjoint identity on an odd grid."""
import numpy as np
import pytest
import torch

import shtns_amd
from shtns_amd.synthetic import spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
GRIDS = {"poles": (shtns_amd.sht_reg_poles, Oracle.REGULAR_POLES), "gauss": (shtns_amd.sht_gauss, Oracle.GAUSS),
         "fejer": (shtns_amd.sht_reg_fast, Oracle.REGULAR)}


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def make(L, M, mres, nlat, nphi, grid, eps=0.0):
    g, og = GRIDS[grid]
    s = shtns_amd.sht(L, M, mres)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | g, polar_opt=eps)
    o = Oracle(L, M, mres, nlat, nphi, grid=og)
    if eps > 0:
        o.set_polar_opt(eps)
    return s, o


@pytest.mark.parametrize("case", [(63, 63, 1, 129, 128, "poles"), (63, 63, 1, 128, 128, "poles"),
                                  (100, 100, 1, 151, 216, "gauss"), (40, 20, 2, 61, 84, "gauss"),
                                  (60, 60, 1, 121, 128, "fejer"), (127, 127, 1, 257, 256, "poles")])
def test_scalar_against_oracle(case, device):
    L, M, mres, nlat, nphi, grid = case
    s, o = make(L, M, mres, nlat, nphi, grid)
    q = spectral_field(L, M, seed=80, mres=mres)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    assert rel_l2(v, o.synth(q)) <= TOL
    w = np.random.default_rng(81).standard_normal((nphi, nlat))
    qa = s.analys(torch.from_numpy(w).to(device)).cpu().numpy()
    assert rel_l2(qa, o.analys(w)) <= TOL
    q2 = s.analys(torch.from_numpy(v).to(device)).cpu().numpy()
    assert rel_l2(q2, q) <= TOL


@pytest.mark.parametrize("case", [(31, 63, 64, "poles"), (31, 64, 64, "poles"), (40, 61, 84, "gauss")])
def test_vector_against_oracle(case, device):
    L, nlat, nphi, grid = case
    s, o = make(L, L, 1, nlat, nphi, grid)
    sl = spectral_field(L, L, seed=82, power=3.0)
    tl = spectral_field(L, L, seed=83, power=3.0)
    sl[0] = tl[0] = 0
    vt, vp = s.synth(torch.from_numpy(sl).to(device), torch.from_numpy(tl).to(device))
    rt, rp = o.synth_sphtor(sl, tl)
    assert rel_l2(vt.cpu().numpy(), rt) <= TOL and rel_l2(vp.cpu().numpy(), rp) <= TOL
    rng = np.random.default_rng(84)
    wt, wp = rng.standard_normal((nphi, nlat)), rng.standard_normal((nphi, nlat))
    s2, t2 = s.analys(torch.from_numpy(wt).to(device), torch.from_numpy(wp).to(device))
    sa, ta = o.analys_sphtor(wt, wp)
    assert rel_l2(s2.cpu().numpy(), sa) <= TOL and rel_l2(t2.cpu().numpy(), ta) <= TOL
    s3, t3 = s.analys(vt, vp)
    assert rel_l2(s3.cpu().numpy(), sl) <= TOL and rel_l2(t3.cpu().numpy(), tl) <= TOL


def test_polar_threshold_on_reg_poles(device):
    s, o = make(63, 63, 1, 129, 128, "poles", eps=1e-6)
    assert o.polar_tm().max() > 0
    q = spectral_field(63, 63, seed=85)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    assert rel_l2(v, o.synth(q)) <= TOL


def test_one_order_on_odd_grid(device):
    L, nlat, nphi = 50, 75, 108
    s, o = make(L, L, 1, nlat, nphi, "gauss")
    q = spectral_field(L, L, seed=86)
    w = np.random.default_rng(87).standard_normal((nphi, nlat))
    qa = o.analys(w)
    for m in (0, 7, 50):
        sel = s.m == m
        only = np.zeros_like(q)
        only[sel] = q[sel]
        row = np.fft.fft(o.synth(only), axis=0)[m] / nphi
        G = np.empty(nlat, dtype=np.complex128)
        s.SH_to_spat_m(np.ascontiguousarray(q[sel]), G, m)
        assert rel_l2(G, row.real + 0j if m == 0 else row) <= TOL
        F = np.ascontiguousarray(np.fft.fft(w, axis=0)[m])
        Q = np.empty(L + 1 - m, dtype=np.complex128)
        s.spat_to_SH_m(F, Q, m)
        assert rel_l2(Q / nphi, qa[sel]) <= TOL


def test_adjoint_identity_on_odd_grid(device):
    """<S a, g> = <a, S^T g> and <A v, b> = <v, A^T b> with the real inner products of the
    adjoint entry points, on an odd Gauss grid (the equator row carries its full weight)."""
    L, nlat, nphi = 40, 61, 84
    s, _ = make(L, L, 1, nlat, nphi, "gauss")
    rng = np.random.default_rng(88)
    a = torch.from_numpy(spectral_field(L, L, seed=89)).to(device)
    b = torch.from_numpy(spectral_field(L, L, seed=90)).to(device)
    g = torch.from_numpy(rng.standard_normal((nphi, nlat))).to(device)
    lhs = float(torch.sum(s.synth(a) * g))
    rhs = float(torch.sum((torch.conj(a) * s.synth_adjoint(g)).real))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    lhs = float(torch.sum((torch.conj(s.analys(g)) * b).real))
    rhs = float(torch.sum(g * s.analys_adjoint(b)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
