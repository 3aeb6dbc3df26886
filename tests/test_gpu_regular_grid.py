"""GPU: transforms on the regular grid without poles (equispaced theta, Fejer quadrature).

Synthesis is compared with a direct sum over the oracle's Legendre functions at the grid's
nodes; that direct sum is first checked against the oracle's own synthesis on the Gauss
grid, which pins its conventions.  Analysis is checked by band-limited round trips: the
Fejer rule on n nodes is exact for degree < n, so with nlat > 2 lmax + 2 (the vector
transforms' integrands reach degree 2 (lmax + 1)) analysis(synthesis(q)) must return q."""
import numpy as np
import pytest

import shtns_amd
from shtns_amd.synthetic import spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def direct_synth(o, q, x, nphi):
    """V[phi_j][theta_k] = sum_m F_m(x_k) e^{i m phi_j} (m = 0 once, m > 0 twice, real part),
    F_m(x) = sum_l q_lm P_lm(x) from the oracle's Legendre functions."""
    nlat = len(x)
    F = np.zeros((o.mmax + 1, nlat), dtype=np.complex128)
    for im in range(o.mmax + 1):
        m = im * o.mres
        ql = np.array([q[o.idx(l, m)] for l in range(m, o.lmax + 1)])
        for k in range(nlat):
            F[im, k] = ql @ o.legendre(im, float(x[k]))
    phi = 2 * np.pi * np.arange(nphi) / nphi
    v = np.outer(np.ones(nphi), F[0].real)
    for im in range(1, o.mmax + 1):
        v += 2 * np.real(np.outer(np.exp(1j * im * o.mres * phi), F[im]))
    return v


def regular(lmax, nlat, nphi):
    s = shtns_amd.sht(lmax)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_reg_fast, polar_opt=0.0)
    return s


def test_direct_sum_matches_oracle_on_gauss_grid():
    o = Oracle(15, 15, 1, 32, 64)
    q = spectral_field(15, 15, seed=60)
    s = shtns_amd.sht(15)
    s.set_grid(32, 64, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    assert rel_l2(direct_synth(o, q, np.asarray(s.cos_theta), 64), o.synth(q)) <= 1e-13


@pytest.mark.parametrize("case", [(31, 64, 64), (31, 48, 64)])
def test_regular_synth_matches_direct_sum(case, device):
    import torch
    L, nlat, nphi = case
    s = regular(L, nlat, nphi)
    o = Oracle(L, L, 1, 32 if L < 32 else L + 1, nphi)   # Legendre functions only
    q = spectral_field(L, L, seed=61)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    assert rel_l2(v, direct_synth(o, q, np.asarray(s.cos_theta), nphi)) <= TOL


@pytest.mark.parametrize("case", [(63, 130, 128), (127, 258, 256)])
def test_regular_round_trips(case, device):
    import torch
    L, nlat, nphi = case
    s = regular(L, nlat, nphi)
    q = spectral_field(L, L, seed=62)
    v = s.synth(torch.from_numpy(q).to(device))
    assert rel_l2(s.analys(v).cpu().numpy(), q) <= TOL
    sl = spectral_field(L, L, seed=63, power=3.0)
    tl = spectral_field(L, L, seed=64, power=3.0)
    sl[0] = tl[0] = 0
    vt, vp = s.synth(torch.from_numpy(sl).to(device), torch.from_numpy(tl).to(device))
    s2, t2 = s.analys(vt, vp)
    assert rel_l2(s2.cpu().numpy(), sl) <= TOL and rel_l2(t2.cpu().numpy(), tl) <= TOL
    # host API on the same grid
    vh = s.synth(q)
    assert rel_l2(vh, v.cpu().numpy()) <= 1e-15
