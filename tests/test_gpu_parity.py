"""GPU parity: the HIP transforms (through the C-ABI library) against the CPU oracle on the
same seeded inputs.  FP64 bar: relative L2 error <= 1e-12."""
import numpy as np
import pytest

import shtns_amd
from shtns_amd.synthetic import spectral_batch, spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12          # relative L2, FP64 plain path vs oracle


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def make(lmax, mmax, mres, nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS):
    s = shtns_amd.sht(lmax, mmax, mres)
    s.set_grid(nlat, nphi, flags=flags | shtns_amd.sht_gauss, polar_opt=0.0)
    o = Oracle(lmax, mmax, mres, nlat, nphi)
    return s, o


def to_oracle_spat(s, v):
    """oracle wants [nphi, nlat]"""
    v = np.asarray(v)
    if s.spat_shape == (s.nlat, s.nphi):
        return np.ascontiguousarray(v.T)
    return v


def from_oracle_spat(s, v):
    if s.spat_shape == (s.nlat, s.nphi):
        return np.ascontiguousarray(v.T)
    return v


CASES = [
    # lmax, mmax, mres, nlat, nphi
    (127, 127, 1, 128, 256),      # config 1 (reference CPU-runnable case)
    (31, 31, 1, 32, 64),
    (63, 63, 1, 96, 192),
    (255, 255, 1, 384, 768),
    (100, 40, 1, 120, 100),       # mmax < lmax, nphi = 4*25
    (60, 15, 4, 64, 48),          # mres = 4
    (2, 2, 1, 4, 6),              # tiny
]


@pytest.mark.parametrize("case", CASES)
def test_synth_matches_oracle(case, device):
    import torch
    s, o = make(*case)
    q = spectral_field(case[0], case[1], seed=0, mres=case[2])
    ref = o.synth(q)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    assert rel_l2(v, ref) <= TOL, (case, rel_l2(v, ref))


@pytest.mark.parametrize("case", CASES)
def test_analys_matches_oracle(case, device):
    import torch
    s, o = make(*case)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((s.nphi, s.nlat))     # not band-limited: pins aliasing behaviour too
    ref = o.analys(v)
    q = s.analys(torch.from_numpy(v).to(device)).cpu().numpy()
    assert rel_l2(q, ref) <= TOL, (case, rel_l2(q, ref))


def test_config1_round_trip(device):
    import torch
    s, o = make(127, 127, 1, 128, 256)
    q = spectral_field(127, 127, seed=0)
    qd = torch.from_numpy(q).to(device)
    v = s.synth(qd)
    q2 = s.analys(v).cpu().numpy()
    assert rel_l2(q2, q) <= 1e-13
    # spatial side against the oracle
    assert rel_l2(v.cpu().numpy(), o.synth(q)) <= TOL


@pytest.mark.parametrize("ltr", [0, 1, 17, 64, 126])
def test_truncated_l(ltr, device):
    import torch
    s, o = make(127, 127, 1, 128, 256)
    q = spectral_field(127, 127, seed=3)
    v = s.synth(torch.from_numpy(q).to(device), ltr=ltr).cpu().numpy()
    assert rel_l2(v, o.synth(q, ltr)) <= TOL
    w = np.random.default_rng(2).standard_normal((s.nphi, s.nlat))
    qa = s.analys(torch.from_numpy(w).to(device), ltr=ltr).cpu().numpy()
    qr = o.analys(w, ltr)
    assert rel_l2(qa, qr) <= TOL
    # coefficients above ltr are zero
    assert np.all(qa[s.l > ltr] == 0)


def test_batch_equals_single(device):
    import torch
    s, o = make(63, 63, 1, 96, 192)
    qb = spectral_batch(63, 63, 9, seed=10)
    vb = s.synth(torch.from_numpy(qb).to(device)).cpu().numpy()
    for f in range(9):
        v1 = s.synth(torch.from_numpy(qb[f]).to(device)).cpu().numpy()
        assert np.array_equal(vb[f], v1)
        assert rel_l2(vb[f], o.synth(qb[f])) <= TOL
    qa = s.analys(torch.from_numpy(vb).to(device)).cpu().numpy()
    for f in range(9):
        assert rel_l2(qa[f], qb[f]) <= 1e-13
        # 9 fields run as two field groups (8 + 1) of one launch; each field's result is the
        # single-field one bit for bit, and the oracle's
        q1 = s.analys(torch.from_numpy(vb[f]).to(device)).cpu().numpy()
        assert np.array_equal(qa[f], q1), f
        assert rel_l2(qa[f], o.analys(vb[f])) <= TOL, f


def test_phi_contiguous_and_host_api(device):
    import torch
    s, o = make(95, 95, 1, 96, 192, flags=shtns_amd.SHT_PHI_CONTIGUOUS)
    assert s.spat_shape == (96, 192)
    q = spectral_field(95, 95, seed=4)
    v_host = s.synth(q)                                  # numpy -> host-pointer C API
    v_dev = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    ref = from_oracle_spat(s, o.synth(q))
    assert rel_l2(v_host, ref) <= TOL
    assert np.array_equal(v_host, v_dev)
    q_host = s.analys(v_host)
    assert rel_l2(q_host, o.analys(to_oracle_spat(s, v_host))) <= TOL


def test_south_pole_first(device):
    import torch
    s = shtns_amd.sht(63)
    s.set_grid(64, 128, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.SHT_SOUTH_POLE_FIRST, polar_opt=0.0)
    o = Oracle(63, 63, 1, 64, 128)
    assert s.cos_theta[0] < 0
    q = spectral_field(63, 63, seed=5)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    ref = o.synth(q)[:, ::-1]          # oracle is north-first
    assert rel_l2(v, ref) <= TOL


@pytest.mark.parametrize("case", [(31, 31, 1, 32, 64), (63, 63, 1, 96, 192), (120, 50, 1, 128, 120)])
def test_vector_matches_oracle(case, device):
    import torch
    s, o = make(*case)
    sl = spectral_field(case[0], case[1], seed=20, power=3.0)
    tl = spectral_field(case[0], case[1], seed=21, power=3.0)
    sl[0] = 0
    tl[0] = 0
    vt_ref, vp_ref = o.synth_sphtor(sl, tl)
    vt, vp = s.synth(torch.from_numpy(sl).to(device), torch.from_numpy(tl).to(device))
    assert rel_l2(vt.cpu().numpy(), vt_ref) <= TOL
    assert rel_l2(vp.cpu().numpy(), vp_ref) <= TOL
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal((2, s.nphi, s.nlat))
    s_ref, t_ref = o.analys_sphtor(a, b)
    s2, t2 = s.analys(torch.from_numpy(a).to(device), torch.from_numpy(b).to(device))
    assert rel_l2(s2.cpu().numpy(), s_ref) <= TOL
    assert rel_l2(t2.cpu().numpy(), t_ref) <= TOL


def test_vector_round_trip_host_api(device):
    s, o = make(127, 127, 1, 128, 256)
    sl = spectral_field(127, 127, seed=22, power=3.0)
    tl = spectral_field(127, 127, seed=23, power=3.0)
    sl[0] = tl[0] = 0
    vt, vp = s.synth(sl, tl)
    s2, t2 = s.analys(vt, vp)
    assert rel_l2(s2, sl) <= 1e-12 and rel_l2(t2, tl) <= 1e-12


def test_adjoint_identities(device):
    import torch
    s, o = make(63, 63, 1, 96, 192)
    q = torch.from_numpy(spectral_field(63, 63, seed=30)).to(device)
    g = torch.from_numpy(np.random.default_rng(8).standard_normal((s.nphi, s.nlat))).to(device)
    # <S q, g> = <q, S^T g>  with <a,b> = sum Re(conj a b)
    lhs = float((s.synth(q) * g).sum())
    rhs = float((torch.conj(q) * s.synth_adjoint(g)).real.sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # <A v, b> = <v, A^T b>
    b = torch.from_numpy(spectral_field(63, 63, seed=31)).to(device)
    lhs = float((torch.conj(s.analys(g)) * b).real.sum())
    rhs = float((g * s.analys_adjoint(b)).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_config2_full_size_round_trip(device):
    """lmax=1023, 1536x3072, batch of 16 (seed 1): round trip on the whole batch, and one
    field of the batch against the oracle."""
    import torch
    s = shtns_amd.sht(1023)
    s.set_grid(1536, 3072, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    qb = torch.from_numpy(spectral_batch(1023, 1023, 16, seed=1)).to(device)
    v = s.synth(qb)
    q2 = s.analys(v)
    err = torch.linalg.norm((q2 - qb).flatten()) / torch.linalg.norm(qb.flatten())
    assert float(err) <= 1e-12
    o = Oracle(1023, 1023, 1, 1536, 3072)
    f = 5
    ref = o.synth(qb[f].cpu().numpy())
    assert rel_l2(v[f].cpu().numpy(), ref) <= TOL
