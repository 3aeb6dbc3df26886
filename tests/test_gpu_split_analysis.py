"""GPU: the latitude-split analysis used for one or two fields when an order has more than 12
latitude tiles (here lmax = 2047 on 3072 x 6144: 24 tiles, two splits whose partial spectra
are summed).  Checked against the extended-precision sampled oracle (tests/sampled_oracle.py,
pinned in tests/test_sampled_oracle.py) on four orders, by round trips, and against the
unsplit kernels: a batch of three fields (and a vector batch of two pairs) runs the
one-launch analysis, whose results for each field must equal the split analysis of that
field alone."""
import numpy as np
import pytest
import torch

import shtns_amd
from shtns_amd.synthetic import spectral_batch
from oracle.oracle import Oracle
from sampled_oracle import anal_orders_ld

pytestmark = pytest.mark.gpu
L, NLAT, NPHI = 2047, 3072, 6144


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


@pytest.fixture(scope="module")
def cfg():
    s = shtns_amd.sht(L)
    s.set_grid(NLAT, NPHI, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s


def test_scalar_split(cfg, device):
    s = cfg
    rng = np.random.default_rng(91)
    w = torch.from_numpy(rng.standard_normal((3, NPHI, NLAT))).to(device)
    q3 = s.analys(w).cpu().numpy()                      # 3 fields: one launch
    q1 = np.stack([s.analys(w[f]).cpu().numpy() for f in range(3)])   # 1 field: split
    e_paths = max(rel_l2(q1[f], q3[f]) for f in range(3))
    o = Oracle(L, L, 1, NLAT, NPHI)
    ims = [0, 1, 700, 2047]
    ref = anal_orders_ld(o, w[0].cpu().numpy(), ims)
    e_or = max(rel_l2(q1[0][o.idx(m, m):o.idx(m, m) + L + 1 - m], ref[m]) for m in ims)
    q = torch.from_numpy(spectral_batch(L, L, 1, seed=92)[0]).to(device)
    e_rt = rel_l2(s.analys(s.synth(q)).cpu().numpy(), q.cpu().numpy())
    print(f"split analysis lmax=2047: vs one-launch {e_paths:.2e}, vs extended-precision oracle {e_or:.2e}, "
          f"round trip {e_rt:.2e}")
    assert e_paths <= 1e-13 and e_or <= 2e-12 and e_rt <= 1e-12


def test_vector_split(cfg, device):
    s = cfg
    S = spectral_batch(L, L, 2, seed=93, power=3.0)
    T = spectral_batch(L, L, 2, seed=95, power=3.0)
    S[:, 0] = T[:, 0] = 0
    St, Tt = torch.from_numpy(S).to(device), torch.from_numpy(T).to(device)
    vt, vp = s.synth(St, Tt)
    s2, t2 = (x.cpu().numpy() for x in s.analys(vt, vp))          # two pairs: one launch
    s1, t1 = (x.cpu().numpy() for x in s.analys(vt[0], vp[0]))    # one pair: split
    e_paths = max(rel_l2(s1, s2[0]), rel_l2(t1, t2[0]))
    e_rt = max(rel_l2(s1, S[0]), rel_l2(t1, T[0]))
    print(f"split vector analysis lmax=2047: vs one-launch {e_paths:.2e}, round trip {e_rt:.2e}")
    assert e_paths <= 1e-13 and e_rt <= 1e-12
