"""GPU: analysis blocks that carry two field groups of one order (six-tile grids, more than
eight fields: 12-wave blocks, waves tw + 6g for tile tw of group g).

Checked on lmax=511, Gauss 768 x 1536 (six latitude tiles of 64 pairs):
* fields 0..7 of 9-, 12- and 16-field batches are bit-identical to the analysis of the same
  eight fields as one batch (one group per block: the same per-wave contraction and the same
  fixed order of the six-wave sum), so the second group's waves, barriers and reduction
  leave the first group's results untouched;
* every field of those batches (the second group's fields included, also when it is partly
  empty) against the oracle's analysis of unfiltered data, rel-L2 <= 1e-12;
* a vector batch of 5 (S,T) pairs (10 spin fields: two groups, the second partly empty)
  against the oracle's vector analysis, rel-L2 <= 1e-12 per field."""
import numpy as np
import pytest
import torch

import shtns_amd
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
L, NLAT, NPHI = 511, 768, 1536
TOL = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


@pytest.fixture(scope="module")
def cfg():
    s = shtns_amd.sht(L)
    s.set_grid(NLAT, NPHI, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s


@pytest.fixture(scope="module")
def oracle():
    return Oracle(L, L, 1, NLAT, NPHI)


@pytest.mark.parametrize("nf", [9, 12, 16])
def test_two_groups_scalar(cfg, oracle, device, nf):
    rng = np.random.default_rng(70 + nf)
    w = rng.standard_normal((nf, NPHI, NLAT))   # unfiltered data: every degree is reached
    wt = torch.from_numpy(w).to(device)
    q = cfg.analys(wt).cpu().numpy()
    q8 = cfg.analys(wt[:8].contiguous()).cpu().numpy()
    assert np.array_equal(q[:8], q8), f"{nf} fields: first group differs from the 8-field batch"
    errs = [rel_l2(q[f], oracle.analys(w[f])) for f in range(nf)]
    print(f"\n{nf} fields: max rel-L2 vs oracle {max(errs):.2e}")
    assert max(errs) <= TOL


def test_two_groups_vector(cfg, oracle, device):
    npair = 5
    rng = np.random.default_rng(90)
    vt = rng.standard_normal((npair, NPHI, NLAT))
    vp = rng.standard_normal((npair, NPHI, NLAT))
    s, t = cfg.analys(torch.from_numpy(vt).to(device), torch.from_numpy(vp).to(device))
    s, t = s.cpu().numpy(), t.cpu().numpy()
    for f in range(npair):
        so, to = oracle.analys_sphtor(vt[f], vp[f])
        es, et = rel_l2(s[f], so), rel_l2(t[f], to)
        assert es <= TOL and et <= TOL, (f, es, et)
