"""GPU: uneven batches replayed from a HIP graph (cushtns_batch_*) and the kernel-shape
autotuner (cushtns_tune), against the CPU oracle.

The graph is replayed several times with the program's own eager work and synchronizations
between replays, with the inputs changed between replays, and every replay is checked."""
import numpy as np
import pytest

import shtns_amd
from shtns_amd.batch import UnevenBatch, tune
from shtns_amd.synthetic import spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
TOL_F32 = 2e-6


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def gauss_cfg(lmax):
    """Grid of the uneven-batch workload: nlat = 3(lmax+1)/2, nphi = 2 nlat."""
    nlat = 3 * (lmax + 1) // 2
    s = shtns_amd.sht(lmax)
    s.set_grid(nlat, 2 * nlat, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s, Oracle(lmax, lmax, 1, nlat, 2 * nlat)


def test_uneven_batch_graph_replays(device):
    import torch
    lmaxes = [15, 31, 63]
    cfgs = [gauss_cfg(L) for L in lmaxes]
    order = [2, 0, 1, 1, 2, 0, 0, 1, 2, 2, 1, 0]          # transform i uses cfgs[order[i]]
    b = UnevenBatch([cfgs[k][0] for k in order])
    for rep in range(3):
        qs = []
        for i, k in enumerate(order):
            q = spectral_field(lmaxes[k], lmaxes[k], seed=100 * rep + i)
            b.arrays(i)[0].copy_(torch.from_numpy(q))
            qs.append(q)
        b.synth()
        torch.cuda.synchronize()
        for i, k in enumerate(order):
            v = b.arrays(i)[1].cpu().numpy()
            assert rel_l2(v, cfgs[k][1].synth(qs[i])) <= TOL, (rep, i)
        # eager work on the same device between the replays
        x = torch.randn(1 << 20, device=device, dtype=torch.float64)
        assert torch.isfinite(x.sum())
        b.analys()
        torch.cuda.synchronize()
        for i, k in enumerate(order):
            q2 = b.arrays(i)[0].cpu().numpy()
            v = b.arrays(i)[1].cpu().numpy()
            assert rel_l2(q2, cfgs[k][1].analys(v)) <= TOL, (rep, i)
            assert rel_l2(q2, qs[i]) <= TOL, (rep, i)   # band-limited round trip
    b.close()


def test_uneven_batch_split_precision_and_tuning(device):
    import torch
    lmaxes = [31, 63]
    cfgs = [gauss_cfg(L) for L in lmaxes]
    order = [0, 1, 1, 0, 1]
    b = UnevenBatch([cfgs[k][0] for k in order], prec=shtns_amd.SHT_PREC_SPLIT, tune=True)
    qs = []
    for i, k in enumerate(order):
        q = spectral_field(lmaxes[k], lmaxes[k], seed=7 + i)
        b.arrays(i)[0].copy_(torch.from_numpy(q))
        qs.append(q)
    b.synth()
    b.analys()
    torch.cuda.synchronize()
    for i, k in enumerate(order):
        assert rel_l2(b.arrays(i)[0].cpu().numpy(), qs[i]) <= TOL_F32
    b.close()
    # the caller's configurations are untouched (the batch works on clones)
    v = cfgs[0][0].synth(torch.from_numpy(qs[0]).to(device)).cpu().numpy()
    assert rel_l2(v, cfgs[0][1].synth(qs[0])) <= TOL


def test_autotuner_choices_and_results(device):
    import torch
    s, o = gauss_cfg(127)
    choices = tune(s, 8)
    assert len(choices) == 5
    ntiles = (s.nlat // 2 + 63) // 64
    assert choices[0] in (1, 2) and choices[2] in (1, 2) and choices[3] in range(7) and choices[4] in range(7)
    assert choices[1] == 0 or ntiles % choices[1] == 0
    q = np.stack([spectral_field(127, 127, seed=30 + f) for f in range(8)])
    v = s.synth(torch.from_numpy(q).to(device))
    q2 = s.analys(v).cpu().numpy()
    v = v.cpu().numpy()
    for f in range(8):
        assert rel_l2(v[f], o.synth(q[f])) <= TOL
        assert rel_l2(q2[f], q[f]) <= TOL


@pytest.mark.parametrize("case", [(1023, 1536, 3072), (255, 384, 768), (63, 96, 192), (127, 128, 256)])
def test_every_fft_kernel_matches_oracle(case, device):
    """Each FFT kernel (register, LDS-staged, cooperative 256/128, cooperative with 8-byte
    staging 256/512), forced through cushtns_set_tuning, in both directions."""
    import torch
    from shtns_amd.batch import set_tuning
    L, nlat, nphi = case
    s = shtns_amd.sht(L)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    nf = 3 if L >= 1023 else 5
    q = np.stack([spectral_field(L, L, seed=40 + f) for f in range(nf)])
    o = Oracle(L, L, 1, nlat, nphi) if L < 1023 else None
    qd = torch.from_numpy(q).to(device)
    ref_v, ref_q = None, None
    for kind in range(1, 7):
        set_tuning(s, [0, 0, 0, kind, kind])
        v = s.synth(qd)
        q2 = s.analys(v).cpu().numpy()
        v = v.cpu().numpy()
        if o is not None:
            for f in range(nf):
                assert rel_l2(v[f], o.synth(q[f])) <= TOL, (kind, f)
        if ref_v is None:
            ref_v = v
        assert rel_l2(v, ref_v) <= TOL, kind
        assert rel_l2(q2, q) <= TOL, kind
    set_tuning(s, [0, 0, 0, 0, 0])
