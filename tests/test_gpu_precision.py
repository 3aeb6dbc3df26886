"""GPU: the reduced-precision contractions and the single-precision entry points
(cu_SH_to_spat_float / cu_spat_to_SH_float, shtns_cuda.h:40-44) against the FP64 oracle.

The recurrence, seeds, FFT and final sums stay FP64; only the degree/latitude contraction
changes.  Each test prints the measured relative L2 error; the bounds below are FP32-class
(FP32 rounding of the recurrence values and FP32 partial sums over 16 degree pairs), not the
FP64 1e-12 bar, which only SHT_PREC_F64 meets.
"""
import numpy as np
import pytest

import shtns_amd
from shtns_amd.synthetic import spectral_batch, spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

# relative L2 bounds vs the FP64 oracle
TOL_F64 = 1e-12
TOL_F32 = 2e-6       # FP32 operands, FP32 partial sums
TOL_SPLIT = 2e-6     # hi/lo operands; the FP32 partial sums bound the error
TOL_FLOAT_IO = 2e-6  # FP32 arrays (7 digits) on top of the FP32 contraction


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def make(lmax, nlat, nphi, mmax=None):
    mmax = lmax if mmax is None else mmax
    s = shtns_amd.sht(lmax, mmax)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s, Oracle(lmax, mmax, 1, nlat, nphi)


PRECS = [(shtns_amd.SHT_PREC_F64, TOL_F64), (shtns_amd.SHT_PREC_F32, TOL_F32),
         (shtns_amd.SHT_PREC_SPLIT, TOL_SPLIT)]


@pytest.mark.parametrize("prec,tol", PRECS)
@pytest.mark.parametrize("case", [(127, 128, 256), (255, 384, 768), (100, 120, 100, 40)])
def test_contraction_precision_vs_oracle(prec, tol, case, device):
    import torch
    s, o = make(*case)
    s.set_precision(prec)
    q = spectral_field(case[0], s.mmax, seed=11)
    v_ref = o.synth(q)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    q_ref = o.analys(v_ref)
    q2 = s.analys(torch.from_numpy(v_ref).to(device)).cpu().numpy()
    es, ea = rel_l2(v, v_ref), rel_l2(q2, q_ref)
    print(f"prec={prec} case={case}: synthesis {es:.2e}, analysis {ea:.2e}")
    assert es <= tol and ea <= tol, (prec, case, es, ea)


@pytest.mark.parametrize("prec", [shtns_amd.SHT_PREC_F32, shtns_amd.SHT_PREC_SPLIT])
def test_reduced_precision_batch_and_truncation(prec, device):
    import torch
    s, o = make(63, 96, 192)
    s.set_precision(prec)
    qb = spectral_batch(63, 63, 5, seed=3)
    ltr = 40
    v = s.synth(torch.from_numpy(qb).to(device), ltr=ltr).cpu().numpy()
    for f in range(5):
        qt = qb[f].copy()
        qt[s.l > ltr] = 0
        assert rel_l2(v[f], o.synth(qt)) <= TOL_F32


@pytest.mark.parametrize("case", [(127, 128, 256), (255, 384, 768), (40, 64, 100)])
def test_float_entry_points(case, device):
    """complex64 -> float32 synthesis and back, against the FP64 oracle (nphi=100 takes the
    staged FFT path, the others the register FFT with FP32 loads/stores)."""
    import torch
    s, o = make(*case)
    q = spectral_field(case[0], case[0], seed=12)
    v_ref = o.synth(q)
    v = s.synth(torch.from_numpy(q.astype(np.complex64)).to(device))
    assert v.dtype == torch.float32
    es = rel_l2(v.cpu().numpy(), v_ref)
    q2 = s.analys(torch.from_numpy(v_ref.astype(np.float32)).to(device))
    assert q2.dtype == torch.complex64
    ea = rel_l2(q2.cpu().numpy(), o.analys(v_ref.astype(np.float32).astype(np.float64)))
    print(f"float I/O case={case}: synthesis {es:.2e}, analysis {ea:.2e}")
    assert es <= TOL_FLOAT_IO and ea <= TOL_FLOAT_IO, (case, es, ea)


def test_float_entry_points_with_fp64_contraction(device):
    """SHT_PREC_F64 on float data: error is the FP32 rounding of the arrays only."""
    import torch
    s, o = make(127, 128, 256)
    s.set_precision(shtns_amd.SHT_PREC_F64)
    q = spectral_field(127, 127, seed=13).astype(np.complex64)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    v_ref = o.synth(q.astype(np.complex128))
    assert rel_l2(v, v_ref) <= 1e-7


def test_set_precision_rejects_unknown(device):
    s, _ = make(31, 32, 64)
    with pytest.raises(RuntimeError):
        s.set_precision(7)
    s.set_precision(shtns_amd.SHT_PREC_AUTO)
