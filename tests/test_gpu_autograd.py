"""Differentiation: hand-written adjoint kernels behind torch.autograd.Function.
gradcheck in FP64 at lmax=31 (32x64); finite-difference check of the loss
sum(w_grid * f^2) at lmax=255 (384x768); agreement with the transposed operator built
from the oracle."""
import numpy as np
import pytest

import shtns_amd
from shtns_amd import autograd as sad
from shtns_amd.synthetic import spectral_batch, spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def make(lmax, nlat, nphi):
    s = shtns_amd.sht(lmax)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS, polar_opt=0.0)
    return s


def test_gradcheck_synth_and_analys_lmax31(device):
    import torch
    s = make(31, 32, 64)
    q = torch.from_numpy(spectral_field(31, 31, seed=3)).to(device)
    # Im(Q_l0) is not an input of the transform: keep it out of the perturbed variables
    mask_m0 = torch.from_numpy((s.m == 0)).to(device)
    qv = q.clone().requires_grad_(True)

    def f_synth(x):
        x = torch.where(mask_m0, x.real.to(torch.complex128), x)
        return sad.synth(s, x)

    assert torch.autograd.gradcheck(f_synth, (qv,), eps=1e-6, atol=1e-8, rtol=1e-6)
    v = torch.from_numpy(np.random.default_rng(4).standard_normal((64, 32))).to(device).requires_grad_(True)

    def f_anal(x):
        out = sad.analys(s, x)
        return torch.view_as_real(out)

    assert torch.autograd.gradcheck(f_anal, (v,), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_adjoint_matches_oracle_transpose(device):
    """S^T g from the kernel equals the transpose of the synthesis matrix built column by
    column with the oracle (small size, all columns)."""
    import torch
    L, nlat, nphi = 7, 8, 16
    s = make(L, nlat, nphi)
    o = Oracle(L, L, 1, nlat, nphi)
    nlm = o.nlm
    cols_re, cols_im = [], []
    for i in range(nlm):
        e = np.zeros(nlm, complex)
        e[i] = 1.0
        cols_re.append(o.synth(e).ravel())
        e[i] = 1j
        cols_im.append(o.synth(e).ravel())
    Sre, Sim = np.array(cols_re).T, np.array(cols_im).T     # grid x nlm
    g = np.random.default_rng(9).standard_normal((nphi, nlat))
    ref = Sre.T @ g.ravel() + 1j * (Sim.T @ g.ravel())
    got = s.synth_adjoint(torch.from_numpy(g).to(device)).cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_fd_check_weighted_energy_lmax255(device):
    """loss = sum(w_grid * f^2), f = synth(q), batch of 64 FP32 fields (config 4); the
    gradient w.r.t. q is checked against central finite differences along random
    directions, in FP64."""
    import torch
    L, nlat, nphi, B = 255, 384, 768, 64
    s = make(L, nlat, nphi)
    rng = np.random.default_rng(3)
    w = torch.from_numpy(rng.uniform(0.5, 1.5, (nphi, nlat))).to(device)
    q32 = torch.from_numpy(spectral_batch(L, L, B, seed=3)).to(device).to(torch.complex64).requires_grad_(True)
    f = sad.synth(s, q32)
    assert f.dtype == torch.float32
    loss = (w.float() * f * f).sum()
    loss.backward()
    assert q32.grad is not None and q32.grad.shape == q32.shape
    # FP64 finite differences on a few fields along random directions
    q = q32.detach().to(torch.complex128)[:4].clone().requires_grad_(True)
    f = sad.synth(s, q)
    loss = (w * f * f).sum()
    (gq,) = torch.autograd.grad(loss, q)
    m0 = torch.from_numpy(s.m == 0).to(device)
    for t in range(3):
        d = torch.from_numpy(spectral_batch(L, L, 4, seed=100 + t)).to(device)
        d = torch.where(m0, d.real.to(torch.complex128), d)
        h = 1e-4
        lp = float((w * sad.synth(s, q.detach() + h * d) ** 2).sum())
        lm = float((w * sad.synth(s, q.detach() - h * d) ** 2).sum())
        fd = (lp - lm) / (2 * h)
        an = float((torch.conj(d) * gq).real.sum())
        assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)


def test_analysis_round_trip_gradient(device):
    """d/dV of sum|analys(V) - target|^2 via the adjoint equals the FD derivative."""
    import torch
    s = make(63, 64, 128)
    rng = np.random.default_rng(5)
    v = torch.from_numpy(rng.standard_normal((2, 128, 64))).to(device).requires_grad_(True)
    tgt = torch.from_numpy(spectral_batch(63, 63, 2, seed=6)).to(device)
    loss = (torch.abs(sad.analys(s, v) - tgt) ** 2).sum()
    (gv,) = torch.autograd.grad(loss, v)
    d = torch.from_numpy(rng.standard_normal((2, 128, 64))).to(device)
    h = 1e-5
    lp = float((torch.abs(sad.analys(s, v.detach() + h * d) - tgt) ** 2).sum())
    lm = float((torch.abs(sad.analys(s, v.detach() - h * d) - tgt) ** 2).sum())
    assert abs((lp - lm) / (2 * h) - float((gv * d).sum())) <= 1e-6 * abs(float((gv * d).sum()))
