"""GPU: the C boundary as SHTns callers use it, and entry points not covered elsewhere.

* a C99 program built with gcc against include/shtns.h + include/shtns_hip.h runs the host
  transforms (SH_to_spat / spat_to_SH on shtns_malloc'd arrays), the device transforms
  (cu_SH_to_spat / cu_spat_to_SH on hipMalloc'd arrays) and the single-precision
  cu_SH_to_spat_float / cu_spat_to_SH_float (shtns_cuda.h:38-44), as time_SHT.c:304-315 and
  cudamaxrelerror.cu:43-68 do; outputs are checked against the oracle;
* cushtns_clone (shtns_cuda.h:63) gives the same results on its own stream, and while a clone
  (or an uneven batch, made of clones) lives, the source's grid cannot be changed or
  released (its tables are borrowed);
* SHsph_to_spat / SHtor_to_spat (shtns.h:300-301) equal SHsphtor_to_spat with the other
  field zero;
* shtns_set_grid_auto with nlat = nphi = 0 picks SHTns's sizes (sht_init.c:1309-1345)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import torch

import shtns_amd
from shtns_amd import _lib
from shtns_amd.batch import UnevenBatch
from shtns_amd.synthetic import spectral_field
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def p(a):
    return ctypes.c_void_p(a.data_ptr()) if hasattr(a, "data_ptr") else a.ctypes.data_as(ctypes.c_void_p)


def test_c_caller_runs_transforms(tmp_path, device):
    exe = tmp_path / "gpu_caller"
    subprocess.check_call(["gcc", "-std=c99", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                           "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "c", "gpu_caller.c"), "-o", str(exe),
                           "-L", os.path.join(ROOT, "shtns_amd"), "-lshtns_amd", "-L", "/opt/rocm/lib", "-lamdhip64",
                           "-Wl,-rpath," + os.path.join(ROOT, "shtns_amd"), "-lm"])
    L, nlat, nphi = 127, 128, 256
    o = Oracle(L, L, 1, nlat, nphi)
    q = spectral_field(L, L, seed=50)
    q.tofile(tmp_path / "in.bin")
    r = subprocess.run([str(exe), str(L), str(nlat), str(nphi), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    ns, nq = nlat * nphi, 2 * o.nlm
    assert raw.size == 3 * (ns + nq)
    parts, at = [], 0
    for _ in range(3):
        parts.append((raw[at:at + ns].reshape(nphi, nlat), raw[at + ns:at + ns + nq].view(np.complex128)))
        at += ns + nq
    v_ref = o.synth(q)
    for name, (v, qa) in zip(("host", "device"), parts[:2]):
        assert rel_l2(v, v_ref) <= TOL, name
        assert rel_l2(qa, o.analys(v)) <= TOL and rel_l2(qa, q) <= TOL, name
    vf, qf = parts[2]          # FP32 data and contraction, FP64 recurrence and FFT
    assert rel_l2(vf, v_ref) <= 2e-6 and rel_l2(qf, q) <= 4e-6


def test_clone_and_live_clone_guard(device):
    lib = _lib.load()
    L, nlat, nphi = 63, 96, 192
    s = shtns_amd.sht(L)
    s.set_grid(nlat, nphi, flags=FLAGS, polar_opt=0.0)
    q = torch.from_numpy(np.stack([spectral_field(L, L, seed=51 + f) for f in range(3)])).to(device)
    ref = s.synth(q)
    st = torch.cuda.Stream(device)
    k = lib.cushtns_clone(s._cfg, ctypes.c_void_p(st.cuda_stream), ctypes.c_void_p(st.cuda_stream))
    assert k
    try:
        out = torch.empty_like(ref)
        torch.cuda.synchronize()
        with torch.cuda.stream(st):
            lib.cu_SH_to_spat_batch(k, p(q), p(out), L, 3, s.nlm, s.nspat)
        st.synchronize()
        _lib.check_error("clone")
        assert torch.equal(out, ref)
        # the source's tables are borrowed: re-gridding or releasing it is refused
        assert lib.shtns_set_grid(s._cfg, FLAGS, 0.0, 128, 256) == 0 and lib.shtns_last_error() == 1
        lib.cushtns_release_gpu(s._cfg)
        assert lib.shtns_last_error() == 1
        lib.shtns_unset_grid(s._cfg)
        assert lib.shtns_last_error() == 1
        assert s._cfg.contents.nlat == nlat
        out.zero_()
        with torch.cuda.stream(st):
            lib.cu_SH_to_spat_batch(k, p(q), p(out), L, 3, s.nlm, s.nspat)
        st.synchronize()
        assert torch.equal(out, ref)
        # a clone cannot take a grid of its own
        assert lib.shtns_set_grid(k, FLAGS, 0.0, 128, 256) == 0 and lib.shtns_last_error() == 1
    finally:
        lib.shtns_destroy(k)
    # an uneven batch clones its configurations: same guard while it lives
    b = UnevenBatch([s, s])
    try:
        assert lib.shtns_set_grid(s._cfg, FLAGS, 0.0, 128, 256) == 0 and lib.shtns_last_error() == 1
        b.arrays(1)[0].copy_(q[1])
        b.synth()
        torch.cuda.synchronize()
        assert rel_l2(b.arrays(1)[1].cpu().numpy(), ref[1].cpu().numpy()) <= 1e-14
    finally:
        b.close()
    assert lib.shtns_set_grid(s._cfg, FLAGS, 0.0, 128, 256) == 128 * 256
    assert lib.shtns_last_error() == 0


def test_sph_and_tor_only_synthesis(device):
    lib = _lib.load()
    L, nlat, nphi = 63, 96, 192
    s = shtns_amd.sht(L)
    s.set_grid(nlat, nphi, flags=FLAGS, polar_opt=0.0)
    sl = spectral_field(L, L, seed=55, power=3.0)
    tl = spectral_field(L, L, seed=56, power=3.0)
    sl[0] = tl[0] = 0
    zero = np.zeros_like(sl)
    vt, vp = np.empty(s.spat_shape), np.empty(s.spat_shape)
    lib.SHsph_to_spat(s._cfg, p(sl), p(vt), p(vp))
    _lib.check_error("SHsph_to_spat")
    rt, rp = s.synth(sl, zero)
    assert rel_l2(vt, rt) <= 1e-14 and rel_l2(vp, rp) <= 1e-14
    lib.SHtor_to_spat(s._cfg, p(tl), p(vt), p(vp))
    _lib.check_error("SHtor_to_spat")
    rt, rp = s.synth(zero, tl)
    assert rel_l2(vt, rt) <= 1e-14 and rel_l2(vp, rp) <= 1e-14


@pytest.mark.parametrize("nl_order", [1, 2])
def test_set_grid_auto_sizes(nl_order, device):
    L = 100
    s = shtns_amd.sht(L)
    nlat, nphi = s.set_grid(0, 0, flags=FLAGS, polar_opt=0.0, nl_order=nl_order)   # shtns_set_grid_auto
    _lib.check_error("set_grid_auto")
    # Gauss: nlat > (nl_order+1) lmax / 2, even; nphi > (nl_order+1) mmax, FFT-friendly
    assert nlat % 2 == 0 and nlat > (nl_order + 1) * L / 2 and nlat <= (nl_order + 1) * L / 2 + 4
    assert nphi % 2 == 0 and nphi > (nl_order + 1) * L
    assert s.nspat == nlat * nphi
    o = Oracle(L, L, 1, nlat, nphi)
    q = spectral_field(L, L, seed=57)
    v = s.synth(torch.from_numpy(q).to(device)).cpu().numpy()
    assert rel_l2(v, o.synth(q)) <= TOL
