"""GPU: cu_SH_to_spat_batch_noalloc (no counterpart in SHTns; for callers that cannot
allocate device memory) against cu_SH_to_spat_batch, called through the C ABI.

Both entry points run the same kernels with the same shapes, so the outputs must agree bit
for bit.  Two calls in a row share the library's fixed workspace; the second is smaller
(another configuration, fewer fields), so stale workspace contents from the first call must
not leak into it.  A call that needs more than the fixed workspace is refused on the host
before anything is launched."""
import ctypes

import numpy as np
import pytest

import shtns_amd
from shtns_amd import _lib
from shtns_amd.synthetic import spectral_field

pytestmark = pytest.mark.gpu


def _cfg(lmax, nlat, nphi):
    s = shtns_amd.sht(lmax)
    s.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s


def _run(lib, fn, s, q, nf, torch):
    v = torch.full((nf,) + tuple(s.spat_shape), float("nan"), dtype=torch.float64, device=q.device)
    s._stream(q)
    fn(s._cfg, ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(v.data_ptr()), s.lmax, nf, s.nlm, s.nspat)
    _lib.check_error(fn.__name__)
    torch.cuda.synchronize()
    return v.cpu().numpy()


def test_noalloc_matches_allocating_entry_point(device):
    import torch
    lib = _lib.load()
    # each call needs less workspace than the one before (lmax=1023, one field: about 59 MB)
    cases = [(1023, 1536, 3072, 1, 10), (255, 384, 768, 3, 11), (63, 96, 192, 1, 12)]
    for lmax, nlat, nphi, nf, seed in cases:
        print(f"noalloc case lmax={lmax} nlat={nlat} nphi={nphi} nfields={nf}", flush=True)
        s = _cfg(lmax, nlat, nphi)
        q = torch.from_numpy(np.stack([spectral_field(lmax, lmax, seed=seed + f) for f in range(nf)])).to(device)
        ref = _run(lib, lib.cu_SH_to_spat_batch, s, q, nf, torch)
        out = _run(lib, lib.cu_SH_to_spat_batch_noalloc, s, q, nf, torch)
        assert np.isfinite(ref).all()
        assert np.array_equal(out, ref), float(np.abs(out - ref).max())


def test_noalloc_refuses_oversized_call(device):
    import torch
    lib = _lib.load()
    # 2 fields at lmax=1023 on 1536 x 3072 need about 118 MB of workspace (limit 64 MiB)
    s = _cfg(1023, 1536, 3072)
    q = torch.zeros((2, s.nlm), dtype=torch.complex128, device=device)
    v = torch.zeros((2,) + tuple(s.spat_shape), dtype=torch.float64, device=device)
    lib.cu_SH_to_spat_batch_noalloc(s._cfg, ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(v.data_ptr()), s.lmax, 2,
                                    s.nlm, s.nspat)
    assert lib.shtns_last_error() == 2
    assert b"workspace" in lib.shtns_last_error_string()
    torch.cuda.synchronize()
    assert float(v.abs().max()) == 0.0   # nothing was launched
