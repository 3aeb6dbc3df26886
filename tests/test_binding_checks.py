"""CPU: the Python binding validates arrays before any pointer reaches the C API, as SHTns's
shtns_numpy.i does (check_spatial / check_spectral, :60-90): numpy arrays of the right dtype,
C-contiguous and of exactly nspat / nlm elements; torch tensors of one call share the first
one's shape, dtype and device.  A wrong array raises and never reaches the library."""
import numpy as np
import pytest

import shtns_amd


@pytest.fixture(scope="module")
def sh():
    s = shtns_amd.sht(31)
    s.set_grid(32, 64, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    return s


def _bad_spat(sh):
    good = np.zeros(sh.spat_shape)
    return {
        "float32": good.astype(np.float32),
        "small": np.zeros(sh.nspat - 1),
        "strided": np.zeros((sh.spat_shape[0], 2 * sh.spat_shape[1]))[:, ::2],
        "list": good.tolist(),
    }


def _bad_spec(sh):
    good = sh.spec_array()
    return {
        "complex64": good.astype(np.complex64),
        "real": np.zeros(sh.nlm),
        "small": np.zeros(sh.nlm - 1, dtype=complex),
        "strided": np.zeros(2 * sh.nlm, dtype=complex)[::2],
    }


@pytest.mark.parametrize("kind", ["float32", "small", "strided", "list"])
def test_host_spatial_arrays_rejected(sh, kind):
    bad = _bad_spat(sh)[kind]
    q = sh.spec_array()
    with pytest.raises((TypeError, RuntimeError)):
        sh.SH_to_spat(q, bad)
    with pytest.raises((TypeError, RuntimeError)):
        sh.spat_to_SH(bad, q)
    good = np.zeros(sh.spat_shape)
    with pytest.raises((TypeError, RuntimeError)):
        sh.SHsphtor_to_spat(q, q, good, bad)
    with pytest.raises((TypeError, RuntimeError)):
        sh.SHqst_to_spat(q, q, q, good, good, bad)


@pytest.mark.parametrize("kind", ["complex64", "real", "small", "strided"])
def test_host_spectral_arrays_rejected(sh, kind):
    bad = _bad_spec(sh)[kind]
    v = np.zeros(sh.spat_shape)
    q = sh.spec_array()
    with pytest.raises(RuntimeError):
        sh.SH_to_spat(bad, v)
    with pytest.raises(RuntimeError):
        sh.spat_to_SH(v, bad)
    with pytest.raises(RuntimeError):
        sh.spat_to_SHsphtor(v, v, q, bad)
    with pytest.raises(RuntimeError):
        sh.spat_to_SHqst(v, v, v, bad, q, q)


def test_host_complex_arrays_rejected(sh):
    z = np.zeros(sh.spat_shape, dtype=np.complex128)
    a = np.zeros(sh.nlm_cplx, dtype=np.complex128)
    with pytest.raises(RuntimeError):
        sh.SH_to_spat_cplx(a[:-1], z)
    with pytest.raises(TypeError):
        sh.spat_cplx_to_SH(z.real.copy(), a)
    with pytest.raises(RuntimeError):
        sh.SHsphtor_to_spat_cplx(a, a, z, z[:, :-1].copy())


def test_torch_mismatched_tensors_rejected(sh):
    torch = pytest.importorskip("torch")
    q1 = torch.zeros((8, sh.nlm), dtype=torch.complex128)
    q2 = torch.zeros((sh.nlm,), dtype=torch.complex128)
    with pytest.raises(RuntimeError, match="shape, dtype and device"):
        sh.synth(q1, q2)
    with pytest.raises(RuntimeError, match="shape, dtype and device"):
        sh.synth(q1, q1, q2)
    v1 = torch.zeros((4,) + tuple(sh.spat_shape), dtype=torch.float64)
    v2 = torch.zeros((5,) + tuple(sh.spat_shape), dtype=torch.float64)
    with pytest.raises(RuntimeError, match="shape, dtype and device"):
        sh.analys(v1, v2)
    with pytest.raises(RuntimeError, match="shape, dtype and device"):
        sh.analys(v1, v1.float())
