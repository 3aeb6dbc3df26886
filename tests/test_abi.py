"""The C-ABI library: it loads, exports every symbol include/*.h declares, keeps SHTns's
public struct prefix and macros (a C caller compiled against include/shtns.h works), and
its host-side functions agree with the oracle.  No GPU compute is called here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import shtns_amd
from shtns_amd import _lib
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("shtns.h", "shtns_hip.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, flags=re.M):
            if m.group(0).lstrip().startswith(("typedef", "#")):
                continue
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 50
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED_SYMBOLS) == names


def test_c_caller_against_public_header(tmp_path):
    exe = tmp_path / "abi_caller"
    subprocess.check_call(["gcc", "-std=c99", "-O1", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "c", "abi_caller.c"), "-o", str(exe),
                           "-L", os.path.join(ROOT, "shtns_amd"), "-lshtns_amd",
                           "-Wl,-rpath," + os.path.join(ROOT, "shtns_amd"), "-lm"])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout
    kv = dict(tok.split("=") for tok in out.split())
    o = Oracle(31, 31, 1, 32, 64)
    assert int(kv["nlm"]) == o.nlm == 528
    assert (kv["lmax"], kv["mmax"], kv["mres"], kv["nlat"], kv["nlat_2"], kv["nphi"]) == ("31", "31", "1", "32", "16", "64")
    assert int(kv["nspat"]) == int(kv["ret"]) == 32 * 64
    assert int(kv["LM11"]) == 32 and int(kv["LiM"]) == o.idx(5, 3)
    assert int(kv["nlm_calc"]) == 528 and int(kv["nlm_cplx"]) == 2 * 528 - 32
    assert int(kv["LMc"]) == o.lm_cplx(5, -3) == 27
    assert abs(float(kv["ct0"]) - o.cos_theta[0]) < 2e-16 and abs(float(kv["st0"]) - o.sin_theta[0]) < 1e-15
    assert int(kv["li5"]) == 5 and int(kv["mi40"]) == 1
    assert int(kv["nw"]) == 16 and abs(float(kv["w0"]) - o.gauss_w[0]) < 1e-14 * o.gauss_w[0]
    assert abs(float(kv["y2"]) - o.legendre(2, 0.3)[0]) < 1e-15
    assert abs(float(kv["sh00"]) - np.sqrt(4 * np.pi)) < 1e-15
    assert abs(float(kv["sh11"]) + np.sqrt(2 * np.pi / 3)) < 1e-15
    assert int(kv["loopn"]) == 528 and int(kv["lsum"]) == int(sum(l * min(l + 1, 32) for l in range(32)))
    # Y_1^{-1}(theta, phi) = sqrt(3 / (8 pi)) sin(theta) e^{-i phi} (Condon-Shortley phase)
    zref = np.sqrt(3 / (8 * np.pi)) * np.sqrt(1 - 0.3 ** 2) * np.exp(-0.5j)
    assert abs(complex(float(kv["zre"]), float(kv["zim"])) - zref) < 1e-15
    # public struct prefix offsets (shtns.h:79-94 layout on LP64)
    assert (kv["off_nlat"], kv["off_li"], kv["off_ct"], kv["off_nlm_cplx"]) == ("12", "24", "40", "60")
    assert _lib.ShtnsInfo.nlat.offset == 12 and _lib.ShtnsInfo.li.offset == 24
    assert _lib.ShtnsInfo.ct.offset == 40 and _lib.ShtnsInfo.nlm_cplx.offset == 60


def test_python_mirror_sizes_and_indexing():
    s = shtns_amd.sht(40, 13, 3)
    assert s.nlm == shtns_amd.nlm_calc(40, 13, 3)
    o = Oracle(40, 13, 3, 44, 80)
    assert s.nlm == o.nlm
    for (l, m) in [(0, 0), (40, 0), (3, 3), (40, 39), (17, 12)]:
        assert s.idx(l, m) == o.idx(l, m)
        assert s.l[s.idx(l, m)] == l and s.m[s.idx(l, m)] == m
    with pytest.raises(ValueError):
        s.idx(5, 4)      # m not a multiple of mres
    s.set_grid(44, 80)
    assert s.spat_shape == (44, 80)                      # Python default: phi-contiguous
    s2 = shtns_amd.sht(40, 13, 3)
    s2.set_grid(44, 80, flags=shtns_amd.SHT_THETA_CONTIGUOUS)
    assert s2.spat_shape == (80, 44)
    assert np.allclose(s.cos_theta, o.cos_theta, rtol=0, atol=2e-16)
    assert np.allclose(s.gauss_wts(), o.gauss_w[:22], rtol=1e-12, atol=0)   # Newton weights: gcc vs clang roundings differ ~1e-13


def test_host_legendre_and_point_evaluation():
    s = shtns_amd.sht(63)
    s.set_grid(64, 128, flags=shtns_amd.SHT_THETA_CONTIGUOUS, polar_opt=0.0)
    o = Oracle(63, 63, 1, 64, 128)
    for im in (0, 1, 17, 63):
        for x in (0.9, -0.2):
            assert np.allclose(s.legendre(im, x), o.legendre(im, x), rtol=0, atol=1e-14)
    from shtns_amd.synthetic import spectral_field
    q = spectral_field(63, 63, seed=3)
    v = o.synth(q)
    for (ip, it) in [(0, 0), (5, 17), (127, 63), (64, 31)]:
        val = s.SH_to_point(q, o.cos_theta[it], 2 * np.pi * ip / 128)
        assert abs(val - v[ip, it]) < 1e-12 * np.max(np.abs(v))


def test_sizes_validation_and_errors_without_gpu():
    from conftest import gpu_available
    s = shtns_amd.sht(31)
    with pytest.raises(ValueError):
        s.set_grid(30, 64)       # nlat <= lmax
    s.set_grid(33, 64)           # odd nlat: the equator row is its own mirror
    assert s.nlat == 33 and s.gauss_wts().size == 17
    with pytest.raises(RuntimeError):
        s.set_grid(32, 66)       # nphi/2 = 33 = 3 * 11 has a factor the FFT does not take
    s.set_grid(32, 64)
    if not gpu_available():
        with pytest.raises(RuntimeError, match="no usable GPU"):
            s.synth(np.zeros(s.nlm, dtype=complex))
        assert _lib.load().shtns_last_error() != 0


def test_load_save_flag_accepted_without_gpu(tmp_path, monkeypatch):
    """SHT_LOAD_SAVE_CFG is accepted by shtns_set_grid; with no GPU nothing is read or written."""
    monkeypatch.setenv("SHTNS_AMD_CFG", str(tmp_path / "cfg"))
    s = shtns_amd.sht(31)
    s.set_grid(48, 96, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss | shtns_amd.SHT_LOAD_SAVE_CFG,
               polar_opt=0.0)
    assert s.nlat == 48 and s.nphi == 96
    assert not (tmp_path / "cfg").exists()
