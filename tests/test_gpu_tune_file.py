"""GPU: the tuning file, the GPU counterpart of SHTns's SHT_LOAD_SAVE_CFG (sht_init.c:975-1061,
1430-1471).  With SHT_LOAD_SAVE_CFG given to shtns_set_grid, cushtns_tune appends its
kernel-shape choices to the file ($SHTNS_AMD_CFG), a later configuration with the same sizes
and flags picks them up at shtns_set_grid, and cushtns_tune for a batch size already in the
file returns the saved choices without timing anything."""
import os
import time

import numpy as np
import pytest
import torch

import shtns_amd
from shtns_amd.batch import tune
from shtns_amd.synthetic import spectral_batch
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
FLAGS = shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss


def make(flags):
    s = shtns_amd.sht(127)
    s.set_grid(192, 384, flags=flags, polar_opt=0.0)
    return s


def choices_of(s):
    import ctypes
    from shtns_amd import _lib
    out = (ctypes.c_int * 5)()
    _lib.load().cushtns_tuning(s._cfg, out)
    return tuple(out)


def test_tuning_file_round_trip(device, tmp_path, monkeypatch):
    path = tmp_path / "tune_cfg"
    monkeypatch.setenv("SHTNS_AMD_CFG", str(path))
    # a line for another device is skipped
    path.write_text("shtns_amd-tune 1 gfx000 127 127 1 192 384 %d 8  1 1 1 1 1\n" % FLAGS)
    a = make(FLAGS | shtns_amd.SHT_LOAD_SAVE_CFG)
    assert choices_of(a) == (0, 0, 0, 0, 0)
    t0 = time.perf_counter()
    ca = tune(a, 8)
    t_tune = time.perf_counter() - t0
    lines = path.read_text().splitlines()
    assert len(lines) == 2 and lines[1].split()[-5:] == [str(x) for x in ca]
    # a new configuration with the same sizes and flags starts from the saved shapes ...
    b = make(FLAGS | shtns_amd.SHT_LOAD_SAVE_CFG)
    assert choices_of(b) == ca
    # ... and tuning it for the same batch size reads the file instead of timing
    t0 = time.perf_counter()
    assert tune(b, 8) == ca
    t_load = time.perf_counter() - t0
    assert len(path.read_text().splitlines()) == 2
    print(f"tune: {t_tune * 1e3:.1f} ms, from the file: {t_load * 1e3:.2f} ms")
    assert t_load < t_tune
    # without the flag nothing is read
    c = make(FLAGS)
    assert choices_of(c) == (0, 0, 0, 0, 0)
    # the loaded shapes give the same transforms as the automatic ones
    q = torch.from_numpy(spectral_batch(127, 127, 8, seed=5)).to(device)
    vb, vc = b.synth(q), c.synth(q)
    assert float((vb - vc).abs().max() / vc.abs().max()) <= 1e-13
    o = Oracle(127, 127, 1, 192, 384)
    qb = b.analys(vb).cpu().numpy()
    assert np.linalg.norm(qb[3] - q[3].cpu().numpy()) / np.linalg.norm(q[3].cpu().numpy()) <= 1e-12
    w = vb[3].cpu().numpy()
    assert np.linalg.norm(w - o.synth(q[3].cpu().numpy())) / np.linalg.norm(w) <= 1e-12
