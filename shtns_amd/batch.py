"""Uneven batches and kernel-shape tuning (the additive cushtns_batch_* / cushtns_tune entry
points of shtns_hip.h).

    b = UnevenBatch([sh_a, sh_b, sh_a, ...], prec=SHT_PREC_F64, tune=True)
    q, v = b.arrays(i)          # torch views of transform i's device arrays
    q.copy_(...)                # write inputs
    b.synth()                   # all transforms, one HIP graph launch on torch's stream
    b.analys()

Transforms sharing an `sht` object share its configuration and run as one batched launch;
the groups of different configurations are parallel branches of the graph.
"""
import ctypes

import numpy as np

from . import _lib
from . import SHT_PREC_AUTO


class _DeviceArray:
    """Exposes a device pointer through __cuda_array_interface__ (torch.as_tensor shares it)."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


def tune(sh, nfields):
    """Autotune the kernel shapes of configuration `sh` for batches of `nfields` fields;
    returns the choices (synthesis groups, analysis waves, analysis groups, synthesis FFT,
    analysis FFT), 0 meaning automatic."""
    lib = _lib.load()
    if lib.cushtns_tune(sh._cfg, int(nfields)) != 0:
        _lib.check_error("tune")
    out = (ctypes.c_int * 5)()
    lib.cushtns_tuning(sh._cfg, out)
    return tuple(out)


def set_tuning(sh, choices):
    """Set the five kernel-shape choices of configuration `sh` by hand (see tune)."""
    arr = (ctypes.c_int * 5)(*[int(c) for c in choices])
    if _lib.load().cushtns_set_tuning(sh._cfg, arr) != 0:
        _lib.check_error("set_tuning")


class UnevenBatch:
    def __init__(self, shts, prec=SHT_PREC_AUTO, tune=False):
        import torch
        if not shts:
            raise ValueError("empty batch")
        for s in shts:
            s._need_grid()
        self.shts = list(shts)
        lib = _lib.load()
        arr = (ctypes.c_void_p * len(shts))(*[ctypes.cast(s._cfg, ctypes.c_void_p) for s in shts])
        h = lib.cushtns_batch_create(len(shts), arr, int(prec), 1 if tune else 0)
        if not h:
            _lib.check_error("batch_create")
            raise RuntimeError("shtns_amd: batch creation failed")
        self._h = ctypes.c_void_p(h)
        self._views = []
        dev = torch.device("cuda", torch.cuda.current_device())
        for i, s in enumerate(shts):
            qp, vp = ctypes.c_void_p(), ctypes.c_void_p()
            if lib.cushtns_batch_arrays(self._h, i, ctypes.byref(qp), ctypes.byref(vp)) != 0:
                _lib.check_error("batch_arrays")
            q = torch.as_tensor(_DeviceArray(qp.value, (s.nlm,), "<c16"), device=dev)
            v = torch.as_tensor(_DeviceArray(vp.value, tuple(s.spat_shape), "<f8"), device=dev)
            self._views.append((q, v))

    def __len__(self):
        return len(self.shts)

    def arrays(self, i):
        """(Qlm, Vr) torch views of transform i's device arrays."""
        return self._views[i]

    def _run(self, fn):
        import torch
        s = torch.cuda.current_stream().cuda_stream
        if fn(self._h, ctypes.c_void_p(s)) != 0:
            _lib.check_error("batch")

    def synth(self):
        """SH_to_spat of every transform (reads each Qlm, writes each Vr)."""
        self._run(_lib.load().cushtns_batch_SH_to_spat)

    def analys(self):
        """spat_to_SH of every transform (reads each Vr, writes each Qlm)."""
        self._run(_lib.load().cushtns_batch_spat_to_SH)

    def close(self):
        if getattr(self, "_h", None):
            import torch
            torch.cuda.synchronize()
            self._views = []
            _lib.load().cushtns_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
