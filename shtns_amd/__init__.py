"""shtns_amd -- spherical harmonic transforms on one AMD MI355X (gfx950).

Python mirror of SHTns's ``shtns.sht`` class (shtns_numpy.i:160-420), backed by the C-ABI
library ``libshtns_amd.so`` (include/shtns.h, include/shtns_hip.h):

* numpy arrays go through the host-pointer API (copied to the GPU and back);
* torch CUDA tensors go through the device-pointer API on torch's current stream, with any
  number of leading batch dimensions (one launch for the whole batch).

Spectral arrays are complex128 ``[..., nlm]`` in SHTns's LM order (``sht.idx(l, m)``);
spatial arrays are float64 ``[..., *spat_shape]`` with ``spat_shape = (nphi, nlat)`` for
theta-contiguous grids and ``(nlat, nphi)`` for phi-contiguous grids (the Python default,
as in SHTns, shtns_numpy.i:214).
"""
import ctypes

import numpy as np

from . import _lib

# constants (shtns.h:48-75)
sht_orthonormal, sht_fourpi, sht_schmidt, sht_for_rotations = 0, 1, 2, 3
SHT_NO_CS_PHASE = 256 * 4
SHT_REAL_NORM = 256 * 8
sht_gauss, sht_auto, sht_reg_fast, sht_reg_dct, sht_quick_init, sht_reg_poles, sht_gauss_fly = range(7)
SHT_NATIVE_LAYOUT = 0
SHT_THETA_CONTIGUOUS = 256
SHT_PHI_CONTIGUOUS = 256 * 2
SHT_SOUTH_POLE_FIRST = 256 * 32
SHT_SCALAR_ONLY = 256 * 16
SHT_LOAD_SAVE_CFG = 256 * 64
SHT_ALLOW_GPU = 256 * 128
SHT_ALLOW_PADDING = 256 * 256
# contraction precision of the GPU Legendre step (shtns_hip.h)
SHT_PREC_AUTO, SHT_PREC_F64, SHT_PREC_F32, SHT_PREC_SPLIT, SHT_PREC_SLICED = -1, 0, 1, 2, 3

__all__ = ["sht", "rotation", "nlm_calc", "build_info"]


def nlm_calc(lmax, mmax, mres):
    return int(_lib.load().nlm_calc(lmax, mmax, mres))


def build_info():
    return _lib.load().shtns_get_build_info().decode()


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(a):
    if _is_torch(a):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


class sht:
    """Spherical harmonic transform of given sizes (SHTns ``shtns.sht``)."""

    def __init__(self, lmax, mmax=-1, mres=1, norm=sht_orthonormal, nthreads=0):
        if lmax < 2:
            raise ValueError("lmax < 2 not allowed")
        if mres <= 0:
            raise ValueError("mres <= 0 invalid")
        if mmax < 0:
            mmax = lmax // mres
        if mmax * mres > lmax:
            raise ValueError("lmax < mmax*mres invalid")
        lib = _lib.load()
        lib.shtns_use_threads(nthreads)
        self._cfg = lib.shtns_create(lmax, mmax, mres, norm)
        if not self._cfg:
            _lib.check_error("shtns_create")
            raise RuntimeError("shtns_create failed")
        info = self._cfg.contents
        self.lmax, self.mmax, self.mres, self.nlm = info.lmax, info.mmax, info.mres, info.nlm
        self.nlm_cplx = info.nlm_cplx
        self.nlat = self.nphi = self.nspat = 0
        self.l = np.zeros(self.nlm, dtype=np.int32)
        self.m = np.zeros(self.nlm, dtype=np.int32)
        for mm in range(0, self.mmax * self.mres + 1, self.mres):
            for ll in range(mm, self.lmax + 1):
                i = self.idx(ll, mm)
                self.l[i], self.m[i] = ll, mm
        self.l.flags.writeable = False
        self.m.flags.writeable = False
        self._layout = 0

    def __del__(self):
        cfg = getattr(self, "_cfg", None)
        if cfg:
            try:
                _lib.load().shtns_destroy(cfg)
            except Exception:
                pass
            self._cfg = None

    # ---- grid ----
    def set_grid(self, nlat=0, nphi=0, flags=sht_quick_init, polar_opt=1.0e-10, nl_order=1):
        """Choose the Gauss grid (shtns_numpy.i:205-222).  Defaults to phi-contiguous
        layout unless SHT_THETA_CONTIGUOUS is given, as SHTns's Python binding does."""
        if nlat != 0 and nlat <= self.lmax:
            raise ValueError("nlat <= lmax")
        if nphi != 0 and nphi <= self.mmax * 2:
            raise ValueError("nphi <= 2*mmax")
        if not (flags & SHT_THETA_CONTIGUOUS):
            flags |= SHT_PHI_CONTIGUOUS
        lib = _lib.load()
        a, b = ctypes.c_int(nlat), ctypes.c_int(nphi)
        n = lib.shtns_set_grid_auto(self._cfg, flags, polar_opt, nl_order, ctypes.byref(a), ctypes.byref(b))
        if n == 0:
            _lib.check_error("set_grid")
        info = self._cfg.contents
        self.nlat, self.nphi, self.nspat = info.nlat, info.nphi, info.nspat
        self._layout = flags & (SHT_PHI_CONTIGUOUS | SHT_THETA_CONTIGUOUS)
        self.cos_theta = np.ctypeslib.as_array(info.ct, shape=(self.nlat,)).copy()
        self.cos_theta.flags.writeable = False
        if self._layout & SHT_PHI_CONTIGUOUS:
            self.spat_shape = (self.nlat, self.nphi)
        else:
            self.spat_shape = (self.nphi, self.nlat)
        return self.nlat, self.nphi

    def gauss_wts(self):
        """Quadrature weights of the latitude pairs, north to equator ((nlat+1)/2 values; the
        equator row of an odd grid carries half its weight, as in SHTns)."""
        w = np.empty((self.nlat + 1) // 2)
        _lib.load().shtns_gauss_wts(self._cfg, _ptr(w))
        return w

    def print_info(self):
        _lib.load().shtns_print_cfg(self._cfg)

    def sh00_1(self):
        return _lib.load().sh00_1(self._cfg)

    def sh10_ct(self):
        return _lib.load().sh10_ct(self._cfg)

    def sh11_st(self):
        return _lib.load().sh11_st(self._cfg)

    def shlm_e1(self, l, m):
        return _lib.load().shlm_e1(self._cfg, l, m)

    def idx(self, l, m):
        if l > self.lmax:
            raise ValueError("l invalid")
        if m > l or m > self.mmax * self.mres or m % self.mres != 0:
            raise ValueError("m invalid")
        im = m // self.mres
        return (im * (2 * self.lmax + 2 - (im + 1) * self.mres)) // 2 + l

    def spec_array(self, im=-1):
        if im < 0:
            return np.zeros(self.nlm, dtype=complex)
        return np.zeros(self.lmax + 1 - im * self.mres, dtype=complex)

    def spat_array(self):
        if self.nlat == 0:
            raise RuntimeError("Grid not set. Call .set_grid() method.")
        return np.zeros(self.spat_shape)

    def legendre(self, im, x, lmax=None):
        """Normalized associated Legendre functions Y_l^m(x), l = m..lmax."""
        if lmax is None:
            lmax = self.lmax
        y = np.empty(lmax - im * self.mres + 1)
        _lib.load().legendre_sphPlm_array(self._cfg, lmax, im, float(x), _ptr(y))
        return y

    def SH_to_point(self, qlm, cost, phi):
        q = np.ascontiguousarray(qlm, dtype=np.complex128)
        return _lib.load().SH_to_point(self._cfg, _ptr(q), float(cost), float(phi))

    def mul_ct_matrix(self):
        """Coefficients (2*nlm) of the multiplication by cos(theta), for SH_mul_mx (shtns_numpy.i:272-276)."""
        mx = np.empty(2 * self.nlm)
        _lib.load().mul_ct_matrix(self._cfg, _ptr(mx))
        return mx

    def st_dt_matrix(self):
        """Coefficients (2*nlm) of sin(theta) d/dtheta, for SH_mul_mx (shtns_numpy.i:277-281)."""
        mx = np.empty(2 * self.nlm)
        _lib.load().st_dt_matrix(self._cfg, _ptr(mx))
        return mx

    def SH_mul_mx(self, mx, Qlm):
        """R_lm = mx[2 lm] Q_{lm-1} + mx[2 lm + 1] Q_{lm+1} (shtns_numpy.i:651-659)."""
        q = np.ascontiguousarray(Qlm, dtype=np.complex128)
        m = np.ascontiguousarray(mx, dtype=np.float64)
        if q.size != self.nlm or m.size != 2 * self.nlm:
            raise RuntimeError("wrong array size.")
        r = np.empty_like(q)
        _lib.load().SH_mul_mx(self._cfg, _ptr(m), _ptr(q), _ptr(r))
        return r

    def Zrotate(self, Qlm, alpha):
        """Rotate the field about the z axis by alpha (shtns_numpy.i:609-614)."""
        q = np.ascontiguousarray(Qlm, dtype=np.complex128)
        if q.size != self.nlm:
            raise RuntimeError("spectral array has wrong size.")
        r = np.empty_like(q)
        _lib.load().SH_Zrotate(self._cfg, _ptr(q), float(alpha), _ptr(r))
        return r

    def _rotate(self, fn, Qlm, *args):
        if self.mres != 1 or self.mmax != self.lmax:
            raise RuntimeError("truncature makes rotation not closed.")
        self._host_spec(1, Qlm)
        r = np.empty_like(Qlm)
        fn(self._cfg, _ptr(Qlm), *args, _ptr(r))
        return r

    def Yrotate(self, Qlm, alpha):
        """Rotate the field about the y axis by alpha (shtns_numpy.i:617-627); mres = 1 and
        mmax = lmax are required."""
        return self._rotate(_lib.load().SH_Yrotate, Qlm, float(alpha))

    def Yrotate90(self, Qlm):
        """Rotate the field by 90 degrees about the y axis (shtns_numpy.i:628-638)."""
        return self._rotate(_lib.load().SH_Yrotate90, Qlm)

    def Xrotate90(self, Qlm):
        """Rotate the field by 90 degrees about the x axis (shtns_numpy.i:639-649)."""
        return self._rotate(_lib.load().SH_Xrotate90, Qlm)

    def spat_to_SH_m(self, Vr, Qlm, im):
        """Legendre analysis at order m = im*mres, no FFT: Vr complex[nlat] -> Qlm complex[lmax+1-m]
        (shtns_numpy.i:662-666)."""
        self._need_grid()
        m = int(im) * self.mres
        if Vr.size != self.nlat or Qlm.size != self.lmax + 1 - m or Vr.dtype != np.complex128 \
                or Qlm.dtype != np.complex128 or not Vr.flags.c_contiguous or not Qlm.flags.c_contiguous:
            raise RuntimeError("wrong array size or type.")
        _lib.load().spat_to_SH_ml(self._cfg, int(im), _ptr(Vr), _ptr(Qlm), self.lmax)
        _lib.check_error("spat_to_SH_m")

    def SH_to_spat_m(self, Qlm, Vr, im):
        """Legendre synthesis at order m = im*mres, no FFT (shtns_numpy.i:667-671)."""
        self._need_grid()
        m = int(im) * self.mres
        if Vr.size != self.nlat or Qlm.size != self.lmax + 1 - m or Vr.dtype != np.complex128 \
                or Qlm.dtype != np.complex128 or not Vr.flags.c_contiguous or not Qlm.flags.c_contiguous:
            raise RuntimeError("wrong array size or type.")
        _lib.load().SH_to_spat_ml(self._cfg, int(im), _ptr(Qlm), _ptr(Vr), self.lmax)
        _lib.check_error("SH_to_spat_m")

    def _check_m_arrays(self, im, spat, spec):
        m = int(im) * self.mres
        for a, n in [(v, self.nlat) for v in spat] + [(q, self.lmax + 1 - m) for q in spec]:
            if a.size != n or a.dtype != np.complex128 or not a.flags.c_contiguous:
                raise RuntimeError("wrong array size or type.")

    def spat_to_SHsphtor_m(self, Vt, Vp, Slm, Tlm, im):
        """Vector Legendre analysis at order m = im*mres, no FFT (shtns_numpy.i:672-676)."""
        self._need_grid()
        self._check_m_arrays(im, (Vt, Vp), (Slm, Tlm))
        _lib.load().spat_to_SHsphtor_ml(self._cfg, int(im), _ptr(Vt), _ptr(Vp), _ptr(Slm), _ptr(Tlm), self.lmax)
        _lib.check_error("spat_to_SHsphtor_m")

    def SHsphtor_to_spat_m(self, Slm, Tlm, Vt, Vp, im):
        """Vector Legendre synthesis at order m = im*mres, no FFT (shtns_numpy.i:677-681)."""
        self._need_grid()
        self._check_m_arrays(im, (Vt, Vp), (Slm, Tlm))
        _lib.load().SHsphtor_to_spat_ml(self._cfg, int(im), _ptr(Slm), _ptr(Tlm), _ptr(Vt), _ptr(Vp), self.lmax)
        _lib.check_error("SHsphtor_to_spat_m")

    def SH_to_point_cplx(self, alm, cost, phi):
        """Complex-valued field at one point from its (lmax+1)^2 coefficients (shtns_numpy.i:571-580)."""
        a = np.ascontiguousarray(alm, dtype=np.complex128)
        if a.size != self.nlm_cplx:
            raise RuntimeError("spectral array has wrong size.")
        r = _lib.load().SH_to_point_cplx(self._cfg, _ptr(a), float(cost), float(phi))
        return complex(r.re, r.im)

    def _point3(self, fn, arrays, cost, phi):
        arr = []
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.complex128)
            if a.size != self.nlm:
                raise RuntimeError("spectral array has wrong size.")
            arr.append(a)
        out = [ctypes.c_double() for _ in range(3)]
        fn(self._cfg, *[_ptr(a) for a in arr], float(cost), float(phi), *[ctypes.byref(o) for o in out])
        return tuple(o.value for o in out)

    def SH_to_grad_point(self, DrSlm, Slm, cost, phi):
        """(gr, gt, gp) = (sum DrSlm Y, dS/dtheta, (1/sin) dS/dphi) at one point (shtns_numpy.i:582-586)."""
        return self._point3(_lib.load().SH_to_grad_point, (DrSlm, Slm), cost, phi)

    def SHqst_to_point(self, Qlm, Slm, Tlm, cost, phi):
        """(vr, vt, vp) of the 3-component field at one point (shtns_numpy.i:587-592)."""
        return self._point3(_lib.load().SHqst_to_point, (Qlm, Slm, Tlm), cost, phi)

    def SH_to_lat(self, Qlm, cost, Vr):
        """Fill Vr (float64, nphi = its size) with the field on the latitude cos(theta) = cost at
        equispaced longitudes 2 pi j / nphi (shtns_numpy.i:597-600)."""
        q = np.ascontiguousarray(Qlm, dtype=np.complex128)
        if q.size != self.nlm or Vr.dtype != np.float64 or not Vr.flags.c_contiguous:
            raise RuntimeError("wrong array size or type.")
        _lib.load().SH_to_lat(self._cfg, _ptr(q), float(cost), _ptr(Vr), int(Vr.size), self.lmax, self.mmax)

    def SHqst_to_lat(self, Qlm, Slm, Tlm, cost, Vr, Vt, Vp):
        """3-component field on one latitude (shtns_numpy.i:601-606)."""
        q = [np.ascontiguousarray(a, dtype=np.complex128) for a in (Qlm, Slm, Tlm)]
        n = Vr.size
        if any(a.size != self.nlm for a in q) or any(v.size != n or v.dtype != np.float64 or not v.flags.c_contiguous
                                                     for v in (Vr, Vt, Vp)):
            raise RuntimeError("wrong array size or type.")
        _lib.load().SHqst_to_lat(self._cfg, *[_ptr(a) for a in q], float(cost), _ptr(Vr), _ptr(Vt), _ptr(Vp), int(n),
                                 self.lmax, self.mmax)

    # ---- checks ----
    def _need_grid(self):
        if self.nlat == 0:
            raise RuntimeError("Grid not set. Call .set_grid() method.")

    def _check_spec(self, q):
        if q.shape[-1] != self.nlm:
            raise RuntimeError("spectral array has wrong size.")

    def _check_spat(self, v):
        if tuple(v.shape[-2:]) != tuple(self.spat_shape):
            raise RuntimeError("spatial array has wrong shape.")

    # host arrays handed to the C API (shtns_numpy.i:60-90 check_spatial / check_spectral):
    # numpy, float64 / complex128, C-contiguous, exactly nspat / nlm (or nlm_cplx) elements
    def _host_spat(self, i, a, dtype=np.float64):
        self._need_grid()
        if not isinstance(a, np.ndarray):
            raise TypeError(f"arg #{i} : numpy array expected.")
        if a.dtype != dtype:
            raise TypeError(f"arg #{i} : spatial array must consist of {np.dtype(dtype).name}.")
        if not a.flags.c_contiguous:
            raise RuntimeError(f"arg #{i} : spatial array not contiguous. Use 'b=a.copy()' to copy a to a contiguous array b.")
        if a.size != self.nspat:
            raise RuntimeError(f"arg #{i} : spatial array has wrong size")

    def _host_spec(self, i, a, size=None):
        if not isinstance(a, np.ndarray):
            raise RuntimeError(f"arg #{i} : numpy array expected.")
        if a.dtype != np.complex128:
            raise RuntimeError(f"arg #{i} : spectral array must consist of complex float. Create with: 'sh.spec_array()'")
        if not a.flags.c_contiguous:
            raise RuntimeError(f"arg #{i} : spectral array not contiguous. Use .copy() to copy to a contiguous array.")
        if a.size != (self.nlm if size is None else size):
            raise RuntimeError(f"arg #{i} : spectral array has wrong size")

    def _host_args(self, spec=(), spat=(), cplx=False):
        """Validate the spectral then spatial host arrays of one call (argument numbers as in
        the call)."""
        i = 1
        for a in spec:
            self._host_spec(i, a, self.nlm_cplx if cplx else None)
            i += 1
        for a in spat:
            self._host_spat(i, a, np.complex128 if cplx else np.float64)
            i += 1

    # ---- numpy (host) transforms, SHTns names ----
    def SH_to_spat(self, Qlm, Vr, ltr=-1):
        self._need_grid()
        self._host_args(spec=(Qlm,), spat=(Vr,))
        lib = _lib.load()
        lib.SH_to_spat_l(self._cfg, _ptr(Qlm), _ptr(Vr), self.lmax if ltr < 0 else ltr)
        _lib.check_error("SH_to_spat")

    def spat_to_SH(self, Vr, Qlm, ltr=-1):
        self._need_grid()
        self._host_args(spec=(Qlm,), spat=(Vr,))
        lib = _lib.load()
        lib.spat_to_SH_l(self._cfg, _ptr(Vr), _ptr(Qlm), self.lmax if ltr < 0 else ltr)
        _lib.check_error("spat_to_SH")

    def SHsphtor_to_spat(self, Slm, Tlm, Vt, Vp, ltr=-1):
        self._need_grid()
        self._host_args(spec=(Slm, Tlm), spat=(Vt, Vp))
        _lib.load().SHsphtor_to_spat_l(self._cfg, _ptr(Slm), _ptr(Tlm), _ptr(Vt), _ptr(Vp), self.lmax if ltr < 0 else ltr)
        _lib.check_error("SHsphtor_to_spat")

    def spat_to_SHsphtor(self, Vt, Vp, Slm, Tlm, ltr=-1):
        self._need_grid()
        self._host_args(spec=(Slm, Tlm), spat=(Vt, Vp))
        _lib.load().spat_to_SHsphtor_l(self._cfg, _ptr(Vt), _ptr(Vp), _ptr(Slm), _ptr(Tlm), self.lmax if ltr < 0 else ltr)
        _lib.check_error("spat_to_SHsphtor")

    def SHqst_to_spat(self, Qlm, Slm, Tlm, Vr, Vt, Vp, ltr=-1):
        """3-component synthesis: radial Qlm -> Vr, tangential Slm, Tlm -> Vt, Vp
        (shtns_numpy.i:376-380)."""
        self._need_grid()
        self._host_args(spec=(Qlm, Slm, Tlm), spat=(Vr, Vt, Vp))
        _lib.load().SHqst_to_spat_l(self._cfg, _ptr(Qlm), _ptr(Slm), _ptr(Tlm), _ptr(Vr), _ptr(Vt), _ptr(Vp),
                                    self.lmax if ltr < 0 else ltr)
        _lib.check_error("SHqst_to_spat")

    def spat_to_SHqst(self, Vr, Vt, Vp, Qlm, Slm, Tlm, ltr=-1):
        """3-component analysis (shtns_numpy.i:371-375)."""
        self._need_grid()
        self._host_args(spec=(Qlm, Slm, Tlm), spat=(Vr, Vt, Vp))
        _lib.load().spat_to_SHqst_l(self._cfg, _ptr(Vr), _ptr(Vt), _ptr(Vp), _ptr(Qlm), _ptr(Slm), _ptr(Tlm),
                                    self.lmax if ltr < 0 else ltr)
        _lib.check_error("spat_to_SHqst")

    def SH_to_spat_cplx(self, alm, z):
        """Complex-valued field: alm[LM_cplx(l, m)], -l <= m <= l -> z (shtns_numpy.i:333-337)."""
        self._need_grid()
        self._host_args(spec=(alm,), spat=(z,), cplx=True)
        _lib.load().SH_to_spat_cplx(self._cfg, _ptr(alm), _ptr(z))
        _lib.check_error("SH_to_spat_cplx")

    def spat_cplx_to_SH(self, z, alm):
        """Complex-valued field z -> alm[LM_cplx(l, m)] (shtns_numpy.i:328-332)."""
        self._need_grid()
        self._host_args(spec=(alm,), spat=(z,), cplx=True)
        _lib.load().spat_cplx_to_SH(self._cfg, _ptr(z), _ptr(alm))
        _lib.check_error("spat_cplx_to_SH")

    def zidx(self, l, m):
        """Index of (l, m), -l <= m <= l, in a complex field's spectrum (LM_cplx, shtns.h:114)."""
        if l > self.lmax or abs(m) > min(l, self.mmax):
            raise ValueError("l or m invalid")
        return l * (l + 1) + m if l <= self.mmax else self.mmax * (2 * l - self.mmax) + l + m

    def SHsphtor_to_spat_cplx(self, Slm, Tlm, Vt, Vp):
        """Complex-valued 2D vector field from Slm, Tlm in the LM_cplx layout (shtns_numpy.i:365-369)."""
        self._need_grid()
        self._host_args(spec=(Slm, Tlm), spat=(Vt, Vp), cplx=True)
        _lib.load().SHsphtor_to_spat_cplx(self._cfg, _ptr(Slm), _ptr(Tlm), _ptr(Vt), _ptr(Vp))
        _lib.check_error("SHsphtor_to_spat_cplx")

    def spat_cplx_to_SHsphtor(self, Vt, Vp, Slm, Tlm):
        """Complex-valued 2D vector field -> Slm, Tlm (shtns_numpy.i:360-364)."""
        self._need_grid()
        self._host_args(spec=(Slm, Tlm), spat=(Vt, Vp), cplx=True)
        _lib.load().spat_cplx_to_SHsphtor(self._cfg, _ptr(Vt), _ptr(Vp), _ptr(Slm), _ptr(Tlm))
        _lib.check_error("spat_cplx_to_SHsphtor")

    def _need_cplx(self, n):
        self._need_grid()
        if self.lmax != self.mmax or self.mres != 1:
            raise RuntimeError("complex SH requires lmax=mmax and mres=1.")
        if n not in (1, 2, 3):
            raise RuntimeError("1, 2 or 3 arrays required.")

    def synth_cplx(self, *arg):
        """z = synth_cplx(zlm) ; zt, zp = synth_cplx(slm, tlm) ; zr, zt, zp = synth_cplx(qlm, slm, tlm):
        complex-valued fields from their (lmax+1)^2 coefficients per array
        (shtns_numpy.i:469-500).  numpy arrays use the host API; torch complex128 tensors on
        the GPU stay on the device and may carry leading batch dimensions (one batched call)."""
        n = len(arg)
        self._need_cplx(n)
        if _is_torch(arg[0]):
            import torch
            q = []
            for a in arg:
                if not _is_torch(a) or not a.is_cuda or a.dtype != torch.complex128:
                    raise RuntimeError("spectral tensors should be torch.complex128 on the GPU")
                if a.shape[-1] != self.nlm_cplx or a.shape != arg[0].shape:
                    raise RuntimeError("spectral array has wrong size.")
                q.append(a.contiguous())
            batch = tuple(q[0].shape[:-1])           # leading batch dimensions, one launch sequence
            nf = int(np.prod(batch)) if batch else 1
            out = [torch.empty(batch + tuple(self.spat_shape), dtype=torch.complex128, device=q[0].device) for _ in q]
            if nf == 0:
                return out[0] if n == 1 else tuple(out)
            self._stream(q[0])
            lib = _lib.load()
            if n != 2:
                lib.cu_SH_to_spat_cplx_batch(self._cfg, _ptr(q[0]), _ptr(out[0]), nf, self.nlm_cplx, self.nspat)
                _lib.check_error("synth_cplx")
            if n >= 2:
                lib.cu_SHsphtor_to_spat_cplx_batch(self._cfg, _ptr(q[-2]), _ptr(q[-1]), _ptr(out[-2]), _ptr(out[-1]),
                                                   nf, self.nlm_cplx, self.nspat)
                _lib.check_error("synth_cplx")
            return out[0] if n == 1 else tuple(out)
        for a in arg:
            if len(a.shape) != 1 or a.shape[0] != self.nlm_cplx:
                raise RuntimeError("spectral array has wrong size.")
        q = []
        for a in arg:
            if a.dtype != np.complex128:
                raise RuntimeError("spectral array should be dtype=complex.")
            q.append(np.ascontiguousarray(a))
        out = [np.empty(self.spat_shape, dtype=np.complex128) for _ in q]
        if n != 2:
            self.SH_to_spat_cplx(q[0], out[0])
        if n >= 2:
            self.SHsphtor_to_spat_cplx(q[-2], q[-1], out[-2], out[-1])
        return out[0] if n == 1 else tuple(out)

    def analys_cplx(self, *arg):
        """zlm = analys_cplx(z) ; slm, tlm = analys_cplx(zt, zp) ; qlm, slm, tlm = analys_cplx(zr, zt, zp)
        (shtns_numpy.i:502-535)."""
        n = len(arg)
        self._need_cplx(n)
        if _is_torch(arg[0]):
            import torch
            v = []
            for z in arg:
                if not _is_torch(z) or not z.is_cuda or z.dtype != torch.complex128:
                    raise RuntimeError("spatial tensors should be torch.complex128 on the GPU")
                if tuple(z.shape[-2:]) != tuple(self.spat_shape) or z.shape != arg[0].shape:
                    raise RuntimeError("spatial array has wrong shape.")
                v.append(z.contiguous())
            batch = tuple(v[0].shape[:-2])
            nf = int(np.prod(batch)) if batch else 1
            out = [torch.empty(batch + (self.nlm_cplx,), dtype=torch.complex128, device=v[0].device) for _ in v]
            if nf == 0:
                return out[0] if n == 1 else tuple(out)
            self._stream(v[0])
            lib = _lib.load()
            if n != 2:
                lib.cu_spat_cplx_to_SH_batch(self._cfg, _ptr(v[0]), _ptr(out[0]), nf, self.nspat, self.nlm_cplx)
                _lib.check_error("analys_cplx")
            if n >= 2:
                lib.cu_spat_cplx_to_SHsphtor_batch(self._cfg, _ptr(v[-2]), _ptr(v[-1]), _ptr(out[-2]), _ptr(out[-1]),
                                                   nf, self.nspat, self.nlm_cplx)
                _lib.check_error("analys_cplx")
            return out[0] if n == 1 else tuple(out)
        for z in arg:
            if tuple(z.shape) != tuple(self.spat_shape):
                raise RuntimeError("spatial array has wrong shape.")
        v = []
        for z in arg:
            if z.dtype != np.complex128:
                raise RuntimeError("spatial array should be dtype=complex128.")
            v.append(np.ascontiguousarray(z))
        out = [np.empty(self.nlm_cplx, dtype=np.complex128) for _ in v]
        if n != 2:
            self.spat_cplx_to_SH(v[0], out[0])
        if n >= 2:
            self.spat_cplx_to_SHsphtor(v[-2], v[-1], out[-2], out[-1])
        return out[0] if n == 1 else tuple(out)

    # ---- generic entry points (numpy or torch) ----
    def synth(self, *arg, ltr=-1):
        """v = synth(qlm) ; vt, vp = synth(slm, tlm) ; vr, vt, vp = synth(qlm, slm, tlm)
        (shtns_numpy.i:396-426)."""
        self._need_grid()
        n = len(arg)
        if n not in (1, 2, 3):
            raise RuntimeError("1, 2 or 3 spectral arrays required.")
        if _is_torch(arg[0]):
            self._same_tensors(arg)
            if n == 3:   # radial scalar, then the tangential vector pair (one stream)
                return (self._synth_torch(arg[:1], ltr),) + tuple(self._synth_torch(arg[1:], ltr))
            return self._synth_torch(arg, ltr)
        q = []
        for a in arg:
            self._check_spec(a)
            if a.dtype != np.complex128:
                raise RuntimeError("spectral array should be dtype=complex.")
            q.append(np.ascontiguousarray(a))
        if n == 1:
            vr = np.empty(self.spat_shape)
            self.SH_to_spat(q[0], vr, ltr)
            return vr
        vt, vp = np.empty(self.spat_shape), np.empty(self.spat_shape)
        if n == 3:
            vr = np.empty(self.spat_shape)
            self.SHqst_to_spat(q[0], q[1], q[2], vr, vt, vp, ltr)
            return vr, vt, vp
        self.SHsphtor_to_spat(q[0], q[1], vt, vp, ltr)
        return vt, vp

    def analys(self, *arg, ltr=-1):
        """qlm = analys(v) ; slm, tlm = analys(vt, vp) ; qlm, slm, tlm = analys(vr, vt, vp)
        (shtns_numpy.i:427-457)."""
        self._need_grid()
        n = len(arg)
        if n not in (1, 2, 3):
            raise RuntimeError("1, 2 or 3 spatial arrays required.")
        if _is_torch(arg[0]):
            self._same_tensors(arg)
            if n == 3:
                return (self._analys_torch(arg[:1], ltr),) + tuple(self._analys_torch(arg[1:], ltr))
            return self._analys_torch(arg, ltr)
        v = []
        for a in arg:
            self._check_spat(a)
            v.append(np.ascontiguousarray(a, dtype=np.float64))
        if n == 1:
            q = self.spec_array()
            self.spat_to_SH(v[0], q, ltr)
            return q
        s, t = self.spec_array(), self.spec_array()
        if n == 3:
            q = self.spec_array()
            self.spat_to_SHqst(v[0], v[1], v[2], q, s, t, ltr)
            return q, s, t
        self.spat_to_SHsphtor(v[0], v[1], s, t, ltr)
        return s, t

    # ---- torch (device) transforms ----
    @staticmethod
    def _same_tensors(arg):
        """Every tensor of one call must be a torch tensor of the first one's shape, dtype and
        device: the batch size and strides handed to the kernels come from the first."""
        a0 = arg[0]
        for i, a in enumerate(arg[1:], start=2):
            if not _is_torch(a) or a.shape != a0.shape or a.dtype != a0.dtype or a.device != a0.device:
                raise RuntimeError(f"arg #{i} : every tensor of a call must have the shape, dtype and device of the "
                                   f"first ({tuple(a0.shape)}, {a0.dtype}, {a0.device})")

    def _stream(self, t):
        """Run this configuration's next kernels on torch's current stream.  All calls on one
        configuration share its device scratch, so when the stream changes the new stream
        first waits for the work queued on the previous one (no overlap, no race)."""
        import torch
        cur = torch.cuda.current_stream(t.device)
        prev = getattr(self, "_last_stream", None)
        if prev is not None and prev != cur:
            cur.wait_stream(prev)
        self._last_stream = cur
        s = cur.cuda_stream
        _lib.load().cushtns_set_streams(self._cfg, ctypes.c_void_p(s), ctypes.c_void_p(s))

    def _torch_spec(self, q, allow_f32=False):
        import torch
        if not q.is_cuda:
            raise RuntimeError("torch tensors must be on the GPU")
        if q.dtype != torch.complex128 and not (allow_f32 and q.dtype == torch.complex64):
            raise RuntimeError("spectral tensor should be torch.complex128"
                               + (" or torch.complex64" if allow_f32 else ""))
        self._check_spec(q)
        return q.contiguous()

    def _torch_spat(self, v, allow_f32=False):
        import torch
        if not v.is_cuda:
            raise RuntimeError("torch tensors must be on the GPU")
        if v.dtype != torch.float64 and not (allow_f32 and v.dtype == torch.float32):
            raise RuntimeError("spatial tensor should be torch.float64" + (" or torch.float32" if allow_f32 else ""))
        self._check_spat(v)
        return v.contiguous()

    def set_precision(self, prec):
        """Precision of the Legendre contraction on the GPU: SHT_PREC_F64, SHT_PREC_F32,
        SHT_PREC_SPLIT or SHT_PREC_AUTO (FP64 for double data, FP32 for float data)."""
        if _lib.load().cushtns_set_precision(self._cfg, int(prec)) != 0:
            _lib.check_error("set_precision")

    def _synth_torch(self, arg, ltr):
        import torch
        q = [self._torch_spec(a, allow_f32=(len(arg) == 1)) for a in arg]
        f32 = q[0].dtype == torch.complex64
        batch = tuple(q[0].shape[:-1])
        nf = int(np.prod(batch)) if batch else 1
        self._stream(q[0])
        lib = _lib.load()
        lt = self.lmax if ltr < 0 else ltr
        rdt = torch.float32 if f32 else torch.float64
        out = [torch.empty(batch + tuple(self.spat_shape), dtype=rdt, device=q[0].device) for _ in q]
        if nf > 0:
            if f32:
                lib.cu_SH_to_spat_float_batch(self._cfg, _ptr(q[0]), _ptr(out[0]), lt, nf, self.nlm, self.nspat)
            elif len(q) == 1:
                lib.cu_SH_to_spat_batch(self._cfg, _ptr(q[0]), _ptr(out[0]), lt, nf, self.nlm, self.nspat)
            else:
                lib.cu_SHsphtor_to_spat_batch(self._cfg, _ptr(q[0]), _ptr(q[1]), _ptr(out[0]), _ptr(out[1]), lt, nf,
                                              self.nlm, self.nspat)
            _lib.check_error("synth")
        return out[0] if len(q) == 1 else tuple(out)

    def _analys_torch(self, arg, ltr):
        import torch
        v = [self._torch_spat(a, allow_f32=(len(arg) == 1)) for a in arg]
        f32 = v[0].dtype == torch.float32
        batch = tuple(v[0].shape[:-2])
        nf = int(np.prod(batch)) if batch else 1
        self._stream(v[0])
        lib = _lib.load()
        lt = self.lmax if ltr < 0 else ltr
        cdt = torch.complex64 if f32 else torch.complex128
        out = [torch.empty(batch + (self.nlm,), dtype=cdt, device=v[0].device) for _ in v]
        if nf > 0:
            if f32:
                lib.cu_spat_to_SH_float_batch(self._cfg, _ptr(v[0]), _ptr(out[0]), lt, nf, self.nspat, self.nlm)
            elif len(v) == 1:
                lib.cu_spat_to_SH_batch(self._cfg, _ptr(v[0]), _ptr(out[0]), lt, nf, self.nspat, self.nlm)
            else:
                lib.cu_spat_to_SHsphtor_batch(self._cfg, _ptr(v[0]), _ptr(v[1]), _ptr(out[0]), _ptr(out[1]), lt, nf,
                                              self.nspat, self.nlm)
            _lib.check_error("analys")
        return out[0] if len(v) == 1 else tuple(out)

    # adjoint operators on torch tensors (used by shtns_amd.autograd)
    def synth_adjoint(self, v, ltr=-1):
        import torch
        v = self._torch_spat(v)
        batch = tuple(v.shape[:-2])
        nf = int(np.prod(batch)) if batch else 1
        self._stream(v)
        out = torch.empty(batch + (self.nlm,), dtype=torch.complex128, device=v.device)
        if nf > 0:
            _lib.load().cu_SH_to_spat_adjoint_batch(self._cfg, _ptr(v), _ptr(out), self.lmax if ltr < 0 else ltr, nf,
                                                    self.nspat, self.nlm)
            _lib.check_error("synth_adjoint")
        return out

    def analys_adjoint(self, q, ltr=-1):
        import torch
        q = self._torch_spec(q)
        batch = tuple(q.shape[:-1])
        nf = int(np.prod(batch)) if batch else 1
        self._stream(q)
        out = torch.empty(batch + tuple(self.spat_shape), dtype=torch.float64, device=q.device)
        if nf > 0:
            _lib.load().cu_spat_to_SH_adjoint_batch(self._cfg, _ptr(q), _ptr(out), self.lmax if ltr < 0 else ltr, nf,
                                                    self.nlm, self.nspat)
            _lib.check_error("analys_adjoint")
        return out


class rotation:
    """Rotation of spherical harmonic expansions (SHTns's ``shtns.rotation``,
    shtns_numpy.i:715-800).  A rotation maps the field f to f(R^-1 x); set_angles_ZYZ(a, b, g)
    rotates by a about z, then b about y, then g about z."""

    def __init__(self, lmax, mmax=-1, norm=0):
        if lmax < 2:
            raise ValueError("lmax < 2 not allowed")
        if mmax < 0:
            mmax = lmax
        if mmax > lmax:
            raise ValueError("lmax < mmax invalid")
        self.lmax, self.mmax = int(lmax), int(mmax)
        self._r = _lib.load().shtns_rotation_create(self.lmax, self.mmax, int(norm))
        if not self._r:
            _lib.check_error("rotation")
        self.nlm = nlm_calc(self.lmax, self.mmax, 1)
        self.nlm_cplx = int(_lib.load().nlm_cplx_calc(self.lmax, self.mmax, 1))

    def __del__(self):
        r = getattr(self, "_r", None)
        if r:
            try:
                _lib.load().shtns_rotation_destroy(r)
            except Exception:
                pass
            self._r = None

    def set_angles_ZYZ(self, alpha, beta, gamma):
        if abs(beta) > np.pi:
            raise ValueError("beta must be between -pi and pi")
        _lib.load().shtns_rotation_set_angles_ZYZ(self._r, float(alpha), float(beta), float(gamma))

    def set_angles_ZXZ(self, alpha, beta, gamma):
        if abs(beta) > np.pi:
            raise ValueError("beta must be between -pi and pi")
        _lib.load().shtns_rotation_set_angles_ZXZ(self._r, float(alpha), float(beta), float(gamma))

    def set_angle_axis(self, theta, Vx, Vy, Vz):
        _lib.load().shtns_rotation_set_angle_axis(self._r, float(theta), float(Vx), float(Vy), float(Vz))
        _lib.check_error("set_angle_axis")

    def wigner_d_matrix(self, l):
        """d^l_{m'm}(beta) as a (2l+1, 2l+1) array, rows m' = -l..l, columns m = -l..l."""
        if l < 0 or l > self.lmax:
            raise ValueError("l must be between 0 and lmax")
        mx = np.empty((2 * l + 1, 2 * l + 1))
        _lib.load().shtns_rotation_wigner_d_matrix(self._r, int(l), _ptr(mx))
        return mx

    def _apply(self, fn, Qlm, n):
        q = np.ascontiguousarray(Qlm, dtype=np.complex128)
        if q.size != n:
            raise RuntimeError("spectral array has wrong size")
        r = np.empty(n, dtype=np.complex128)
        fn(self._r, _ptr(q), _ptr(r))
        return r

    def apply_real(self, Qlm):
        """Rotated coefficients of a real field (nlm_calc(lmax, mmax, 1) values, LM order)."""
        return self._apply(_lib.load().shtns_rotation_apply_real, Qlm, self.nlm)

    def apply_cplx(self, Zlm):
        """Rotated coefficients of a complex field (LM_cplx layout)."""
        return self._apply(_lib.load().shtns_rotation_apply_cplx, Zlm, self.nlm_cplx)
