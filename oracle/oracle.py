"""ctypes binding of the CPU oracle (oracle/sht_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package
``shtns_amd`` never imports it.

Arrays follow SHTns's layouts (shtns.h:104-106, doc/main.dox:174-189):
spectral ``complex128[nlm]`` in LM order, spatial ``float64[nphi, nlat]``
(theta-contiguous, north first).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        lib.orc_create.restype = P
        lib.orc_create.argtypes = [I, I, I, I, I]
        lib.orc_create_grid.restype = P
        lib.orc_create_grid.argtypes = [I, I, I, I, I, I]
        lib.orc_destroy.argtypes = [P]
        lib.orc_nlm.restype = ctypes.c_long
        lib.orc_nlm.argtypes = [P]
        lib.orc_set_threads.argtypes = [I]
        lib.orc_grid.argtypes = [P, P, P, P]
        lib.orc_legendre.argtypes = [P, I, D, D, I, P]
        lib.orc_SH_to_spat.argtypes = [P, P, P, I]
        lib.orc_spat_to_SH.argtypes = [P, P, P, I]
        lib.orc_SHsphtor_to_spat.argtypes = [P, P, P, P, P, I]
        lib.orc_spat_to_SHsphtor.argtypes = [P, P, P, P, P, I]
        lib.orc_fft.argtypes = [P, I, I]
        lib.orc_set_polar_opt.argtypes = [P, D]
        lib.orc_polar_tm.argtypes = [P, P]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    _load().orc_set_threads(int(n))


def fft(x, sign=-1):
    """Unnormalized complex DFT of a 1-D complex array (sign -1 forward)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    _load().orc_fft(_p(a), a.size, int(sign))
    return a


class Oracle:
    """One Gauss-grid configuration of the oracle."""

    GAUSS, REGULAR, REGULAR_POLES = 0, 1, 2

    def __init__(self, lmax, mmax=None, mres=1, nlat=None, nphi=None, grid=0):
        if mmax is None:
            mmax = lmax // mres
        if nlat is None:
            nlat = lmax + 1 + ((lmax + 1) & 1)
        if nphi is None:
            nphi = 2 * mmax + 2
        self.lmax, self.mmax, self.mres, self.nlat, self.nphi = lmax, mmax, mres, nlat, nphi
        lib = _load()
        self._h = lib.orc_create_grid(lmax, mmax, mres, nlat, nphi, grid)
        if not self._h:
            raise ValueError("invalid sizes for oracle config")
        self.nlm = int(lib.orc_nlm(self._h))
        ct = np.empty(nlat)
        st = np.empty(nlat)
        w = np.empty(nlat)
        lib.orc_grid(self._h, _p(ct), _p(st), _p(w))
        self.cos_theta, self.sin_theta, self.gauss_w = ct, st, w

    def __del__(self):
        if getattr(self, "_h", None):
            _load().orc_destroy(self._h)
            self._h = None

    def set_polar_opt(self, eps):
        """SHTns's polar optimization threshold (sht_init.c:522-557); 0 disables it."""
        _load().orc_set_polar_opt(self._h, float(eps))

    def polar_tm(self):
        """First latitude index of each order im = 0..mmax used by the transforms."""
        tm = np.empty(self.mmax + 1, dtype=np.int32)
        _load().orc_polar_tm(self._h, _p(tm))
        return tm

    def idx(self, l, m):
        im = m // self.mres
        return (im * (2 * self.lmax + 2 - (im + 1) * self.mres)) // 2 + l

    def legendre(self, im, x, sint=None, ltop=None):
        if sint is None:
            sint = np.sqrt((1.0 - x) * (1.0 + x))
        m = im * self.mres
        if ltop is None:
            ltop = self.lmax
        y = np.empty(ltop - m + 1)
        _load().orc_legendre(self._h, im, float(x), float(sint), ltop, _p(y))
        return y

    def _spec(self, q):
        q = np.ascontiguousarray(q, dtype=np.complex128)
        assert q.shape == (self.nlm,)
        return q

    def _spat(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == (self.nphi, self.nlat)
        return v

    def synth(self, qlm, ltr=-1):
        q = self._spec(qlm)
        v = np.empty((self.nphi, self.nlat))
        _load().orc_SH_to_spat(self._h, _p(q), _p(v), int(ltr))
        return v

    def analys(self, vr, ltr=-1):
        v = self._spat(vr)
        q = np.empty(self.nlm, dtype=np.complex128)
        _load().orc_spat_to_SH(self._h, _p(v), _p(q), int(ltr))
        return q

    def synth_sphtor(self, slm, tlm, ltr=-1):
        s, t = self._spec(slm), self._spec(tlm)
        vt = np.empty((self.nphi, self.nlat))
        vp = np.empty((self.nphi, self.nlat))
        _load().orc_SHsphtor_to_spat(self._h, _p(s), _p(t), _p(vt), _p(vp), int(ltr))
        return vt, vp

    def analys_sphtor(self, vt, vp, ltr=-1):
        a, b = self._spat(vt), self._spat(vp)
        s = np.empty(self.nlm, dtype=np.complex128)
        t = np.empty(self.nlm, dtype=np.complex128)
        _load().orc_spat_to_SHsphtor(self._h, _p(a), _p(b), _p(s), _p(t), int(ltr))
        return s, t

    # ---- complex-valued fields (mres = 1) ----
    def lm_cplx(self, l, m):
        """Index of (l, m), -l <= m <= l, in the complex-field spectral layout (shtns.h:114)."""
        M = self.mmax
        return l * (l + 1) + m if l <= M else M * (2 * l - M) + l + m

    @property
    def nlm_cplx(self):
        return self.lm_cplx(self.lmax, min(self.lmax, self.mmax)) + 1

    def _plm(self, m):
        """P_lm(cos theta_k) for l = m..lmax at every grid latitude: float64[nlat, lmax-m+1]."""
        return np.array([self.legendre(m, float(x), float(s)) for x, s in zip(self.cos_theta, self.sin_theta)])

    def synth_cplx(self, alm):
        """SH_to_spat_cplx (sht_func.c:1000-1052).  For each order m >= 0, the Legendre sums of
        a_{l,m} go to Fourier bin m and those of (-1)^m a_{l,-m} to bin nphi - m; then a
        backward (e^{+i m phi}, unnormalized) complex DFT along phi.
        Returns complex128[nphi, nlat]."""
        assert self.mres == 1
        a = np.ascontiguousarray(alm, dtype=np.complex128)
        assert a.shape == (self.nlm_cplx,)
        G = np.zeros((self.nphi, self.nlat), dtype=np.complex128)
        for m in range(self.mmax + 1):
            P = self._plm(m)
            ls = range(m, self.lmax + 1)
            G[m] = P @ a[[self.lm_cplx(l, m) for l in ls]]
            if m > 0:
                G[self.nphi - m] = (-1) ** m * (P @ a[[self.lm_cplx(l, -m) for l in ls]])
        return np.fft.ifft(G, axis=0) * self.nphi

    def analys_cplx(self, z):
        """spat_cplx_to_SH (sht_func.c:943-995).  DFT along phi (bin m = sum_j z_j e^{-i m phi_j}),
        then per order the Gauss-weighted Legendre projections, scaled by 2 pi / nphi: bin m
        gives a_{l,m}, bin nphi - m gives (-1)^m a_{l,-m}."""
        assert self.mres == 1
        z = np.ascontiguousarray(z, dtype=np.complex128)
        assert z.shape == (self.nphi, self.nlat)
        Z = np.fft.fft(z, axis=0) * (2 * np.pi / self.nphi)
        w = self.gauss_w
        a = np.zeros(self.nlm_cplx, dtype=np.complex128)
        for m in range(self.mmax + 1):
            P = self._plm(m)
            ls = list(range(m, self.lmax + 1))
            a[[self.lm_cplx(l, m) for l in ls]] = P.T @ (w * Z[m])
            if m > 0:
                a[[self.lm_cplx(l, -m) for l in ls]] = (-1) ** m * (P.T @ (w * Z[self.nphi - m]))
        return a

    def _plm_dtheta(self, m):
        """P_lm and dP_lm/dtheta at every latitude, l = m..lmax (two float64[nlat, lmax-m+1]):
        sin(theta) dP_l/dtheta = l cos(theta) P_l - sqrt((2l+1)/(2l-1) (l^2 - m^2)) P_{l-1}."""
        P = self._plm(m)
        ls = np.arange(m, self.lmax + 1, dtype=np.float64)
        Pm1 = np.zeros_like(P)
        Pm1[:, 1:] = P[:, :-1]
        c = np.sqrt(np.abs((2 * ls + 1) / (2 * ls - 1) * (ls * ls - m * m)))
        dP = (ls * self.cos_theta[:, None] * P - c * Pm1) / self.sin_theta[:, None]
        return P, dP

    def synth_sphtor_cplx(self, slm, tlm):
        """SHsphtor_to_spat_cplx (sht_func.c:1060-1126): per order, bins m and nphi - m from
        orders +m and -m (coefficients times (-1)^m for -m), then a backward complex DFT.
        Vt = dS/dtheta + (1/sin) dT/dphi, Vp = (1/sin) dS/dphi - dT/dtheta, d/dphi = i m."""
        assert self.mres == 1
        s = np.ascontiguousarray(slm, dtype=np.complex128)
        t = np.ascontiguousarray(tlm, dtype=np.complex128)
        Gt = np.zeros((self.nphi, self.nlat), dtype=np.complex128)
        Gp = np.zeros_like(Gt)
        for m in range(self.mmax + 1):
            P, dP = self._plm_dtheta(m)
            ls = range(m, self.lmax + 1)
            for mm, row, sign in ([(0, 0, 1.0)] if m == 0 else [(m, m, 1.0), (-m, self.nphi - m, (-1.0) ** m)]):
                sl = sign * s[[self.lm_cplx(l, mm) for l in ls]]
                tl = sign * t[[self.lm_cplx(l, mm) for l in ls]]
                imP = (1j * mm) * P / self.sin_theta[:, None]
                Gt[row] = dP @ sl + imP @ tl
                Gp[row] = imP @ sl - dP @ tl
        return np.fft.ifft(Gt, axis=0) * self.nphi, np.fft.ifft(Gp, axis=0) * self.nphi

    def analys_sphtor_cplx(self, zt, zp):
        """spat_cplx_to_SHsphtor (sht_func.c:1133-1200): DFT along phi, then per order the
        Gauss-quadrature projections on grad Y and r x grad Y, divided by l(l+1); bin m gives
        order +m, bin nphi - m order -m (times (-1)^m).  S_00 = T_00 = 0."""
        assert self.mres == 1
        Zt = np.fft.fft(np.ascontiguousarray(zt, dtype=np.complex128), axis=0) * (2 * np.pi / self.nphi)
        Zp = np.fft.fft(np.ascontiguousarray(zp, dtype=np.complex128), axis=0) * (2 * np.pi / self.nphi)
        w = self.gauss_w
        s = np.zeros(self.nlm_cplx, dtype=np.complex128)
        t = np.zeros_like(s)
        for m in range(self.mmax + 1):
            P, dP = self._plm_dtheta(m)
            ls = np.arange(m, self.lmax + 1)
            inv = np.zeros(len(ls))
            inv[ls > 0] = 1.0 / (ls[ls > 0] * (ls[ls > 0] + 1.0))
            for mm, row, sign in ([(0, 0, 1.0)] if m == 0 else [(m, m, 1.0), (-m, self.nphi - m, (-1.0) ** m)]):
                a, b = w * Zt[row], w * Zp[row]
                imP = (-1j * mm) * P / self.sin_theta[:, None]     # (1/sin) d/dphi of conj(Y)
                idx = [self.lm_cplx(l, mm) for l in ls]
                s[idx] = sign * inv * (a @ dP + b @ imP)
                t[idx] = sign * inv * (a @ imP - b @ dP)
        return s, t


# ---- normalizations other than orthonormal (shtns.h:48-55) ----
NO_CS_PHASE, REAL_NORM = 256 * 4, 256 * 8


def norm_factors(lmax, mmax, mres, norm):
    """c_lm with Y^norm_lm = c_lm Y^orthonormal_lm, in LM order, following legendre_precomp
    (sht_legendre.c:586-605, 635-660): sht_fourpi (1) multiplies by sqrt(4 pi), sht_schmidt (2)
    and sht_for_rotations (3) by sqrt(4 pi / (2l+1)); SHT_NO_CS_PHASE drops the (-1)^m and
    SHT_REAL_NORM scales m > 0 by sqrt(1/2).  Pinned by tests/test_norms.py against the closed
    forms of these functions built from scipy's associated Legendre functions."""
    base = norm & 255
    c = []
    for im in range(mmax + 1):
        m = im * mres
        for l in range(m, lmax + 1):
            f = 1.0
            if base == 1:
                f = np.sqrt(4 * np.pi)
            elif base in (2, 3):
                f = np.sqrt(4 * np.pi / (2 * l + 1))
            if (norm & NO_CS_PHASE) and (m & 1):
                f = -f
            if (norm & REAL_NORM) and m > 0:
                f *= np.sqrt(0.5)
            c.append(f)
    return np.array(c)


class NormOracle(Oracle):
    """The oracle in another normalization: synthesis of Q is the orthonormal synthesis of
    c Q, analysis returns the orthonormal analysis divided by c (the inverse pair)."""

    def __init__(self, lmax, mmax=None, mres=1, nlat=None, nphi=None, grid=0, norm=0):
        super().__init__(lmax, mmax, mres, nlat, nphi, grid)
        self.norm = norm
        self.c = norm_factors(self.lmax, self.mmax, self.mres, norm)

    def synth(self, qlm, ltr=-1):
        return super().synth(self._spec(qlm) * self.c, ltr)

    def analys(self, vr, ltr=-1):
        return super().analys(vr, ltr) / self.c

    def synth_sphtor(self, slm, tlm, ltr=-1):
        return super().synth_sphtor(self._spec(slm) * self.c, self._spec(tlm) * self.c, ltr)

    def analys_sphtor(self, vt, vp, ltr=-1):
        s, t = super().analys_sphtor(vt, vp, ltr)
        return s / self.c, t / self.c
