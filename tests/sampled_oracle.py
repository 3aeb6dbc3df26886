"""Oracle values at sampled latitudes and orders, for grids too large for the oracle's full
transforms in a test (lmax = 4095 on 6144 x 12288).  Test infrastructure only.

Real field f(theta_k, phi_j) = sum_l Q_l0 Y_l0 + 2 Re sum_{m>0} Q_lm Y_lm e^{i m phi_j}
(doc/main.dox:263-303), phi_j = 2 pi j / nphi.  Its DFT over phi, divided by nphi, holds
G_m(theta_k) = sum_l Q_lm Y_lm(cos theta_k) in bin m (the real part at m = 0), and the
analysis is Q_lm = (2 pi / nphi) sum_k w_k Y_lm(cos theta_k) DFT_m(f)(theta_k) (Gauss weights
w_k summing to 2; real part at m = 0).  Both are pinned against the oracle's full transforms
in tests/test_sampled_oracle.py."""
import numpy as np


def spat_rows(V, lats):
    """DFT over phi / nphi of a theta-contiguous field V[nphi, nlat] at latitudes `lats`:
    complex [nphi, len(lats)]."""
    V = np.asarray(V)
    return np.fft.fft(V[:, lats], axis=0) / V.shape[0]


def synth_rows(o, q, lats):
    """G[m, k] = sum_l Q_lm Y_lm(x_k) for every order m = im*mres and the given latitudes."""
    G = np.zeros((o.mmax + 1, len(lats)), dtype=np.complex128)
    for j, k in enumerate(lats):
        x, s = o.cos_theta[k], o.sin_theta[k]
        for im in range(o.mmax + 1):
            m = im * o.mres
            y = o.legendre(im, x, s)
            b = o.idx(m, m)
            g = np.dot(y, q[b:b + o.lmax + 1 - m])
            G[im, j] = g.real if m == 0 else g
    return G


def anal_orders(o, V, ims):
    """Q_lm (l = m..lmax) of the analysis of V[nphi, nlat] for the order indices `ims`."""
    V = np.asarray(V)
    nphi = V.shape[0]
    out = {}
    for im in ims:
        m = im * o.mres
        # DFT bin im over phi (phi spans 2 pi / mres), computed for this bin only
        F = np.exp(-2j * np.pi * im * np.arange(nphi) / nphi) @ V
        acc = np.zeros(o.lmax + 1 - m, dtype=np.complex128)
        for k in range(o.nlat):
            acc += o.gauss_w[k] * F[k] * o.legendre(im, o.cos_theta[k], o.sin_theta[k])
        acc *= 2.0 * np.pi / nphi
        out[im] = acc.real + 0j if m == 0 else acc
    return out


# ---- extended-precision references (x87 long double, 64-bit mantissa) -------------------
# The classic orthonormal recurrence (sht_legendre.c:646-657), run in long double: at degree
# 4095 near the poles an FP64 recurrence (the oracle's, the GPU's) drifts by up to ~1e-10
# relative; these references drift ~2000x less.  Pinned against the oracle at low degree in
# tests/test_sampled_oracle.py.
def _plm_ld_all_m(L, ct, st):
    """Generator over l = 0..L of (orders 0..min(l, L), P_lm(ct)) in long double."""
    ld = np.longdouble
    ms = np.arange(L + 1)
    x, s = ld(ct), ld(st)
    k = np.arange(1, L + 1, dtype=ld)
    lg = np.concatenate([[ld(0)], np.cumsum(np.log2((k + ld(0.5)) / k))])
    lseed = 0.5 * (lg + np.log2(ld(0.25) / ld(np.pi))) + ms * np.log2(s) if s > 0 else None
    if lseed is None:
        seed = np.zeros(L + 1, dtype=ld)
        seed[0] = np.sqrt(ld(0.25) / ld(np.pi))
    else:
        seed = np.exp2(np.maximum(lseed, ld(-16000)))
        seed = np.where(lseed > -16000, seed, ld(0))
    seed = seed * np.where(ms % 2 == 1, -1, 1)
    ym2 = np.zeros(L + 1, dtype=ld)
    ym1 = seed.copy()
    yield 0, ym1[:1]
    for l in range(1, L + 1):
        m = ms[:l].astype(ld)
        ll = ld(l)
        a = np.sqrt((4 * ll * ll - 1) / (ll * ll - m * m))
        b = -np.sqrt(((2 * ll + 1) * ((ll - 1) ** 2 - m * m)) / ((2 * ll - 3) * (ll * ll - m * m)))
        b = np.where(l - 1 > ms[:l], b, 0)
        yn = a * x * ym1[:l] + b * ym2[:l]
        ym2[:l] = ym1[:l]
        ym1[:l] = yn
        ym1[l] = seed[l]
        yield l, ym1[:l + 1]


def synth_rows_ld(o, q, lats):
    """synth_rows in extended precision (mmax = lmax, mres = 1)."""
    L = o.lmax
    ms = np.arange(L + 1)
    base = (ms * (2 * L + 1 - ms)) // 2
    qr, qi = q.real.astype(np.longdouble), q.imag.astype(np.longdouble)
    G = np.zeros((L + 1, len(lats)), dtype=np.complex128)
    for j, k in enumerate(lats):
        ar = np.zeros(L + 1, dtype=np.longdouble)
        ai = np.zeros(L + 1, dtype=np.longdouble)
        for l, y in _plm_ld_all_m(L, o.cos_theta[k], o.sin_theta[k]):
            idx = base[:l + 1] + l
            ar[:l + 1] += y * qr[idx]
            ai[:l + 1] += y * qi[idx]
        G[:, j] = ar.astype(np.float64) + 1j * ai.astype(np.float64)
        G[0, j] = G[0, j].real
    return G


def anal_orders_ld(o, V, ims):
    """anal_orders in extended precision: P_lm at every latitude from the recurrence over l
    run for one order, all latitudes at once (mres = 1)."""
    V = np.asarray(V)
    nphi = V.shape[0]
    ld = np.longdouble
    x, s = o.cos_theta.astype(ld), o.sin_theta.astype(ld)
    w = o.gauss_w.astype(ld)
    out = {}
    for m in ims:
        F = np.exp(-2j * np.pi * m * np.arange(nphi) / nphi) @ V
        Fr, Fi = F.real.astype(ld) * w, F.imag.astype(ld) * w
        k = np.arange(1, m + 1, dtype=ld)
        lseed = 0.5 * (np.sum(np.log2((k + ld(0.5)) / k)) + np.log2(ld(0.25) / ld(np.pi))) + m * np.log2(s)
        y1 = np.where(lseed > -16000, np.exp2(np.maximum(lseed, ld(-16000))), ld(0)) * (-1 if m % 2 else 1)
        y0 = np.zeros_like(y1)
        acc = np.zeros(o.lmax + 1 - m, dtype=np.complex128)
        acc[0] = complex(np.sum(y1 * Fr), np.sum(y1 * Fi))
        for l in range(m + 1, o.lmax + 1):
            ll, mm = ld(l), ld(m)
            a = np.sqrt((4 * ll * ll - 1) / (ll * ll - mm * mm))
            b = -np.sqrt(((2 * ll + 1) * ((ll - 1) ** 2 - mm * mm)) / ((2 * ll - 3) * (ll * ll - mm * mm))) if l - 1 > m else ld(0)
            yn = a * x * y1 + b * y0
            y0, y1 = y1, yn
            acc[l - m] = complex(np.sum(y1 * Fr), np.sum(y1 * Fi))
        acc *= 2.0 * np.pi / nphi
        out[m] = acc.real + 0j if m == 0 else acc
    return out
