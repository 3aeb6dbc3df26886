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
