"""CPU: the sampled-oracle helpers (tests/sampled_oracle.py) agree with the oracle's full
transforms, so GPU tests at sizes beyond the full oracle can rely on them."""
import numpy as np

from oracle.oracle import Oracle
from shtns_amd.synthetic import spectral_field
from sampled_oracle import anal_orders, anal_orders_ld, spat_rows, synth_rows, synth_rows_ld


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def test_synth_rows_match_full_synthesis():
    for (L, M, mres, nlat, nphi) in [(63, 63, 1, 96, 192), (60, 15, 4, 64, 48)]:
        o = Oracle(L, M, mres, nlat, nphi)
        q = spectral_field(L, M, seed=3, mres=mres)
        lats = [0, 5, nlat // 2 - 1, nlat // 2, nlat - 1]
        G = synth_rows(o, q, lats)
        R = spat_rows(o.synth(q), lats)[:M + 1]
        R[0] = R[0].real
        assert rel(G, R) <= 1e-13


def test_anal_orders_match_full_analysis():
    for (L, M, mres, nlat, nphi) in [(63, 63, 1, 96, 192), (60, 15, 4, 64, 48)]:
        o = Oracle(L, M, mres, nlat, nphi)
        V = np.random.default_rng(4).standard_normal((nphi, nlat))
        ref = o.analys(V)
        got = anal_orders(o, V, [0, 1, M])
        for im, q in got.items():
            m = im * mres
            b = o.idx(m, m)
            assert rel(q, ref[b:b + L + 1 - m]) <= 1e-13


def test_extended_precision_references_match_oracle():
    L, nlat, nphi = 63, 96, 192
    o = Oracle(L, L, 1, nlat, nphi)
    q = spectral_field(L, L, seed=5)
    lats = [0, 7, 47, 48, 95]
    assert rel(synth_rows_ld(o, q, lats), synth_rows(o, q, lats)) <= 1e-14
    V = np.random.default_rng(6).standard_normal((nphi, nlat))
    ref = anal_orders(o, V, [0, 3, 63])
    got = anal_orders_ld(o, V, [0, 3, 63])
    for m in ref:
        assert rel(got[m], ref[m]) <= 1e-14
