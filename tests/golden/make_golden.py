"""Writes tests/golden/legendre_golden.json: normalized associated Legendre functions
Y_l^m(x) = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x), with the Condon-Shortley phase.

For l <= 255 the values come from the explicit polynomial
P_l^m(x) = (-1)^m (1-x^2)^(m/2) d^m/dx^m P_l(x),  P_l(x) = 2^-l sum_j (-1)^j C(l,j) C(2l-2j,l) x^(l-2j)
with exact integer coefficients, evaluated in 500-digit arithmetic -- independent of any
recurrence.  For larger l (up to 4095) the
hypergeometric series needs thousands of digits, so the values come from the three-term
recurrence carried out in 80-digit arithmetic (no rounding or exponent-range effects at
double-precision level); the script checks the two methods agree to 1e-40 wherever both
are evaluated (l <= 255).  Points cover orders up to l and latitudes from the equator to
very near the pole, where sin(theta)^m underflows double precision.

Run:  python tests/golden/make_golden.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 60
HERE = os.path.dirname(os.path.abspath(__file__))


def ylm_hyp(l, m, x):
    from math import comb, factorial
    with mp.workdps(500):
        x = mp.mpf(float(x))
        acc = mp.mpf(0)
        for j in range(l // 2 + 1):
            n = l - 2 * j
            if n < m:
                continue
            c = (-1) ** j * comb(l, j) * comb(2 * l - 2 * j, l) * (factorial(n) // factorial(n - m))
            acc += c * x ** (n - m)
        p = acc / mp.mpf(2) ** l * mp.sqrt((1 - x) * (1 + x)) ** m * (-1) ** m
        norm = mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.factorial(l - m) / mp.factorial(l + m))
        return +(norm * p)


def ylm_rec(l, m, x):
    with mp.workdps(80):
        x = mp.mpf(float(x))   # the exact double the tests feed in
        s = mp.sqrt((1 - x) * (1 + x))
        t = mp.mpf(1) / (4 * mp.pi)
        for k in range(1, m + 1):
            t *= mp.mpf(2 * k + 1) / (2 * k)
        y0 = mp.sqrt(t) * s ** m * (-1 if m & 1 else 1)
        if l == m:
            return +y0
        y1 = mp.sqrt(2 * m + 3) * x * y0
        for ll in range(m + 2, l + 1):
            a = mp.sqrt(mp.mpf((2 * ll + 1) * (2 * ll - 1)) / ((ll + m) * (ll - m)))
            b = -mp.sqrt(mp.mpf((2 * ll + 1) * (ll - 1 + m) * (ll - 1 - m)) / ((2 * ll - 3) * (ll + m) * (ll - m)))
            y0, y1 = y1, a * x * y1 + b * y0
        return +y1


def ylm(l, m, x):
    r = ylm_rec(l, m, x)
    if l <= 255:
        h = ylm_hyp(l, m, x)
        assert abs(h - r) <= mp.mpf("1e-40") * max(abs(h), mp.mpf("1e-300")), (l, m, x, h, r)
    return r


def main():
    pts = []
    xs = ["0.0", "0.123456789", "-0.5", "0.7071067811865476", "0.95", "0.999", "0.99999", "0.9999999"]
    ls = [0, 1, 2, 3, 10, 31, 100, 127, 255, 511, 1023, 2047, 4095]
    for l in ls:
        ms = sorted(set([0, 1, 2, l // 7, l // 3, l // 2, (3 * l) // 4, l - 1, l]) - {-1})
        for m in ms:
            if m < 0 or m > l:
                continue
            for x in xs:
                v = ylm(l, m, x)
                pts.append({"l": l, "m": m, "x": x, "value": mp.nstr(v, 25, min_fixed=-1, max_fixed=-1)})
    with open(os.path.join(HERE, "legendre_golden.json"), "w") as f:
        json.dump({"description": "Y_l^m(x) orthonormal with Condon-Shortley phase; l<=255 explicit polynomial (500 digits), larger l 80-digit recurrence",
                   "points": pts}, f, indent=0)
    print(len(pts), "points")


if __name__ == "__main__":
    main()
