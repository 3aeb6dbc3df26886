"""Seeded synthetic inputs: SplitMix64 counter stream -> Box-Muller normals.

A spectral field has complex coefficients with Re and Im ~ N(0, sigma_l^2) independently,
sigma_l = (l+1)^(-power/2) (power 2 for scalar fields, 3 for spheroidal/toroidal
fields), and Im = 0 at m = 0.  Field f of a batch uses seed ``seed + f``.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    """The first n outputs of SplitMix64 started at state `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + _GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = s
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def normals(seed, n):
    """n standard normal doubles (Box-Muller on pairs of 53-bit uniforms in (0,1])."""
    k = (n + 1) // 2
    u = splitmix64(seed, 2 * k)
    u1 = ((u[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (u[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * k)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:n]


def lm_arrays(lmax, mmax, mres=1):
    ls, ms = [], []
    for im in range(mmax + 1):
        m = im * mres
        ls.append(np.arange(m, lmax + 1))
        ms.append(np.full(lmax + 1 - m, m))
    return np.concatenate(ls), np.concatenate(ms)


def spectral_field(lmax, mmax, seed, mres=1, power=2.0):
    """complex128[nlm] with Re, Im ~ N(0, (l+1)^-power) and Im(m=0) = 0."""
    l, m = lm_arrays(lmax, mmax, mres)
    g = normals(seed, 2 * l.size)
    sig = (l + 1.0) ** (-0.5 * power)
    q = (g[0::2] + 1j * g[1::2]) * sig
    q[m == 0] = q[m == 0].real
    return q


def spectral_batch(lmax, mmax, nfields, seed, mres=1, power=2.0):
    return np.stack([spectral_field(lmax, mmax, seed + f, mres, power) for f in range(nfields)])


# Uneven-lmax workload: 256 scalar transforms with lmax uniform over these degrees
# (SplitMix64, seed 4), each on a Gauss grid nlat = 3(lmax+1)/2, nphi = 2 nlat; field i
# uses data seed 5 + i.
UNEVEN_LMAX = (63, 127, 255, 511, 1023)
UNEVEN_N, UNEVEN_SEED_L, UNEVEN_SEED_DATA = 256, 4, 5


def uneven_lmaxes(n=UNEVEN_N, seed=UNEVEN_SEED_L):
    """lmax of the n transforms of the uneven workload."""
    u = splitmix64(seed, n)
    return [UNEVEN_LMAX[int(x % len(UNEVEN_LMAX))] for x in u]
