"""Seeded synthetic inputs shared by the benchmark harness and the tests.

Every random number comes from SplitMix64 (Steele, Lea and Flood, 2014), specified bit for
bit so that the same seed gives the same field on any host: state s advances by
0x9E3779B97F4A7C15 per draw, each draw is mixed by the two xor-shift-multiply rounds below,
and a uniform double is (x >> 11) * 2**-53.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, start=0):
    """Return draws start..start+n-1 of the SplitMix64 stream seeded with `seed` (uint64)."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, n, start=0):
    return (splitmix64(seed, n, start) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def interior_uniform(shape, ng, seed, lo=-1.0, hi=1.0, fill=np.nan):
    """Array of `shape` = (nblocks, ncomp, Ek, Ej, Ei) with U(lo, hi) interiors drawn in
    (block, comp, k, j, i) order and `fill` in the ghost zones."""
    nb, nc, ek, ej, ei = shape
    gk, gj, gi = ng[2], ng[1], ng[0]
    nk, nj, ni = ek - 2 * gk, ej - 2 * gj, ei - 2 * gi
    vals = uniform(seed, nb * nc * nk * nj * ni)
    out = np.full(shape, fill, dtype=np.float64)
    out[:, :, gk:gk + nk, gj:gj + nj, gi:gi + ni] = (lo + (hi - lo) * vals).reshape(
        nb, nc, nk, nj, ni)
    return out
