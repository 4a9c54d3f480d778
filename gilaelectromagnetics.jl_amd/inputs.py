"""Seeded synthetic source fields (column-major (c1,c2,c3,3[,n]) arrays).

SplitMix64 -> uniform doubles -> Box-Muller: portable, so the same seed gives
the same field on any host.
"""
import numpy as np


def splitmix64(seed, n):
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def complex_gaussian(shape, seed):
    """Standard complex Gaussian (re, im ~ N(0, 1/2) each)."""
    n = int(np.prod(shape))
    u = uniform(seed, 2 * n)
    r = np.sqrt(-np.log(u[0::2]))
    ph = 2 * np.pi * u[1::2]
    return np.asfortranarray((r * np.cos(ph) + 1j * r * np.sin(ph)).reshape(shape, order="F"))


def cell_centres(cel, scl, org=(0, 0, 0)):
    """Cell-centre coordinates per axis (wavelengths), GlaVol's grid."""
    out = []
    for c, s, o in zip(cel, scl, org):
        s, o = float(s), float(o)
        brd = s * (c // 2 - (0.5 if c % 2 == 0 else 0.0))
        out.append(o - brd + s * np.arange(c))
    return out


def plane_wave(cel, scl, khat, pol, org=(0, 0, 0)):
    """E(r) = pol * exp(2 pi i khat . r), r the cell centres in wavelengths."""
    gx, gy, gz = cell_centres(cel, scl, org)
    ph = np.exp(2j * np.pi * (khat[0] * gx[:, None, None] + khat[1] * gy[None, :, None] + khat[2] * gz[None, None, :]))
    out = np.empty(tuple(cel) + (3,), dtype=np.complex128, order="F")
    for c in range(3):
        out[..., c] = pol[c] * ph
    return out


def point_dipole(cel, cell, orient):
    out = np.zeros(tuple(cel) + (3,), dtype=np.complex128, order="F")
    out[tuple(cell) + (slice(None),)] = orient
    return out


def random_unit(seed):
    g = np.real(complex_gaussian((3,), seed)) + np.imag(complex_gaussian((3,), seed + 7919))
    return g / np.linalg.norm(g)
