"""Seeded synthetic workloads for the tracer (host side, numpy).

Every draw comes from a counter-based generator: value = U(seed, stream, i)
with a splitmix64 hash, so particle i of a workload is the same whatever the
batch size (a 16-particle subsample is the first 16 particles of the full
batch).  Streams name the variables; they do not depend on draw order.

Workloads (sizes are the full ones; pass n to take a prefix):

- ``dipole_protons_1mev``: 1024 protons, K = 1 MeV, EarthDipole, equatorial
  start at r ~ U[4,7] Re, azimuth ~ U[0,2pi), pitch ~ U[10,90] deg, gyrophase
  ~ U[0,2pi); Particle.advance(10) at cyclotronresolution 20.  Seed 0.
- ``doubledipole_electrons_gc``: 65,536 electrons, log10 K[eV] ~ U[4,7],
  L ~ U[3,8], azimuth ~ U[0,2pi), equatorial pitch ~ U[15,85] deg,
  DoubleDipole, GuidingCenter.advance(60) with rtol = atol = 1e-8.  Seed 1.
- ``current_sheet_protons``: 262,144 protons in B = 20 nT tanh(z/0.5 Re) x
  + 1 nT z, E = 0.3 mV/m y; log10 K[eV] ~ U[3,5], isotropic directions,
  x, y ~ U[-5,5] Re, z ~ U[-1.5,1.5] Re; Adaptive.advance(120).  Seed 2.
- ``mixed_dipole_uniform``: 1,048,576 particles, half protons half
  electrons, EarthDipole + (0,0,-10 nT), log10 K ~ U[5, log10 5e6],
  L ~ U[3,7], pitch ~ U[10,90] deg; advance(30).  Seed 3.
"""
import numpy as np

from . import constants as K

_M64 = (1 << 64) - 1


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, stream, n, start=0):
    """n doubles in [0,1) for (seed, stream), indices start..start+n-1."""
    key = _splitmix64(_splitmix64(np.uint64(seed)) ^ np.uint64(stream))
    idx = np.arange(start, start + n, dtype=np.uint64)
    bits = _splitmix64(key ^ _splitmix64(idx))
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(seed, stream, n, lo, hi):
    return lo + (hi - lo) * uniform01(seed, stream, n)


def speedfromKE(KE_eV, mass):
    """Vectorised rapt/utils.py:218-249 (eV input)."""
    KE = np.asarray(KE_eV, dtype=np.float64) * K.e
    mc2 = mass * K.c ** 2
    rel = KE / mc2 >= 1e-6
    out = np.sqrt(2 * KE / mass)
    out = np.where(rel, K.c * np.sqrt(1 - (mc2 / (mc2 + KE)) ** 2), out)
    return out


def dipole_protons_1mev(n=1024, seed=0):
    r = uniform(seed, 0, n, 4.0, 7.0) * K.Re
    phi = uniform(seed, 1, n, 0.0, 2 * np.pi)
    pa = uniform(seed, 2, n, 10.0, 90.0)
    psi = uniform(seed, 3, n, 0.0, 2 * np.pi)
    v = speedfromKE(np.full(n, 1e6), K.m_pr)
    pos = np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(n)], axis=1)
    rhat = np.stack([np.cos(phi), np.sin(phi), np.zeros(n)], axis=1)
    bhat = np.array([0.0, 0.0, 1.0])
    brh = np.cross(bhat, rhat)
    a = np.radians(pa)
    vel = v[:, None] * (np.cos(a)[:, None] * bhat +
                        np.sin(a)[:, None] * (np.cos(psi)[:, None] * rhat +
                                              np.sin(psi)[:, None] * brh))
    return dict(pos=pos, vel=vel, mass=np.full(n, K.m_pr), charge=np.full(n, K.e),
                t0=0.0, delta=10.0, cyclotronresolution=20)


def doubledipole_electrons_gc(n=65536, seed=1):
    logK = uniform(seed, 0, n, 4.0, 7.0)
    L = uniform(seed, 1, n, 3.0, 8.0)
    phi = uniform(seed, 2, n, 0.0, 2 * np.pi)
    pa = uniform(seed, 3, n, 15.0, 85.0)
    v = speedfromKE(10.0 ** logK, K.m_el)
    pos = np.stack([L * K.Re * np.cos(phi), L * K.Re * np.sin(phi), np.zeros(n)], axis=1)
    return dict(pos=pos, v=v, pa=pa, mass=np.full(n, K.m_el), charge=np.full(n, -K.e),
                t0=0.0, delta=60.0, rtol=1e-8, atol=1e-8, logK=logK, L=L)


def current_sheet_protons(n=262144, seed=2):
    logK = uniform(seed, 0, n, 3.0, 5.0)
    cth = uniform(seed, 1, n, -1.0, 1.0)
    phv = uniform(seed, 2, n, 0.0, 2 * np.pi)
    x = uniform(seed, 3, n, -5.0, 5.0) * K.Re
    y = uniform(seed, 4, n, -5.0, 5.0) * K.Re
    z = uniform(seed, 5, n, -1.5, 1.5) * K.Re
    v = speedfromKE(10.0 ** logK, K.m_pr)
    sth = np.sqrt(1 - cth ** 2)
    vel = v[:, None] * np.stack([sth * np.cos(phv), sth * np.sin(phv), cth], axis=1)
    pos = np.stack([x, y, z], axis=1)
    return dict(pos=pos, vel=vel, mass=np.full(n, K.m_pr), charge=np.full(n, K.e),
                t0=0.0, delta=120.0, B0=20e-9, Bn=1e-9, lam=0.5 * K.Re, Ey=0.3e-3)


def mixed_dipole_uniform(n=1048576, seed=3):
    logK = uniform(seed, 0, n, 5.0, np.log10(5e6))
    L = uniform(seed, 1, n, 3.0, 7.0)
    phi = uniform(seed, 2, n, 0.0, 2 * np.pi)
    pa = uniform(seed, 3, n, 10.0, 90.0)
    psi = uniform(seed, 4, n, 0.0, 2 * np.pi)
    species = (np.arange(n) % 2).astype(np.int32)  # 0 proton, 1 electron
    mass = np.where(species == 0, K.m_pr, K.m_el)
    charge = np.where(species == 0, K.e, -K.e)
    v = np.where(species == 0, speedfromKE(10.0 ** logK, K.m_pr),
                 speedfromKE(10.0 ** logK, K.m_el))
    r = L * K.Re
    pos = np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(n)], axis=1)
    rhat = np.stack([np.cos(phi), np.sin(phi), np.zeros(n)], axis=1)
    bhat = np.array([0.0, 0.0, 1.0])
    brh = np.cross(bhat, rhat)
    a = np.radians(pa)
    vel = v[:, None] * (np.cos(a)[:, None] * bhat +
                        np.sin(a)[:, None] * (np.cos(psi)[:, None] * rhat +
                                              np.sin(psi)[:, None] * brh))
    return dict(pos=pos, vel=vel, v=v, pa=pa, mass=mass, charge=charge, species=species,
                t0=0.0, delta=30.0, dBz=-10e-9, L=L, logK=logK)
