"""Initial conditions and droplet populations (host side, numpy float64).

The flow recipe restates forced-dns-sm.py:691-737 followed by the loop's
``uk = rfft(u) * dealias`` (:567-569); the droplet recipe restates
particles.py:91-96 with a seeded generator and a choice of Stokes-number
distribution.  These produce the inputs of a run; they are not on the timed
path, and the same arrays are handed to the GPU and to any checker.
"""
from __future__ import annotations

import os

import numpy as np
from scipy.fft import fft, fftfreq, ifft, irfft2, rfft2

from .config import Config

# scipy.fft threads: the host's cores, at most 16 (one GPU's share of a shared host)
_WORKERS = min(16, os.cpu_count() or 1)


class _Spec:
    """Wavenumber arrays and operators of forced-dns-sm.py:55-62, 123-138."""

    def __init__(self, N: int, L: float):
        self.N, self.Nf = N, N // 2 + 1
        TWO_PI = 2 * np.pi
        Kx = fftfreq(N, 1.0 / N) * TWO_PI / L
        Kz = np.abs(Kx[: self.Nf])
        self.kx, self.ky, self.kz = np.meshgrid(Kx, Kx, Kz, indexing="ij")
        lap = -(self.kx ** 2 + self.ky ** 2 + self.kz ** 2)
        self.k = (-lap) ** 0.5
        self.kint = np.clip(np.round(self.k, 0).astype(int), None, N // 2)
        self.dealias = self.kint <= N / 3
        self.invlap = self.dealias / np.where(lap == 0, np.inf, lap)
        self.normalize = np.where((self.kz == 0) + (self.kz == N // 2), 1 / (N ** 6 / TWO_PI ** 3),
                                  2 / (N ** 6 / TWO_PI ** 3))
        self.shells = np.arange(-0.5, self.Nf, 1.0)
        self.shells[0] = 0.0

    def rfft(self, u):
        return fft(rfft2(u, axes=(-2, -1), workers=_WORKERS), axis=-3, workers=_WORKERS)

    def irfft(self, fu):
        return irfft2(ifft(fu, axis=-3, workers=_WORKERS), (self.N, self.N), axes=(-2, -1), workers=_WORKERS)

    def e1d(self, x):
        return np.histogram(self.k.ravel(), bins=self.shells, weights=x.ravel())[0]

    def energy3d(self, uk):
        return 0.5 * (np.abs(uk[0]) ** 2 + np.abs(uk[1]) ** 2 + np.abs(uk[2]) ** 2) * self.normalize


def initial_flow(cfg: Config, rng: np.random.Generator):
    """Random-phase solenoidal field of energy einit; returns ``(uk, n, u)``.

    uk: (3, N, N, N/2+1) complex128 as the step loop sees it; n: ones (N,N,N);
    u: the physical velocity before the loop's final rfft (used for the CFL dt).
    """
    S = _Spec(cfg.N, cfg.L)
    N = cfg.N
    TWO_PI = 2 * np.pi
    thu = rng.uniform(0, TWO_PI, S.k.shape)                               # :695-697
    thv = rng.uniform(0, TWO_PI, S.k.shape)
    thw = rng.uniform(0, TWO_PI, S.k.shape)
    kint = S.kint
    if cfg.profile == "reference":
        eprofile = kint ** 2 * np.exp(-kint ** 2 / 2) / S.normalize         # :700
    elif cfg.profile == "k4":
        eprofile = kint ** 4.0 * np.exp(-2.0 * (kint / 3.0) ** 2) / S.normalize
    else:
        raise ValueError(f"unknown profile {cfg.profile!r}")
    amp = (eprofile / np.where(kint == 0, np.inf, kint ** 2)) ** 0.5       # :703
    sel = (kint ** 2 < cfg.kinit ** 2) * (kint > 0) * S.dealias
    uk = np.empty((3, N, N, S.Nf), dtype=np.complex128)
    for c, th in enumerate((thu, thv, thw)):                               # :705-707
        uk[c] = amp * np.exp(1j * th) * sel
    del thu, thv, thw
    u = np.stack([S.irfft(uk[c]) for c in range(3)])                       # :709-711
    uk = np.stack([S.rfft(u[c]) for c in range(3)])                        # :713-715
    trm = S.kx * uk[0] + S.ky * uk[1] + S.kz * uk[2]                       # :717
    uk[0] = uk[0] + S.invlap * S.kx * trm                                  # :718-720
    uk[1] = uk[1] + S.invlap * S.ky * trm
    uk[2] = uk[2] + S.invlap * S.kz * trm
    e0 = np.sum(S.e1d(S.energy3d(uk)))                                     # :722-725
    uk = uk * (cfg.einit / e0) ** 0.5                                      # :726-728
    u = np.stack([S.irfft(uk[c]) for c in range(3)])                       # :731-733
    n = np.ones((N, N, N))                                                 # :736
    uk = np.stack([S.rfft(u[c]) * S.dealias for c in range(3)])            # :567-569
    return uk, n, u


def cfl_dt(cfg: Config, u: np.ndarray) -> float:
    """dt = cfl * dx / max(|u| + |v| + |w|) at t = 0 (fixed afterwards)."""
    if cfg.dt is not None:
        return float(cfg.dt)
    umax = float(np.max(np.abs(u[0]) + np.abs(u[1]) + np.abs(u[2])))
    return cfg.cfl * cfg.dx / umax


def stokes_numbers(cfg: Config, rng: np.random.Generator, n: int) -> np.ndarray:
    kind = cfg.st_dist[0]
    if kind == "mono":
        return np.full(n, float(cfg.st_dist[1]))
    if kind == "uniform":
        return rng.uniform(cfg.st_dist[1], cfg.st_dist[2], n)
    if kind == "lognormal":
        return cfg.st_dist[1] * np.exp(cfg.st_dist[2] * rng.standard_normal(n))
    if kind == "bi":
        return np.where(rng.random(n) < 0.5, cfg.st_dist[1], cfg.st_dist[2])
    raise ValueError(f"unknown Stokes distribution {cfg.st_dist!r}")


def initial_droplets(cfg: Config, rng: np.random.Generator, n: int | None = None):
    """Droplet rows (n, 7) = [x, y, z, vx, vy, vz, m/M0] and ids (particles.py:91-96).

    All seven columns start U[0, L) as in the reference; the mass column is then
    St**1.5 * factor for the drawn Stokes numbers.
    """
    n = cfg.n_particles if n is None else n
    coord = rng.uniform(0, cfg.L, (n, 7))                                  # :91
    st = stokes_numbers(cfg, rng, n)
    coord[:, -1] = st ** 1.5 * cfg.factor                                  # :94
    prtclid = np.arange(n, dtype=np.int64)                                 # :92
    return coord, prtclid
