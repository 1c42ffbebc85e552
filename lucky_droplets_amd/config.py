"""Run configuration: the reference's module constants as a dataclass, plus
named presets for the benchmark workloads.

Reference constants (forced-dns-sm.py:11-38, 78-97; particles.py:74-75):
  nu0 = 0.8, m = 1.5, nu = nu0 (3 m / N)^(2 (lp - 1/3))          :81-83
  tf  = 9 (m / N)^(2/3) / nu0           (Kolmogorov time)          :92
  tps = sts * tf                        (small-droplet time)        :93
  g   = 9.81 * 0.1 / tf if gravity else 0                           :94
  f0  = nu0^3 (2 pi)^3 / nshells        (power per forced shell)    :97
  einit = (2 pi)^3                                                  :86
  factor = (nu tf / (2 rho_p))**3 / 2 * 36 pi rho_p / M0            particles.py:74
  growthfactor = 9 pi nu / (2 rho_p)                                particles.py:75
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Tuple


@dataclasses.dataclass
class Config:
    N: int = 64
    precision: str = "f64"          # "f32" or "f64": fields and transforms
    nu: Optional[float] = None      # None -> the reference's resolution scaling
    lp: int = 1
    nu0: float = 0.8
    m_eta: float = 1.5              # the reference's `m` (desired kmax*eta)
    nshells: int = 1
    isforcing: bool = True
    explicit: bool = True
    gravity: bool = False
    sts: float = 0.001              # small-droplet Stokes number
    rhop: float = 1000.0
    M0: float = 4e-6
    L: float = 2 * math.pi
    # initial condition
    profile: str = "reference"      # "reference" (k^2 e^{-k^2/2}) or "k4" (k^4 e^{-2 (k/3)^2})
    kinit: int = 31
    seed: int = 0
    # droplets
    n_particles: int = 0
    st_dist: Tuple = ("mono", 0.05)  # ("mono", St) | ("uniform", lo, hi) | ("lognormal", median, sigma) | ("bi", St1, St2)
    particle_velocity: str = "uniform"   # "uniform": U[0, L) as particles.py:91
    # time stepping
    dt: Optional[float] = None      # None -> CFL
    cfl: float = 0.5
    steps: int = 10

    # ----------------------------------------------------------- derived
    @property
    def nu_eff(self) -> float:
        if self.nu is not None:
            return float(self.nu)
        return self.nu0 * (3 * self.m_eta / self.N) ** (2 * (self.lp - 1 / 3))   # :83

    @property
    def tf(self) -> float:
        return 9 * (self.m_eta / self.N) ** (2 / 3) / self.nu0                     # :92

    @property
    def tps(self) -> float:
        return self.sts * self.tf                                                  # :93

    @property
    def g(self) -> float:
        return 9.81 * 0.1 / self.tf if self.gravity else 0.0                       # :94

    @property
    def shell_no(self) -> Tuple[int, ...]:
        return tuple(range(1, 1 + self.nshells))                                  # :90

    @property
    def f0(self) -> float:
        return self.nu0 ** 3 * (2 * math.pi) ** 3 / max(self.nshells, 1)           # :97

    @property
    def einit(self) -> float:
        return (2 * math.pi) ** 3                                                  # :86

    @property
    def factor(self) -> float:
        nu, tf, rp = self.nu_eff, self.tf, self.rhop
        return (nu * tf / (2 * rp)) ** 3 / 2 * (36 * math.pi * rp) / self.M0       # particles.py:74

    @property
    def growthfactor(self) -> float:
        return 9 * math.pi * self.nu_eff / (2 * self.rhop)                         # particles.py:75

    @property
    def dx(self) -> float:
        return self.L / self.N

    def replace(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)


def preset(name: str) -> Config:
    """Named workloads.

    c1   64^3 FP64, nu = 0.01, shells |k| <= 2, E(k) ~ k^4 e^{-2(k/3)^2}, CFL 0.5,
         200 steps, 1e4 droplets with St ~ U[0.05, 1], drag + settling.
    c2   256^3 FP32, nu = 2.5e-3, same forcing/spectrum, 500 steps, 1e6 droplets,
         St lognormal (median 0.3, sigma_lnSt = 0.6).  The benchmark workload.
    c2_64  the c2 workload in FP64 (the GPU oracle form).
    c3   one 64^3 FP32 realization of the ensemble (nu = 0.01, 1e5 droplets).
    c4   512^3 FP32, nu = 1e-3, 1e7 droplets, bidisperse St 0.2 / 0.6.
    ref  the reference's own defaults at N = 64 (forced-dns-sm.py:28-97).
    """
    if name == "c1":
        return Config(N=64, precision="f64", nu=0.01, nshells=2, profile="k4", n_particles=10_000,
                      st_dist=("uniform", 0.05, 1.0), gravity=True, steps=200)
    if name == "c2":
        return Config(N=256, precision="f32", nu=2.5e-3, nshells=2, profile="k4", n_particles=1_000_000,
                      st_dist=("lognormal", 0.3, 0.6), gravity=False, steps=500)
    if name == "c2_64":
        return preset("c2").replace(precision="f64")
    if name == "c3":
        return Config(N=64, precision="f32", nu=0.01, nshells=2, profile="k4", n_particles=100_000,
                      st_dist=("uniform", 0.05, 1.0), gravity=False, steps=100)
    if name == "c4":
        return Config(N=512, precision="f32", nu=1e-3, nshells=2, profile="k4", n_particles=10_000_000,
                      st_dist=("bi", 0.2, 0.6), gravity=False, steps=2000)
    if name == "ref":
        return Config(N=64, precision="f64", nshells=1, profile="reference", n_particles=0, dt=0.256 / 64)
    raise KeyError(name)
