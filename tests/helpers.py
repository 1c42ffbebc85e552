"""Shared builders for the parity tests: one seeded input set, handed to the
GPU path and to the CPU oracle."""
import numpy as np

import oracle
from lucky_droplets_amd import ic
from lucky_droplets_amd.config import Config


def make_flow_oracle(cfg: Config) -> "oracle.Flow":
    return oracle.Flow(cfg.N, cfg.nu_eff, cfg.shell_no, cfg.f0, cfg.tps, cfg.g, cfg.isforcing,
                       1.0 if cfg.explicit else 0.0, cfg.lp, cfg.L)


def make_drops_oracle(cfg: Config, coord, prtclid) -> "oracle.Droplets":
    return oracle.Droplets(cfg.L, cfg.N, coord.copy(), cfg.sts, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                           prtclid.reshape(-1, 1).copy())


def inputs(cfg: Config, seed: int = 0, n_particles=None, n_noise=0.0):
    """(uk, n, coord, prtclid) float64/complex128 numpy inputs."""
    rng = np.random.default_rng(seed)
    uk, n, _ = ic.initial_flow(cfg, rng)
    coord, pid = ic.initial_droplets(cfg, rng, n_particles)
    if n_noise:
        n = n + n_noise * (rng.random(n.shape) - 0.5)
    return uk, n, coord, pid


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)


def by_id(rows, ids):
    ids = np.asarray(ids).reshape(-1).astype(np.int64)
    return np.asarray(rows)[np.argsort(ids, kind="stable")]


def min_image(a, b, L):
    """Largest periodic position difference between matching rows (columns 0..2)."""
    d = np.asarray(a)[:, :3] - np.asarray(b)[:, :3]
    d = d - L * np.round(d / L)
    return float(np.max(np.abs(d))) if d.size else 0.0
