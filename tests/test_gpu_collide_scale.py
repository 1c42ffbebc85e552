"""Cell-list collision pairs at 10^5 droplets against an independent periodic
k-d tree search (oracle.collide.pairs_kdtree), bit-exact; and the pair count
of the 10^6-droplet benchmark population with its own radii."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_collision_pairs_1e5_against_kdtree(gpu):
    import torch
    from lucky_droplets_amd.config import Config
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    from oracle import collide
    n = 100_000
    cfg = Config(N=64, precision="f32", nu=0.01, nshells=2, profile="k4", n_particles=n)
    rng = np.random.default_rng(40)
    rows = rng.uniform(0, cfg.L, (n, 7))
    rows[:, :3] += rng.integers(-1, 2, (n, 3)) * cfg.L
    rows[: n // 20, 2] = rng.uniform(-1e-3, 1e-3, n // 20)      # on the periodic seam
    r = rng.uniform(0.005, 0.02, n)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, n, cfg.sts, 0.05, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=rows)
    got = stb.collision_pairs(torch.from_numpy(r).to(gpu), max_pairs=1 << 18).cpu().numpy()
    expect = collide.pairs_kdtree(rows, r, cfg.L)
    assert expect.shape[0] > 500
    assert np.array_equal(got, expect)
    d.close()


def test_collision_pairs_benchmark_population(gpu):
    """10^6 droplets of the 256^3 workload (St lognormal, median 0.3), radius
    r = sqrt(9 nu tau_eta St / (2 rho_p)): the same pairs as the k-d tree."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    from oracle import collide
    cfg = preset("c2")
    rng = np.random.default_rng(cfg.seed)
    coord, pid = ic.initial_droplets(cfg, rng)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, cfg.n_particles, cfg.sts, 0.3, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop,
                        cfg.M0, dns=d, coord=coord, prtclid=pid)
    rad = stb.radius(cfg.nu_eff, cfg.rhop)
    got = stb.collision_pairs(rad, max_pairs=1 << 16).cpu().numpy()
    expect = collide.pairs_kdtree(coord, rad.cpu().numpy(), cfg.L)
    assert np.array_equal(got, expect)
    d.close()
