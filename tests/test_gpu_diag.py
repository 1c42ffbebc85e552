"""Diagnostics of save() and plot.py on the device against the oracle:
dissipation spectrum, divergence field (forced-dns-sm.py:216-229, 472),
resident droplet mass (plot.py:56-72)."""
import numpy as np
import pytest

from helpers import inputs, make_flow_oracle, relerr
from lucky_droplets_amd.config import Config

pytestmark = pytest.mark.gpu


def _t(x, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


def _cfg(N, prec, nparts=0):
    return Config(N=N, precision=prec, nu=0.02, nshells=2, profile="k4", n_particles=nparts,
                  st_dist=("uniform", 0.05, 1.0), gravity=True)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dissipation_and_divergence_match_oracle(gpu, prec):
    from lucky_droplets_amd.dns import ForcedDNS
    N = 32
    cfg = _cfg(N, prec)
    uk, _, _, _ = inputs(cfg, seed=3)
    fl = make_flow_oracle(cfg)
    G = fl.grid
    d = ForcedDNS(cfg, max_particles=0)
    tol = 1e-12 if prec == "f64" else 2e-5
    D = d.dissipation(_t(uk, gpu, d.cdtype)).cpu().numpy()
    D_ref = G.dissipation_spectrum(uk, cfg.nu_eff)
    assert relerr(D, D_ref) < tol
    # a field with divergence: add a gradient part i k phi to the solenoidal one
    rng = np.random.default_rng(5)
    phi = (rng.standard_normal(uk.shape[1:]) + 1j * rng.standard_normal(uk.shape[1:])) * G.dealias
    wk = uk + 1j * np.stack([G.kx, G.ky, G.kz]) * phi * np.abs(uk).max()
    w = np.stack([G.irfft_mpi(wk[c]) for c in range(3)])
    div_ref = G.divergence(w)
    div = d.divergence(_t(wk, gpu, d.cdtype)).cpu().numpy()
    assert relerr(div, div_ref) < (1e-11 if prec == "f64" else 1e-5)
    assert abs(d.divergence_max(_t(wk, gpu, d.cdtype)) - np.abs(div_ref).max()) < 1e-5 * np.abs(div_ref).max()
    # the solenoidal state: divergence at round-off
    assert d.divergence_max(_t(uk, gpu, d.cdtype)) < (1e-12 if prec == "f64" else 1e-5) * np.abs(div_ref).max()
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_droplet_mass_of_resident_state(gpu, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    N, npart = 32, 5000
    cfg = _cfg(N, prec, npart)
    uk, n, coord, pid = inputs(cfg, seed=4)
    d = ForcedDNS(cfg)
    d.set_state(_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
    s, mn, mx, q = d.droplet_stats()
    m = coord[:, 6].astype(np.float32 if prec == "f32" else np.float64).astype(np.float64)
    assert abs(s - m.sum()) < 1e-12 * m.sum()
    assert mn == m.min() and mx == m.max()
    assert abs(q - (m * m).sum()) < 1e-12 * (m * m).sum()
    # after steps the total grows by what the field lost (sum n dx^3 + sum m is invariant)
    dx3 = cfg.dx ** 3
    t0 = float(n.sum()) * dx3 + s
    d.advance(2e-3, 3)
    s1 = d.droplet_stats()[0]
    ng = d.zeros_real()
    d.get_state(None, ng, None)
    t1 = float(ng.double().sum()) * dx3 + s1
    assert s1 > s
    assert abs(t1 - t0) < (1e-9 if prec == "f64" else 1e-3) * (s1 - s)
    d.close()


def test_state_file_umat_is_fluid_velocity_at_droplets(gpu, tmp_path):
    """save(..., dns=, stb=) writes umat as the reference does
    (forced-dns-sm.py:455-460): interp_cosine of u at the droplets."""
    import torch
    from lucky_droplets_amd import checkpoint
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    from helpers import make_drops_oracle
    N, npart = 32, 400
    cfg = _cfg(N, "f64", npart)
    uk, n, coord, pid = inputs(cfg, seed=6)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, N, npart, cfg.sts, 0.1, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0, dns=d,
                        coord=coord, prtclid=pid)
    ukg = _t(uk, gpu, d.cdtype)
    base = checkpoint.save(tmp_path, 0.5, ukg, _t(n, gpu, d.rdtype), coord=stb.coord, prtclid=stb.prtclid, dns=d,
                           stb=stb)
    _, _, sdir = checkpoint.run_dirs(tmp_path, 0.5, "wo_g", 0.001, 0.05, True)
    st = np.load(sdir / "state_0.npz")
    G = make_flow_oracle(cfg).grid
    u = np.stack([G.irfft_mpi(uk[c]) for c in range(3)])
    dr = make_drops_oracle(cfg, coord, pid)
    ref = dr.interp_cosine(np.mod(coord, cfg.L), u)
    order = np.argsort(st["prtclid"].reshape(-1))
    assert relerr(st["umat"][order], ref[np.argsort(pid)]) < 1e-12
    d.close()


def test_kernel_class_timing_counts_launches_and_keeps_results(gpu):
    """ld_profile (the bench's per-kernel HIP-event timing): every class
    records one launch per RK stage (the y passes five per stage: u/omega and
    v inverse, nonlinear and density forward, kn inverse), a positive total
    duration, and the steps taken while profiling equal unprofiled ones bit for
    bit."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(32, "f32", 2000)
    rng = np.random.default_rng(8)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, _ = ic.initial_droplets(cfg, rng)
    per_stage = {"droplets": 1, "deposit": 1, "x_rhs": 1, "x_uw": 1, "y_pass": 5, "z_cross": 1, "z_vel": 1,
                 "x_kn": 1, "z_kn": 1}
    ref = None
    for name in [None] + list(per_stage):
        d = ForcedDNS(cfg)
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        if name:
            d.profile(name, True)
        d.advance(dt, 2)
        if name:
            ms, launches = d.profile_read()
            d.profile(name, False)
            assert launches == 2 * 4 * per_stage[name], (name, launches)
            assert ms > 0.0
        d.get_state(*t)
        torch.cuda.synchronize()
        out = [x.cpu().numpy() for x in t]
        d.close()
        if ref is None:
            ref = out
        else:
            for a, b in zip(ref, out):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mass_history_conserves_the_sum(gpu, prec):
    """tracking.mass_history over 20 steps with growing droplets: the
    droplets gain mass, the density loses it, and the sum stays within 1e-12
    (FP64) / 1e-5 (FP32) of the mass exchanged."""
    import torch
    from lucky_droplets_amd import ic, tracking
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(32, prec, 4000)
    rng = np.random.default_rng(12)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, _ = ic.initial_droplets(cfg, rng)
    d = ForcedDNS(cfg)
    d.set_state(_t(uk, gpu, d.cdtype), _t(np.ones_like(n), gpu, d.rdtype), _t(coord, gpu, torch.float64))
    hist = tracking.mass_history(d, dt, 20, 5)
    d.close()
    assert hist.shape == (5, 4) and list(hist[:, 0]) == [0, 5, 10, 15, 20]
    s = tracking.summary(hist)
    assert s["droplet_mass_gained"] > 0 and hist[-1, 2] < hist[0, 2]
    assert s["max_total_drift_per_gained"] < (1e-12 if prec == "f64" else 1e-5)
