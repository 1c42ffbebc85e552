"""GPU path (libld.so) vs the CPU FP64 oracle on identical seeded inputs.

Tolerances: FP64 1e-10 relative (max-abs error over max-abs value, per
field); FP32 2e-5 relative for single transforms and 1e-4 for whole RHS/step
results (single-precision rounding of O(1e2) transforms)."""
import numpy as np
import pytest

from helpers import by_id, inputs, make_drops_oracle, make_flow_oracle, min_image, relerr
from lucky_droplets_amd.config import Config

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-10, "f32": 1e-4}


def _t(x, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


def _cfg(N, prec, nparts=0, **kw):
    base = dict(N=N, precision=prec, nu=0.02, nshells=2, profile="k4", n_particles=nparts,
                st_dist=("uniform", 0.05, 1.0), gravity=True)
    base.update(kw)
    return Config(**base)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("N", [16, 32, 64, 128])
def test_fft_matches_scipy(gpu, N, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(N, prec)
    d = ForcedDNS(cfg, max_particles=0)
    G = make_flow_oracle(cfg).grid
    rng = np.random.default_rng(N)
    u = rng.standard_normal((3, N, N, N))
    ref = np.stack([G.rfft_mpi(u[c]) for c in range(3)])
    fu = d.zeros_spec(3)
    d.rfft_mpi(_t(u, gpu, d.rdtype), fu)
    tol = 1e-13 if prec == "f64" else 2e-6
    assert relerr(fu.cpu().numpy(), ref) < tol
    # inverse of a general (non-Hermitian) half spectrum: c2r reads Re of kz = 0, N/2
    X = rng.standard_normal((2, N, N, N // 2 + 1)) + 1j * rng.standard_normal((2, N, N, N // 2 + 1))
    refi = np.stack([G.irfft_mpi(X[c]) for c in range(2)])
    uo = d.zeros_real(2)
    d.irfft_mpi(_t(X, gpu, d.cdtype), uo)
    assert relerr(uo.cpu().numpy(), refi) < tol
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_forcing(gpu, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(32, prec)
    uk, n, coord, pid = inputs(cfg, seed=1)
    fl = make_flow_oracle(cfg)
    ref = fl.forcing(uk)
    d = ForcedDNS(cfg, max_particles=0)
    fk = d.zeros_spec(3)
    ek = torch.zeros(cfg.N // 2 + 1, dtype=torch.float64, device=gpu)
    d.forcing(_t(uk, gpu, d.cdtype), fk, ek)
    assert relerr(fk.cpu().numpy(), ref) < (1e-12 if prec == "f64" else 1e-5)
    e_ref = fl.grid.e3d_to_e1d(fl.grid.energy3d(uk))
    assert relerr(ek.cpu().numpy()[1:3], e_ref[1:3]) < (1e-12 if prec == "f64" else 1e-5)
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_clip_zero(gpu, prec):
    """ld_clip_zero: x clipped in place, out = x * oldmean / newmean; misaligned
    field pointers are refused."""
    import ctypes
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(32, prec)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((32, 32, 32)) + 0.5
    fl = make_flow_oracle(cfg)
    xr = x.copy()
    ref = fl.clip_zero(xr)
    d = ForcedDNS(cfg, max_particles=0)
    xg = _t(x, gpu, d.rdtype)
    out = d.clip_zero(xg)
    tol = 1e-14 if prec == "f64" else 1e-6
    assert relerr(out.cpu().numpy(), ref) < tol
    assert relerr(xg.cpu().numpy(), xr) < tol
    raw = torch.zeros(32 ** 3 + 8, dtype=d.rdtype, device=gpu)
    rc = d.lib.ld_clip_zero(d.ctx, ctypes.c_void_p(raw.data_ptr() + raw.element_size()),
                            ctypes.c_void_p(raw.data_ptr()))
    assert rc == 1
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nparts", [0, 40])
def test_full_rhs(gpu, prec, nparts):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(32, prec, nparts)
    uk, n, coord, pid = inputs(cfg, seed=2, n_noise=0.8)
    fl = make_flow_oracle(cfg)
    dr = make_drops_oracle(cfg, coord, pid) if nparts else None
    ku_r, kn_r, kp_r, _ = fl.full_RHS(0.0, uk, n.copy(), coord.copy(), coord.copy(), dr)
    d = ForcedDNS(cfg, max_particles=nparts)
    u_out = d.zeros_real(3)
    v_out = d.zeros_real(3)
    fc = d.zeros_real()
    ku, kn, kp, _ = d.full_RHS(0.0, _t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), None,
                               _t(coord, gpu, torch.float64) if nparts else None, fc=fc, u_out=u_out, v_out=v_out)
    tol = TOL[prec]
    assert relerr(u_out.cpu().numpy(), fl.last["u"]) < tol
    assert relerr(v_out.cpu().numpy(), fl.last["v"]) < tol
    assert relerr(ku.cpu().numpy(), ku_r) < tol
    assert relerr(kn.cpu().numpy(), kn_r) < tol
    if nparts:
        assert relerr(fc.cpu().numpy(), fl.last["fc"]) < tol
        kp_id = by_id(kp.cpu().numpy(), pid)
        kp_ref = by_id(kp_r, dr.prtclid)
        for cols in (slice(0, 3), slice(3, 6), slice(6, 7)):
            assert relerr(kp_id[:, cols], kp_ref[:, cols]) < tol
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rk4(gpu, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    nparts = 40
    cfg = _cfg(32, prec, nparts)
    uk, n, coord, pid = inputs(cfg, seed=3, n_noise=0.8)
    n = np.abs(n)
    fl = make_flow_oracle(cfg)
    dr = make_drops_oracle(cfg, coord, pid)
    h = 0.01
    sump_r, uknew_r, nnew_r = fl.RK4(0.0, h, dr, uk, n.copy())
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, nparts, cfg.sts, 0.05, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=coord, prtclid=pid)
    sump, uknew, nnew = d.RK4(0.0, h, stb, _t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype))
    tol = TOL[prec]
    assert relerr(uknew.cpu().numpy(), uknew_r) < tol
    assert relerr(nnew.cpu().numpy(), nnew_r) < tol
    a = by_id(sump.cpu().numpy(), pid)
    b = by_id(sump_r, dr.prtclid)
    assert min_image(a, b, cfg.L) < tol * cfg.L
    assert relerr(a[:, 3:], b[:, 3:]) < tol
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_steps(gpu, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    nparts = 60
    cfg = _cfg(32, prec, nparts)
    uk, n, coord, pid = inputs(cfg, seed=4)
    fl = make_flow_oracle(cfg)
    dr = make_drops_oracle(cfg, coord, pid)
    h = 0.01
    uk_r, n_r = uk.copy(), n.copy()
    for _ in range(3):
        uk_r, n_r = fl.step(0.0, h, dr, uk_r, n_r)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, nparts, cfg.sts, 0.05, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=coord, prtclid=pid)
    ukg = _t(uk, gpu, d.cdtype)
    ng = _t(n, gpu, d.rdtype)
    d.step(0.0, h, stb, ukg, ng, nsteps=3)
    tol = TOL[prec]
    assert relerr(ukg.cpu().numpy(), uk_r) < tol
    assert relerr(ng.cpu().numpy(), n_r) < tol
    a = by_id(stb.coord.cpu().numpy(), pid)
    b = by_id(dr.coord, dr.prtclid)
    assert min_image(a, b, cfg.L) < tol * cfg.L
    assert relerr(a[:, 3:], b[:, 3:]) < tol
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_steps_dense_droplets(gpu, prec):
    """Many droplets per cell: exercises the spatial sort, shared deposit nodes
    and the overlapped density stream; rows come back in input order."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    nparts = 3000
    cfg = _cfg(16, prec, nparts, nu=0.05)
    uk, n, coord, pid = inputs(cfg, seed=9)
    fl = make_flow_oracle(cfg)
    dr = make_drops_oracle(cfg, coord, pid)
    h = 0.01
    uk_r, n_r = uk.copy(), n.copy()
    for _ in range(4):
        uk_r, n_r = fl.step(0.0, h, dr, uk_r, n_r)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, nparts, cfg.sts, 0.05, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=coord, prtclid=pid)
    ukg = _t(uk, gpu, d.cdtype)
    ng = _t(n, gpu, d.rdtype)
    d.step(0.0, h, stb, ukg, ng, nsteps=2)
    d.step(0.0, h, stb, ukg, ng, nsteps=2)
    tol = TOL[prec]
    assert relerr(ukg.cpu().numpy(), uk_r) < tol
    assert relerr(ng.cpu().numpy(), n_r) < tol
    a = stb.coord.cpu().numpy()               # input row order is kept
    b = by_id(dr.coord, dr.prtclid)
    assert min_image(a, b, cfg.L) < tol * cfg.L
    assert relerr(a[:, 3:], b[:, 3:]) < tol
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_plain_form_matches_oracle(gpu, prec):
    """The unfused reference-dataflow path (ld_advance_plain) on the same inputs."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    nparts = 50
    cfg = _cfg(32, prec, nparts)
    uk, n, coord, pid = inputs(cfg, seed=12)
    fl = make_flow_oracle(cfg)
    dr = make_drops_oracle(cfg, coord, pid)
    h = 0.01
    uk_r, n_r = uk.copy(), n.copy()
    for _ in range(2):
        uk_r, n_r = fl.step(0.0, h, dr, uk_r, n_r)
    d = ForcedDNS(cfg)
    d.set_state(_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
    d.advance_plain(h, 2)
    ukg = d.zeros_spec(3)
    ng = d.zeros_real()
    cg = torch.zeros((nparts, 7), dtype=torch.float64, device=gpu)
    d.get_state(ukg, ng, cg)
    torch.cuda.synchronize()
    tol = TOL[prec]
    assert relerr(ukg.cpu().numpy(), uk_r) < tol
    assert relerr(ng.cpu().numpy(), n_r) < tol
    a = cg.cpu().numpy()
    b = by_id(dr.coord, dr.prtclid)
    assert min_image(a, b, cfg.L) < tol * cfg.L
    assert relerr(a[:, 3:], b[:, 3:]) < tol
    d.close()


def _collide_case(gpu, n, rlo, rhi, seed, seam=False):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    from oracle import collide
    cfg = _cfg(32, "f64", n)
    rng = np.random.default_rng(seed)
    rows = rng.uniform(0, cfg.L, (n, 7))
    rows[:, :3] += rng.integers(-2, 3, (n, 3)) * cfg.L        # unwrapped positions
    if seam:
        rows[: n // 4, 0] = rng.uniform(-1e-3, 1e-3, n // 4)
    r = rng.uniform(rlo, rhi, n)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, n, cfg.sts, 0.05, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=rows)
    got = stb.collision_pairs(torch.from_numpy(r).to(gpu), max_pairs=1 << 16).cpu().numpy()
    expect = collide.pairs_bruteforce(rows, r, cfg.L)
    return d, stb, rows, r, got, expect


@pytest.mark.parametrize("n,rlo,rhi,seed", [(3000, 0.005, 0.03, 0), (2000, 0.2, 0.6, 1), (5000, 0.001, 0.02, 2)])
def test_collision_pairs_bit_exact(gpu, n, rlo, rhi, seed):
    d, stb, rows, r, got, expect = _collide_case(gpu, n, rlo, rhi, seed, seam=(seed == 2))
    assert expect.shape[0] > 0
    assert np.array_equal(got, expect)
    d.close()


def test_coalescence_bit_exact(gpu):
    from oracle import collide
    d, stb, rows, r, pairs, expect = _collide_case(gpu, 2000, 0.05, 0.15, 3)
    assert np.array_equal(pairs, expect) and pairs.shape[0] > 10
    import torch
    ev = stb.coalesce(torch.from_numpy(pairs).to(gpu)).cpu().numpy()
    ref_rows, ref_ev = collide.coalesce(rows, expect, stb.L)
    assert np.array_equal(ev, ref_ev)
    assert np.array_equal(stb.coord.cpu().numpy(), ref_rows)
    d.close()


def test_collision_capacity_error(gpu):
    import torch
    from lucky_droplets_amd import _lib
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    cfg = _cfg(32, "f64", 500)
    rows = np.random.default_rng(5).uniform(0, cfg.L, (500, 7))
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, 500, cfg.sts, 0.05, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=rows)
    with pytest.raises(_lib.LdError):
        stb.collision_pairs(torch.full((500,), 0.5, dtype=torch.float64, device=gpu), max_pairs=4)
    d.close()


@pytest.mark.parametrize("n,rmax,nbins,seed", [(4000, 0.6, 48, 0), (1500, 2.5, 20, 1), (3000, 0.2, 512, 2)])
def test_pair_statistics(gpu, n, rmax, nbins, seed):
    """Separation histogram bit-exact, radial-velocity sums to 1e-12, g(r) and
    the collision kernel derived from them (rmax = 2.5 > L/3 takes the
    one-cell path)."""
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    from oracle import collide
    cfg = _cfg(32, "f64", n)
    rng = np.random.default_rng(seed)
    rows = rng.uniform(0, cfg.L, (n, 7))
    rows[:, :3] += rng.integers(-2, 3, (n, 3)) * cfg.L
    rows[:, 3:6] = rng.normal(0, 1, (n, 3))
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, n, cfg.sts, 0.05, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=rows)
    got = stb.pair_statistics(rmax, nbins)
    c, ws, wn = collide.pair_stats_bruteforce(rows, cfg.L, rmax, nbins)
    assert c.sum() > 100
    assert np.array_equal(got["counts"], c)
    assert np.allclose(got["wsum"], ws, rtol=1e-12, atol=0)
    assert np.allclose(got["wneg"], wn, rtol=1e-12, atol=0)
    # uniform random positions: g(r) ~ 1 where the bins hold many pairs
    full = got["counts"] > 2000
    if full.any():
        assert np.all(np.abs(got["g"][full] - 1.0) < 0.15)
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_diagnostics(gpu, prec):
    """E(k), the n spectrum, Pi(k) and field statistics of save() / plot.py."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(32, prec)
    uk, n, coord, pid = inputs(cfg, seed=21, n_noise=0.6)
    fl = make_flow_oracle(cfg)
    G = fl.grid
    d = ForcedDNS(cfg, max_particles=0)
    ukg = _t(uk, gpu, d.cdtype)
    ng = _t(n, gpu, d.rdtype)
    tol = 1e-12 if prec == "f64" else 2e-5
    e = d.spectrum(ukg, 0.5).cpu().numpy()
    assert relerr(e, G.e3d_to_e1d(G.energy3d(uk))) < tol
    nk = G.rfft_mpi(n)
    nkg = d.zeros_spec()
    d.rfft_mpi(ng, nkg)
    assert relerr(d.spectrum(nkg, 1.0).cpu().numpy(), G.e3d_to_e1d(np.abs(nk) ** 2 * G.normalize)) < 10 * tol
    ku, _, _, _ = fl.full_RHS(0.0, uk, n.copy(), np.zeros((0, 7)), np.zeros((0, 7)), None, visc=0, forc=0)
    pik = np.real(np.conj(uk[0]) * ku[0] + np.conj(uk[1]) * ku[1] + np.conj(uk[2]) * ku[2]) * G.dealias * G.normalize
    ref = np.cumsum(G.e3d_to_e1d(pik)[::-1])[::-1]
    got = d.flux(ukg, ng).cpu().numpy()
    assert np.max(np.abs(got - ref)) < (1e-9 if prec == "f64" else 1e-3) * np.max(np.abs(ref))
    s, mn, mx, sq = d.field_stats(ng)
    nn = ng.double().cpu().numpy()
    assert abs(s - nn.sum()) < 1e-9 * abs(nn.sum()) and mn == nn.min() and mx == nn.max()
    assert abs(sq - (nn ** 2).sum()) < 1e-9 * (nn ** 2).sum()
    d.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_advance_bitwise_reproducible(gpu, prec):
    """Two runs of the fused step from the same state give identical bits
    (fixed-point deposit: the integer sums do not depend on atomic order)."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c1").replace(precision=prec, n_particles=20000)
    rng = np.random.default_rng(7)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, pid = ic.initial_droplets(cfg, rng)
    outs = []
    for _ in range(2):
        d = ForcedDNS(cfg)
        ukg, ng, cg = _t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64)
        d.set_state(ukg, ng, cg)
        d.advance(dt, 5)
        d.get_state(ukg, ng, cg)
        torch.cuda.synchronize()
        outs.append((ukg.cpu().numpy().copy(), ng.cpu().numpy().copy(), cg.cpu().numpy().copy()))
        d.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("N", [64, 128])
def test_density_chain_order_keeps_results(gpu, N, monkeypatch):
    """Queueing each stage's density chain after the next flow x pass
    (LD_LATE_DENSITY=1, the default from N = 128) or right after its own flow
    stage (0) runs the same kernels on the same data: identical bits."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c1").replace(precision="f32", N=N, n_particles=20000)
    rng = np.random.default_rng(3)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, _ = ic.initial_droplets(cfg, rng)
    outs = []
    for late in ("0", "1"):
        monkeypatch.setenv("LD_LATE_DENSITY", late)
        d = ForcedDNS(cfg)
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        d.advance(dt, 3)
        d.get_state(*t)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in t])
        d.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("N,prec,outside", [(64, "f64", False), (64, "f32", False), (256, "f32", False),
                                            (64, "f64", True), (64, "f32", True)])
def test_dealias_truncation_keeps_results(gpu, N, prec, outside, monkeypatch):
    """The passes that skip the kz > N/3 lines (LD_TRUNCATE=1, the default)
    give the same values as running every line (0): those lines carry only
    modes that the dealiasing mask removes (the nonlinear term) or that are
    zero (v_hat, and u_hat when the state has no mode outside the mask).
    With `outside` the state carries small modes outside the mask at
    kz <= N/3, which the state check must see (the u, omega transforms then
    move every entry).  Compared as values: a
    skipped zero may differ in sign."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c1").replace(precision=prec, N=N, n_particles=20000)
    rng = np.random.default_rng(5)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, _ = ic.initial_droplets(cfg, rng)
    uk0 = uk
    if outside:
        # small modes outside the mask only at kz <= N/3 (the lines the y passes keep)
        kc = N // 3
        k = np.fft.fftfreq(N, 1.0 / N).astype(int)
        k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + np.arange(N // 2 + 1)[None, None, :] ** 2
        sel = (k2 > kc * (kc + 1)) & (np.arange(N // 2 + 1)[None, None, :] <= kc)
        uk = uk.copy()
        noise = rng.standard_normal(uk.shape) + 1j * rng.standard_normal(uk.shape)
        uk = np.where(sel[None], 1e-3 * np.abs(uk).max() * noise, uk)
    outs = []
    for trunc in ("0", "1"):
        monkeypatch.setenv("LD_TRUNCATE", trunc)
        d = ForcedDNS(cfg, graph=True)
        if outside and trunc == "1":
            # a step graph captured for a band-limited state must not be replayed for this one
            d.set_state(_t(uk0, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
            d.advance(dt, 1)
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        d.advance(dt, 3)
        d.get_state(*t)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in t])
        d.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ensemble_matches_single_runs(gpu, prec):
    """A batched ensemble (one launch per kernel for all realizations) evolves
    each realization bit for bit as a context holding it alone."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c1").replace(precision=prec, N=32, n_particles=3000)
    B, steps = 3, 4
    uks, ns, cs, dts = [], [], [], []
    for b in range(B):
        rng = np.random.default_rng(100 + b)
        uk, n, u = ic.initial_flow(cfg, rng)
        dts.append(ic.cfl_dt(cfg, u))
        coord, _ = ic.initial_droplets(cfg, rng)
        uks.append(uk); ns.append(n); cs.append(coord)
    dt = min(dts)
    e = ForcedDNS(cfg, batch=B)
    uke = _t(np.stack(uks), gpu, e.cdtype)
    ne = _t(np.stack(ns), gpu, e.rdtype)
    ce = _t(np.stack(cs), gpu, torch.float64)
    e.set_state(uke, ne, ce)
    e.advance(dt, steps)
    e.get_state(uke, ne, ce)
    lm, bl = e.status_all()
    torch.cuda.synchronize()
    assert lm.shape == (B,) and not bl.any()
    with pytest.raises(Exception):
        e.forcing(uke[0], torch.zeros_like(uke[0]))        # single-realization entry points refuse
    e.close()
    for b in range(B):
        d = ForcedDNS(cfg)
        ukg, ng, cg = _t(uks[b], gpu, d.cdtype), _t(ns[b], gpu, d.rdtype), _t(cs[b], gpu, torch.float64)
        d.set_state(ukg, ng, cg)
        d.advance(dt, steps)
        d.get_state(ukg, ng, cg)
        torch.cuda.synchronize()
        for a, ref in ((uke[b], ukg), (ne[b], ng), (ce[b], cg)):
            assert np.array_equal(a.cpu().numpy().view(np.uint8), ref.cpu().numpy().view(np.uint8))
        d.close()


def test_ensemble_truncation_keeps_results(gpu, monkeypatch):
    """A batched ensemble whose realizations alternate between a state with no
    mode outside the mask and one with such modes: the state check covers every
    realization, the zeroed RK buffers are per realization, and the steps give
    the same values as with the truncation off (graph replay, set twice)."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c1").replace(precision="f32", N=32, n_particles=2000)
    B, steps = 3, 3
    uks, ns, cs, dts = [], [], [], []
    for b in range(B):
        rng = np.random.default_rng(200 + b)
        uk, n, u = ic.initial_flow(cfg, rng)
        dts.append(ic.cfl_dt(cfg, u))
        coord, _ = ic.initial_droplets(cfg, rng)
        uks.append(uk); ns.append(n); cs.append(coord)
    dt = min(dts)
    outside = np.stack(uks).copy()
    outside[1, :, 10, 10, 2] = 1e-3 * np.abs(outside).max()   # |k|^2 = 204 > 10 * 11: outside the mask
    outs = []
    for trunc in ("0", "1"):
        monkeypatch.setenv("LD_TRUNCATE", trunc)
        e = ForcedDNS(cfg, batch=B, graph=True)
        res = []
        for start in (np.stack(uks), outside):
            t = (_t(start, gpu, e.cdtype), _t(np.stack(ns), gpu, e.rdtype), _t(np.stack(cs), gpu, torch.float64))
            e.set_state(*t)
            e.advance(dt, steps)
            e.get_state(*t)
            torch.cuda.synchronize()
            res.append([x.cpu().numpy() for x in t])
        e.close()
        outs.append(res)
    for r0, r1 in zip(outs[0], outs[1]):
        for a, b in zip(r0, r1):
            assert np.array_equal(a, b)


def test_metric_workload_fp32_against_fp64_and_invariants(gpu):
    """The metric workload itself (256^3 FP32 fields + 10^6 droplets), where the
    CPU oracle is too slow: two fused steps in FP32 against the same steps of the
    FP64 engine (validated against the oracle at small sizes) within the FP32
    tolerance 1e-4, and size-independent properties of the FP32 result:
    solenoidal u_hat, Hermitian kz = 0 and N/2 planes, n >= 0, finite rows in
    input order."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c2")
    rng = np.random.default_rng(cfg.seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, pid = ic.initial_droplets(cfg, rng)
    del u
    res = {}
    for prec in ("f32", "f64"):
        d = ForcedDNS(cfg.replace(precision=prec))
        ukg, ng, cg = _t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64)
        d.set_state(ukg, ng, cg)
        d.advance(dt, 2)
        d.get_state(ukg, ng, cg)
        torch.cuda.synchronize()
        assert not bool(d.status()[1])
        d.close()
        res[prec] = (ukg.to(torch.complex128), ng.to(torch.float64), cg)

    def rel(a, b):
        return float((a - b).abs().max() / b.abs().max())
    u32, n32, c32 = res["f32"]
    u64, n64, c64 = res["f64"]
    assert rel(u32, u64) < 1e-4
    assert rel(n32, n64) < 1e-4
    assert rel(c32, c64) < 1e-4
    assert bool(torch.isfinite(c32).all()) and bool(torch.isfinite(u32.abs()).all())
    assert float(n32.min()) >= 0.0
    N = cfg.N
    k = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=gpu)
    kz = torch.arange(N // 2 + 1, dtype=torch.float64, device=gpu)
    div = (k[:, None, None] * u32[0] + k[None, :, None] * u32[1] + kz[None, None, :] * u32[2]).abs().max()
    kmag = torch.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    assert float(div) < 1e-5 * float((kmag * u32.abs().amax(dim=0)).max())
    for iz in (0, N // 2):
        pl = u32[:, :, :, iz]
        partner = torch.roll(torch.flip(pl, dims=(1, 2)), shifts=(1, 1), dims=(1, 2))
        assert float((pl - partner.conj()).abs().max()) <= 1e-6 * float(u32.abs().max())


def _c1_state(gpu, prec, npart, seed=11):
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    cfg = preset("c1").replace(precision=prec, n_particles=npart)
    rng = np.random.default_rng(seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    coord, _ = ic.initial_droplets(cfg, rng)
    return cfg, uk, n, coord, ic.cfl_dt(cfg, u)


def test_graph_replays_match_eager(gpu):
    """advance() through captured HIP graphs equals eager launches bit for bit,
    replay after replay (odd step counts alternate the droplet buffers)."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg, uk, n, coord, dt = _c1_state(gpu, "f32", 8000)
    ctxs = [ForcedDNS(cfg), ForcedDNS(cfg, graph=True)]
    st = []
    for d in ctxs:
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        st.append(t)
    for k, nsteps in enumerate((1, 2, 3, 1)):
        for d, t in zip(ctxs, st):
            d.advance(dt, nsteps)
            d.get_state(*t)
        torch.cuda.synchronize()
        for a, b in zip(st[0], st[1]):
            assert np.array_equal(a.cpu().numpy().view(np.uint8), b.cpu().numpy().view(np.uint8)), f"replay {k}"
        x = torch.ones(1 << 20, device=gpu).sum()        # eager work between replays
        assert float(x) == float(1 << 20)
    for d in ctxs:
        d.close()


def test_autotune_keeps_results(gpu):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg, uk, n, coord, dt = _c1_state(gpu, "f32", 4000)
    outs = []
    for tune in (False, True):
        d = ForcedDNS(cfg, autotune=tune)
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        d.advance(dt, 2)
        d.get_state(*t)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in t])
        if tune:
            yl, us = d.tuning()
            assert yl in (16, 8, 32) and min(us[:2]) > 0
        d.close()
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_step_queue_orders_give_the_same_bits(gpu):
    """At N >= 128, where a stage's density chain is queued after the next
    stage's x pass, one call of 3 steps, 3 calls of one step and HIP-graph
    replay give the same state bit for bit."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(128, "f32", 20000, nu=0.01)
    rng = np.random.default_rng(41)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, _ = ic.initial_droplets(cfg, rng)
    outs = []
    for mode in ("one_call", "per_step", "graph"):
        d = ForcedDNS(cfg, graph=(mode == "graph"))
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        if mode == "one_call":
            d.advance(dt, 3)
        else:
            for _ in range(3):
                d.advance(dt, 1)
        d.get_state(*t)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in t])
        d.close()
    assert np.all(np.isfinite(outs[0][0].view(np.float32)))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("N,prec", [(64, "f32"), (32, "f64"), (256, "f32")])
def test_every_launch_geometry_gives_the_same_bits(gpu, N, prec, monkeypatch):
    """The z-tile heights and x-pass line counts the autotuner chooses between
    (defaults, and all halved at once) give bit-identical steps; the autotuner
    reports a choice and a time for every candidate it ran."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(N, prec, 3000 if N < 256 else 20000, nu=0.01)
    rng = np.random.default_rng(31)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, _ = ic.initial_droplets(cfg, rng)
    outs = []
    for force in ("0,0,0,0,0", "1,1,1,1,1", "2,2,0,0,0"):
        monkeypatch.setenv("LD_TUNE_FORCE", force)
        d = ForcedDNS(cfg)
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        d.advance(dt, 2)
        d.get_state(*t)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in t])
        d.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    monkeypatch.delenv("LD_TUNE_FORCE")
    d = ForcedDNS(cfg, autotune=True)
    choice, times = d.tuning_all()
    assert choice["y_lines"] in (4, 8, 16, 32)
    for k in ("z_cross_tile", "z_vel_tile", "x_uw_lines", "x_rhs_lines", "x_kn_lines"):
        assert choice[k] >= 2 and times[k][0] > 0
    assert choice["y_split_n1"] in (4, 8, 16) and times["y_half_split_at_chosen_lines"][0] > 0
    d.close()


@pytest.mark.parametrize("N,prec", [(64, "f32"), (32, "f64"), (256, "f32")])
def test_y_pass_other_line_split_matches_default(gpu, N, prec, monkeypatch):
    """The y passes with the other line split (N1/2 threads per line, each
    holding 2 N2 elements; LD_YSPLIT=1) give the default split's two steps to
    rounding (another factorisation: not bit for bit): 1e-13 relative in norm
    for FP64, 2e-6 for FP32, on u_hat, n and the droplet rows."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = _cfg(N, prec, 3000 if N < 256 else 20000, nu=0.01)
    rng = np.random.default_rng(32)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, _ = ic.initial_droplets(cfg, rng)
    outs = []
    for split in ("0", "1"):
        monkeypatch.setenv("LD_YSPLIT", split)
        d = ForcedDNS(cfg)
        choice, _ = d.tuning_all()
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        d.advance(dt, 2)
        d.get_state(*t)
        torch.cuda.synchronize()
        outs.append(([x.cpu().numpy() for x in t], choice["y_split_n1"]))
        d.close()
    monkeypatch.delenv("LD_YSPLIT")
    (a, n1a), (b, n1b) = outs
    assert n1b * 2 == n1a
    tol = 1e-13 if prec == "f64" else 2e-6
    for x, y in zip(a, b):
        assert relerr(y, x) < tol


def test_c1_200_steps_against_oracle_golden(gpu):
    """The 64^3 FP64 workload (10^4 droplets, 200 steps) against the committed
    CPU-oracle run (tests/golden/make_c1_golden.py): 1e-10 relative on the
    sampled spectral modes, density points and droplet rows (by id)."""
    import os
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_oracle_200.npz"))
    cfg = preset("c1")
    rng = np.random.default_rng(cfg.seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    assert dt == float(g["dt"])
    coord, pid = ic.initial_droplets(cfg, rng)
    d = ForcedDNS(cfg)
    ukg = _t(uk, gpu, d.cdtype)
    ng = _t(n, gpu, d.rdtype)
    cg = _t(coord, gpu, torch.float64)
    d.set_state(ukg, ng, cg)
    d.advance(dt, int(g["steps"]))
    d.get_state(ukg, ng, cg)
    torch.cuda.synchronize()
    uk1, n1, c1 = ukg.cpu().numpy(), ng.cpu().numpy(), cg.cpu().numpy()
    mi, ni = g["mode_ix"], g["n_ix"]
    e_u = np.max(np.abs(uk1[:, mi[0], mi[1], mi[2]] - g["uk_modes"])) / float(g["uk_absmax"])
    e_n = np.max(np.abs(n1[ni[0], ni[1], ni[2]] - g["n_pts"])) / float(g["n_absmax"])
    rows = c1[g["droplet_ids"]]            # rows come back in input order = id order
    ref = g["droplet_rows"]
    e_x = min_image(rows, ref, cfg.L) / cfg.L
    e_v = relerr(rows[:, 3:6], ref[:, 3:6])
    e_m = relerr(rows[:, 6], ref[:, 6])
    print(f"c1 200 steps: uk {e_u:.2e} n {e_n:.2e} x {e_x:.2e} v {e_v:.2e} m {e_m:.2e}")
    assert max(e_u, e_n, e_x, e_v, e_m) < 1e-10
    ek = d.spectrum(ukg, 0.5).cpu().numpy()
    assert relerr(ek, g["ek"]) < 1e-10
    d.close()


# ------------------------------------------------------------ learned subgrid form
def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("N, ch", [(8, (3, 16, 32, 3)), (16, (3, 32, 32, 16, 3)), (32, (3, 16, 32, 32, 3))])
def test_cnn_forward_and_gradients_match_oracle(gpu, N, ch):
    """bf16 MFMA forward and backward against the CPU oracle with the same
    bfloat16 rounding points (FP32 vs exact accumulation left): 1e-2 / 3e-2
    relative in norm.  N = 8 runs the 16x16x16 kernels; N >= 16 the 16x16x32
    convolution (every channel pair 16/32, mirrored taps with the ReLU mask)
    and the node-blocked weight gradient, z-line segments of 16 and 32 nodes."""
    import torch
    from lucky_droplets_amd.cnn import SubgridCNN
    from oracle import cnn as ocnn
    net = SubgridCNN(N, ch)
    rng = np.random.default_rng(5)
    layers = [(rng.normal(0, np.sqrt(2.0 / (27 * ci)), (27, co, ci)).astype(np.float32),
               rng.normal(0, 0.1, co).astype(np.float32)) for ci, co in zip(ch[:-1], ch[1:])]
    net.set_params(layers)
    x = rng.normal(size=(N, N, N, ch[0])).astype(np.float32)
    t = rng.normal(size=(N, N, N, ch[-1])).astype(np.float32)
    xg, tg = torch.from_numpy(x).to(gpu), torch.from_numpy(t).to(gpu)
    y = net.forward(xg).cpu().numpy().reshape(N, N, N, ch[-1])
    yr, _ = ocnn.forward(layers, x, bf16=True)
    assert _rel(y, yr) < 1e-2
    loss, g = net.grad(xg, tg)
    lr, gr = ocnn.loss_and_grads(layers, x, t, bf16=True)
    assert abs(loss - lr) < 1e-2 * lr
    gl = net.unpack(g.cpu().numpy())
    for l in range(len(layers)):
        assert _rel(gl[l][0], gr[l][0]) < 3e-2, l
        assert _rel(gl[l][1], gr[l][1]) < 3e-2, l
    net.close()


def _deconv_data(seed):
    """Coarse 32^3 input (the (1/4, 1/2, 1/4)-filtered 64^3 field at the coarse
    nodes) and the fine field there."""
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from oracle import cnn as ocnn
    cfg = preset("c1").replace(N=64)
    _, _, u = ic.initial_flow(cfg, np.random.default_rng(seed))
    fine = np.ascontiguousarray(u.transpose(1, 2, 3, 0))
    return ocnn.box_filter(fine), fine[::2, ::2, ::2]


def test_cnn_training_beats_plain_filtered_field(gpu):
    """Trained on the GPU by its own forward/backward kernels, the network
    reconstructs the fine field at the coarse nodes from the filtered coarse
    field; on a held-out realization its residual error is well below that of
    the plain coarse field (which is the network with zero output)."""
    import torch
    from lucky_droplets_amd.cnn import SubgridCNN
    xs, ts = [], []
    for seed in (100, 101, 102):
        cb, fn = _deconv_data(seed)
        xs.append(cb); ts.append(fn - cb)
    sx = float(np.std(np.concatenate(xs)))
    sr = float(np.std(np.concatenate(ts)))
    net = SubgridCNN(32, (3, 16, 16, 3))
    net.init_params(seed=0, gain=0.1)
    data = [(torch.from_numpy((x / sx).astype(np.float32)).to(gpu), torch.from_numpy((t / sr).astype(np.float32)).to(gpu))
            for x, t in zip(xs, ts)]
    losses = []
    for it in range(300):
        x, t = data[it % len(data)]
        losses.append(net.train_step(x, t, lr=3e-3))
    assert np.mean(losses[-30:]) < 0.5 * np.mean(losses[:10])
    cb, fn = _deconv_data(131)
    xv = torch.from_numpy((cb / sx).astype(np.float32)).to(gpu)
    pred = net.forward(xv).cpu().numpy().reshape(cb.shape) * sr
    err_cnn = float(np.mean((fn - cb - pred) ** 2))
    err_plain = float(np.mean((fn - cb) ** 2))
    print(f"held-out mean squared error: plain {err_plain:.3e}, with the network {err_cnn:.3e}")
    assert err_cnn < 0.5 * err_plain
    net.close()
