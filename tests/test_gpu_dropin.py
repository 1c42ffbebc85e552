"""The drop-in droplet entry points on their own: ld_interp_cosine,
ld_exterp_cosine_scalar and ld_prhs (with its interpmat / exterpmat outputs)
against the oracle's MPI_particles methods (particles.py:435-577, 817-837),
and the blow-up error of the stepping calls (forced-dns-sm.py:618-620).

Rows are compared by prtclid (the reference reorders rows in send)."""
import numpy as np
import pytest

from helpers import by_id, inputs, make_drops_oracle, make_flow_oracle, relerr
from lucky_droplets_amd.config import Config

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 2e-5}


def _t(x, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


def _cfg(N, prec, nparts):
    return Config(N=N, precision=prec, nu=0.02, nshells=2, profile="k4", n_particles=nparts,
                  st_dist=("uniform", 0.05, 1.0), gravity=True)


def _rows(cfg, n, seed, outside=False):
    """Droplet rows with positions in [0, L) (or, with outside, a third of them
    shifted by -L or +L so that send must wrap and reorder them)."""
    rng = np.random.default_rng(seed)
    _, _, coord, pid = inputs(cfg, seed=seed, n_particles=n)
    coord[:, 3:6] = rng.standard_normal((n, 3))
    if outside:
        k = rng.integers(0, 3, n)
        coord[:, 0] += np.where(k == 1, -cfg.L, np.where(k == 2, cfg.L, 0.0))
    return coord, pid


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("N", [16, 32])
def test_interp_cosine_matches_oracle(gpu, prec, N):
    """ld_interp_cosine of 7 fields at 500 rows (particles.py:435-511)."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    cfg = _cfg(N, prec, 500)
    coord, pid = _rows(cfg, 500, seed=N)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((7, N, N, N))
    dr = make_drops_oracle(cfg, coord, pid)
    ref = dr.interp_cosine(coord, f)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, N, 500, cfg.sts, 0.1, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0, dns=d,
                        coord=coord, prtclid=pid)
    fg = _t(f, gpu, d.rdtype)
    out = stb.interp_cosine(stb.coord, fg).cpu().numpy()
    assert out.shape == (500, 7)
    assert relerr(out, ref) < TOL[prec]
    assert np.array_equal(stb.interpmat.cpu().numpy(), out)
    # positions shifted by whole periods gather the same values (the kernel wraps them)
    sh = coord.copy()
    sh[:, :3] += cfg.L * rng.integers(-2, 3, (500, 3))
    out2 = stb.interp_cosine(_t(sh, gpu, torch.float64), fg).cpu().numpy()
    assert relerr(out2, ref) < (1e-9 if prec == "f64" else 1e-4)
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("N", [16, 32])
def test_exterp_cosine_scalar_matches_oracle(gpu, prec, N):
    """ld_exterp_cosine_scalar: c zeroed, then the deposit of exterpmat / dx^3
    (particles.py:514-577); a field holding garbage is overwritten."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    cfg = _cfg(N, prec, 800)
    coord, pid = _rows(cfg, 800, seed=10 + N)
    rng = np.random.default_rng(4)
    w = rng.random(800) + 0.1
    dr = make_drops_oracle(cfg, coord, pid)
    dr.exterpmat[:, 0] = w
    ref = dr.exterp_cosine_scalar(coord, np.ones((N, N, N)))
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, N, 800, cfg.sts, 0.1, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0, dns=d,
                        coord=coord, prtclid=pid)
    stb.exterpmat = _t(w[:, None], gpu, torch.float64)
    c = torch.full((N, N, N), 7.0, dtype=d.rdtype, device=gpu)
    stb.exterp_cosine_scalar(stb.coord, c)
    cg = c.cpu().numpy()
    assert relerr(cg, ref) < TOL[prec]
    # mass: sum c dx^3 = sum w
    dx = cfg.L / N
    assert abs(cg.astype(np.float64).sum() * dx ** 3 - w.sum()) < (1e-11 if prec == "f64" else 1e-4) * w.sum()
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("outside", [False, True])
def test_prhs_outputs_match_oracle(gpu, prec, outside):
    """ld_prhs: rhs, interpmat, exterpmat and fc against the oracle's pRHS
    (particles.py:817-837) by prtclid, with rows inside [0, L) and with rows
    that send must wrap and reorder."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    N, npart = 32, 600
    cfg = _cfg(N, prec, npart)
    uk, n, _, _ = inputs(cfg, seed=5, n_noise=0.5)
    coord, pid = _rows(cfg, npart, seed=6, outside=outside)
    fl = make_flow_oracle(cfg)
    G = fl.grid
    u = np.stack([G.irfft_mpi(uk[c]) for c in range(3)])
    rng = np.random.default_rng(8)
    us = u + 0.1 * rng.standard_normal(u.shape)
    n = np.abs(n)
    dr = make_drops_oracle(cfg, coord, pid)
    sump = coord.copy()
    fc_ref = np.zeros((N, N, N))
    _, rhs_ref, fc_ref = dr.pRHS(0.0, coord.copy(), u, us, n, fc_ref, sump)
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, N, npart, cfg.sts, 0.1, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0, dns=d,
                        coord=coord, prtclid=pid)
    fc = torch.full((N, N, N), -3.0, dtype=d.rdtype, device=gpu)
    _, rhs, fc = stb.pRHS(0.0, stb.coord, _t(u, gpu, d.rdtype), _t(us, gpu, d.rdtype), _t(n, gpu, d.rdtype), fc,
                          None)
    tol = TOL[prec] if prec == "f64" else 1e-4
    assert relerr(fc.cpu().numpy(), fc_ref) < tol
    got = {"rhs": rhs.cpu().numpy(), "interp": stb.interpmat.cpu().numpy(), "ext": stb.exterpmat.cpu().numpy()}
    ref = {"rhs": rhs_ref, "interp": dr.interpmat, "ext": dr.exterpmat}
    for k in got:
        g_id = by_id(got[k], pid)
        r_id = by_id(ref[k], dr.prtclid)
        assert g_id.shape == r_id.shape
        cols = [slice(0, 3), slice(3, 6), slice(6, 7)] if g_id.shape[1] == 7 else [slice(0, 1)]
        for cs in cols:
            assert relerr(g_id[:, cs], r_id[:, cs]) < tol, k
    # the stage St the engine used is the row's (particles.py:823)
    st_ref = by_id((coord[:, 6] / dr.factor) ** (2 / 3), pid)
    assert relerr(by_id(stb.st.cpu().numpy(), pid), st_ref) < 1e-14
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_blowup_is_reported(gpu, prec):
    """A state whose largest Re(u_hat) exceeds 100 N^3 after a step makes
    ld_step / ld_advance return LD_ERR_BLOWUP (Python: BlowUp), as the
    reference aborts (forced-dns-sm.py:618-620); a normal state does not."""
    import torch
    from lucky_droplets_amd import _lib
    from lucky_droplets_amd.dns import ForcedDNS
    N = 16
    cfg = _cfg(N, prec, 0)
    uk, n, _, _ = inputs(cfg, seed=9)
    d = ForcedDNS(cfg, max_particles=0)
    ukg = _t(uk, gpu, d.cdtype)
    ng = _t(n, gpu, d.rdtype)
    d.step(0.0, 1e-3, None, ukg, ng)
    m, blown = d.status()
    assert not blown and 0 < m < 100 * N ** 3
    # one transverse mode far above the threshold: (kx, ky, kz) = (0, 0, 1), u_x
    big = uk.copy()
    big[0, 0, 0, 1] = 1e3 * N ** 3
    ukg = _t(big, gpu, d.cdtype)
    with pytest.raises(_lib.BlowUp):
        d.step(0.0, 1e-4, None, ukg, _t(n, gpu, d.rdtype))
    m, blown = d.status()
    assert blown and m > 100 * N ** 3
    # the resident-state path reports it too
    d2 = ForcedDNS(cfg, max_particles=0)
    d2.set_state(_t(big, gpu, d2.cdtype), _t(n, gpu, d2.rdtype))
    with pytest.raises(_lib.BlowUp):
        d2.advance(1e-4, 2)
    d.close()
    d2.close()
