"""The engine at the large grids: 256^3 and 512^3.

* 3-D transforms at N = 256 and 512 against scipy.fft (the reference's FFT
  library, forced-dns-sm.py:200-213).
* The 256^3 + 10^6-droplet workload against a sampled golden of the CPU FP64
  oracle (tests/golden/make_metric_golden.py): one full_RHS
  (forced-dns-sm.py:287-332) and one fused step (:335-362, 582-620), FP64 at
  1e-10 and FP32 at 1e-4 relative (error over the field's largest magnitude).
* The 512^3 + 10^7-droplet workload (bidisperse St 0.2 / 0.6, nu = 1e-3):
  two steps in FP32 and FP64 with property checks (FP32 against FP64,
  solenoidal velocity, Hermitian kz = 0 and N/2 planes, n >= 0, and the
  conserved total of small-droplet mass and droplet mass).
"""
import math
import os

import numpy as np
import pytest

from helpers import make_flow_oracle, min_image
from lucky_droplets_amd.config import Config, preset

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "metric_oracle_256.npz")


def _t(x, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("N", [256, 512])
def test_fft_large_matches_scipy(gpu, N, prec):
    from lucky_droplets_amd.dns import ForcedDNS
    import oracle.flow as oflow
    oflow.WORKERS = min(16, os.cpu_count() or 1)
    cfg = Config(N=N, precision=prec, nu=1e-3, nshells=2)
    d = ForcedDNS(cfg, max_particles=0)
    G = make_flow_oracle(cfg).grid
    rng = np.random.default_rng(N + 1)
    u = rng.standard_normal((N, N, N))
    ref = G.rfft_mpi(u)
    fu = d.zeros_spec()
    d.rfft_mpi(_t(u, gpu, d.rdtype), fu)
    tol = 1e-13 if prec == "f64" else 3e-6
    assert np.abs(fu.cpu().numpy() - ref).max() / np.abs(ref).max() < tol
    del ref, fu
    X = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    refi = G.irfft_mpi(X)
    uo = d.zeros_real()
    d.irfft_mpi(_t(X, gpu, d.cdtype), uo)
    assert np.abs(uo.cpu().numpy() - refi).max() / np.abs(refi).max() < tol
    d.close()


def _metric_inputs():
    from lucky_droplets_amd import ic
    cfg = preset("c2")
    rng = np.random.default_rng(cfg.seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    del u
    coord, pid = ic.initial_droplets(cfg, rng)
    return cfg, uk, n, coord, pid


@pytest.fixture(scope="module")
def metric_case():
    if not os.path.exists(GOLDEN):
        pytest.skip("golden fixture not generated")
    g = dict(np.load(GOLDEN))
    return g, _metric_inputs()


def _err(got, g, name):
    return float(np.abs(got - g[name]).max() / g[name + "_absmax"])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_metric_full_rhs_against_oracle_golden(gpu, metric_case, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    g, (cfg, uk, n, coord, pid) = metric_case
    cfg = cfg.replace(precision=prec)
    tol = 1e-10 if prec == "f64" else 1e-4
    d = ForcedDNS(cfg)
    u_out, v_out, fc = d.zeros_real(3), d.zeros_real(3), d.zeros_real()
    ku, kn, kp, _ = d.full_RHS(0.0, _t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), None,
                               _t(coord, gpu, torch.float64), fc=fc, u_out=u_out, v_out=v_out)
    si, pi = tuple(g["spec_ix"]), tuple(g["phys_ix"])
    assert _err(ku.cpu().numpy()[si], g, "rhs_ku") < tol
    assert _err(kn.cpu().numpy()[pi[1:]], g, "rhs_kn") < tol
    assert _err(u_out.cpu().numpy()[pi], g, "rhs_u") < tol
    assert _err(v_out.cpu().numpy()[pi], g, "rhs_v") < tol
    assert _err(fc.cpu().numpy()[pi[1:]], g, "rhs_fc") < tol
    kpg = kp.cpu().numpy()[g["droplet_ids"]]      # rows keep their input order (id = row)
    for cols in (slice(0, 3), slice(3, 6), slice(6, 7)):
        e = np.abs(kpg[:, cols] - g["rhs_kp"][:, cols]).max() / g["rhs_kp_absmax"][cols].max()
        assert e < tol
    d.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_metric_step_against_oracle_golden(gpu, metric_case, prec):
    """One fused step (sorted droplets, brick deposit, dealiasing truncation)
    against the oracle's step."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    g, (cfg, uk, n, coord, pid) = metric_case
    cfg = cfg.replace(precision=prec)
    tol = 1e-10 if prec == "f64" else 1e-4
    d = ForcedDNS(cfg)
    d.set_state(_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
    d.advance(float(g["dt"]), 1)
    ukg, ng, cg = d.zeros_spec(3), d.zeros_real(), torch.empty_like(_t(coord, gpu, torch.float64))
    d.get_state(ukg, ng, cg)
    si, pi = tuple(g["spec_ix"]), tuple(g["phys_ix"])
    assert _err(ukg.cpu().numpy()[si], g, "step_uk") < tol
    assert _err(ng.cpu().numpy()[pi[1:]], g, "step_n") < tol
    rows = cg.cpu().numpy()
    sel = rows[g["droplet_ids"]]
    ref = g["step_rows"]
    assert min_image(sel, ref, cfg.L) < (1e-10 if prec == "f64" else 1e-5)
    for cols in (slice(3, 6), slice(6, 7)):
        assert np.abs(sel[:, cols] - ref[:, cols]).max() / np.abs(ref[:, cols]).max() < tol
    assert abs(rows[:, 6].sum() - g["step_mass"]) < tol * g["step_mass"]
    assert abs(float(ng.double().mean()) - g["step_n_mean"]) < tol
    d.close()


# ---------------------------------------------------------------- 512^3 + 10^7 droplets
def _gpu_initial_flow(cfg, seed, dev):
    """The initial-condition recipe of lucky_droplets_amd.ic (forced-dns-sm.py:691-737,
    567-569) built on the GPU with torch.fft, as test input at 512^3; returns
    (uk complex128, n float64, dt)."""
    import torch
    N, Nf = cfg.N, cfg.N // 2 + 1
    kx = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=dev)
    kz = torch.arange(Nf, dtype=torch.float64, device=dev)
    KX, KY, KZ = kx[:, None, None], kx[None, :, None], kz[None, None, :]
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    kint = torch.clamp(torch.round(torch.sqrt(k2)), max=N // 2)
    dealias = kint <= N / 3
    norm = torch.where((KZ == 0) | (KZ == N // 2), 1.0, 2.0) * (2 * math.pi) ** 3 / float(N) ** 6
    epro = kint ** 4 * torch.exp(-2.0 * (kint / 3.0) ** 2) / norm
    amp = torch.sqrt(epro / torch.where(kint == 0, torch.inf, kint ** 2))
    sel = (kint ** 2 < cfg.kinit ** 2) & (kint > 0) & dealias
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    uk = torch.empty((3, N, N, Nf), dtype=torch.complex128, device=dev)
    for c in range(3):
        th = torch.rand((N, N, Nf), generator=gen, dtype=torch.float64, device=dev) * (2 * math.pi)
        uk[c] = amp * torch.polar(torch.ones_like(th), th) * sel
    del th, epro, amp
    inv = torch.where(k2 == 0, 0.0, 1.0 / k2) * dealias
    for it in range(2):
        u = torch.fft.irfftn(uk, s=(N, N, N), dim=(1, 2, 3))
        uk = torch.fft.rfftn(u, dim=(1, 2, 3))
        del u
        if it == 0:
            trm = KX * uk[0] + KY * uk[1] + KZ * uk[2]
            uk[0] -= inv * KX * trm
            uk[1] -= inv * KY * trm
            uk[2] -= inv * KZ * trm
            del trm
            e0 = float((0.5 * (uk.abs() ** 2).sum(0) * norm).sum())
            uk *= math.sqrt(cfg.einit / e0)
    u = torch.fft.irfftn(uk, s=(N, N, N), dim=(1, 2, 3))
    dt = cfg.cfl * cfg.dx / float(u.abs().sum(0).max())
    del u
    uk = uk * dealias
    return uk.contiguous(), torch.ones((N, N, N), dtype=torch.float64, device=dev), dt


def _props(uk, N):
    """(divergence over |k||u_hat| scale, Hermitian defect of the kz = 0 and N/2 planes)."""
    import torch
    dev = uk.device
    kx = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=dev)
    kz = torch.arange(N // 2 + 1, dtype=torch.float64, device=dev)
    ukd = uk.to(torch.complex128)
    div = kx[:, None, None] * ukd[0] + kx[None, :, None] * ukd[1] + kz[None, None, :] * ukd[2]
    scale = float(ukd.abs().max()) * N
    dv = float(div.abs().max()) / scale
    del div
    herm = 0.0
    for iz in (0, N // 2):
        P = ukd[:, :, :, iz]
        Q = torch.conj(torch.roll(torch.flip(P, dims=(1, 2)), shifts=(1, 1), dims=(1, 2)))
        herm = max(herm, float((P - Q).abs().max()) / float(ukd.abs().max()))
    return dv, herm


def test_c4_512_droplets_fp32_against_fp64_and_invariants(gpu):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c4")
    N, npart = cfg.N, cfg.n_particles
    uk, n, dt = _gpu_initial_flow(cfg, 0, gpu)
    gen = torch.Generator(device=gpu)
    gen.manual_seed(1)
    coord = torch.rand((npart, 7), generator=gen, dtype=torch.float64, device=gpu) * cfg.L
    st = torch.where(torch.rand(npart, generator=gen, dtype=torch.float64, device=gpu) < 0.5, 0.2, 0.6)
    coord[:, 6] = st ** 1.5 * cfg.factor
    dx3 = cfg.dx ** 3
    total0 = float(n.sum()) * dx3 + float(coord[:, 6].sum())
    mass0 = float(coord[:, 6].sum())
    out = {}
    for prec in ("f64", "f32"):
        c = cfg.replace(precision=prec)
        d = ForcedDNS(c)
        d.set_state(uk.to(d.cdtype).contiguous(), n.to(d.rdtype).contiguous(), coord)
        d.advance(dt, 2)
        ukg, ng = d.zeros_spec(3), d.zeros_real()
        cg = torch.empty_like(coord)
        d.get_state(ukg, ng, cg)
        d.close()
        del d
        out[prec] = (ukg, ng, cg)
        dv, herm = _props(ukg, N)
        tol = 1e-13 if prec == "f64" else 1e-6
        assert dv < tol and herm < tol, (prec, dv, herm)
        assert float(ng.min()) >= 0.0
        mass = float(cg[:, 6].sum())
        total = float(ng.double().sum()) * dx3 + mass
        assert mass > mass0                      # droplets collect small-droplet mass
        # sum(n) dx^3 + sum(m) is invariant: kn has mean -mean(fc) and sum(fc) dx^3 = sum(W)
        # (forced-dns-sm.py:325-330, particles.py:828-835); clip_zero keeps the mean
        assert abs(total - total0) < (1e-9 if prec == "f64" else 2e-3) * (mass - mass0), (prec, total - total0,
                                                                                           mass - mass0)
    (u64, n64, c64), (u32, n32, c32) = out["f64"], out["f32"]
    assert float((u32.to(torch.complex128) - u64).abs().max()) / float(u64.abs().max()) < 1e-4
    assert float((n32.double() - n64).abs().max()) / float(n64.abs().max()) < 1e-4
    dpos = c32[:, :3] - c64[:, :3]
    dpos = dpos - cfg.L * torch.round(dpos / cfg.L)
    assert float(dpos.abs().max()) < 1e-5
    assert float((c32[:, 3:6] - c64[:, 3:6]).abs().max()) / float(c64[:, 3:6].abs().max()) < 1e-4
    assert float((c32[:, 6] - c64[:, 6]).abs().max()) / float(c64[:, 6].abs().max()) < 1e-4
