"""The learned subgrid form's kernels and pipeline.

* Spectral resampling between grids: a band-limited field resampled to a
  finer grid and transformed back is the same function sampled there
  (closed form); truncation is the inverse of padding on those modes.
* CNN row layouts: fields -> rows, subgrid residual -> compose round trip,
  and the channel order against a numpy reshape.
* Droplets in a frozen field: RK4 with the oracle's cosine-kernel gather
  (particles.py:435-511) and the drag/settling RHS (particles.py:831-835).
* The pipeline on a small case: trained on fine DNS states, the network's
  subgrid velocity beats the plain coarse field on a held-out state, and the
  three forms give finite pair statistics.
"""
import numpy as np
import pytest

from helpers import make_drops_oracle, relerr
from lucky_droplets_amd.config import Config

pytestmark = pytest.mark.gpu


def _t(x, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


def _grid(N):
    x = np.arange(N) * 2 * np.pi / N
    return np.meshgrid(x, x, x, indexing="ij")


def test_spectral_resample_closed_form(gpu):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    Nc, Nf = 16, 32
    fields = lambda x, y, z: np.stack([np.sin(x) * np.cos(2 * y) + np.cos(3 * z), np.sin(5 * y - 2 * x),
                                       np.cos(x + y + z) * np.sin(4 * z)])
    cc = Config(N=Nc, precision="f64", nu=0.01)
    cf = Config(N=Nf, precision="f64", nu=0.01)
    dc, df = ForcedDNS(cc, max_particles=0), ForcedDNS(cf, max_particles=0)
    uc = _t(fields(*_grid(Nc)), gpu, torch.float64)
    ukc = dc.rfft_mpi(uc, dc.zeros_spec(3))
    ukf = df.zeros_spec(3)
    df._stream()
    assert df.lib.ld_spectral_resample(df.ctx, ukc.data_ptr(), Nc, ukf.data_ptr(), 3) == 0
    uf = df.irfft_mpi(ukf, df.zeros_real(3)).cpu().numpy()
    assert np.max(np.abs(uf - fields(*_grid(Nf)))) < 1e-13
    # back to the coarse grid: the same spectrum
    back = dc.zeros_spec(3)
    dc._stream()
    assert dc.lib.ld_spectral_resample(dc.ctx, ukf.data_ptr(), Nf, back.data_ptr(), 3) == 0
    assert relerr(back.cpu().numpy(), ukc.cpu().numpy()) < 1e-14
    dc.close(); df.close()


def test_subgrid_rows_round_trip_and_layout(gpu):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    Nf, r = 32, 2
    Nc = Nf // r
    df = ForcedDNS(Config(N=Nf, precision="f32", nu=0.01), max_particles=0)
    rng = np.random.default_rng(3)
    fine = rng.standard_normal((3, Nf, Nf, Nf)).astype(np.float32)
    base = rng.standard_normal((3, Nf, Nf, Nf)).astype(np.float32)
    fg, bg = _t(fine, gpu, torch.float32), _t(base, gpu, torch.float32)
    net = torch.empty((Nc ** 3, 3 * r ** 3), dtype=torch.float32, device=gpu)
    df._stream()
    assert df.lib.ld_subgrid_residual(df.ctx, fg.data_ptr(), bg.data_ptr(), r, net.data_ptr(), 0.5) == 0
    # channel c r^3 + ((X%r) r + Y%r) r + Z%r of coarse node (X/r, Y/r, Z/r)
    d = (fine - base) * 0.5
    exp = d.reshape(3, Nc, r, Nc, r, Nc, r).transpose(1, 3, 5, 0, 2, 4, 6).reshape(Nc ** 3, 3 * r ** 3)
    assert np.max(np.abs(net.cpu().numpy() - exp)) < 1e-6
    out = df.zeros_real(3)
    assert df.lib.ld_subgrid_compose(df.ctx, net.data_ptr(), r, bg.data_ptr(), out.data_ptr(), 2.0) == 0
    assert np.max(np.abs(out.cpu().numpy() - fine)) < 1e-5
    rows = torch.empty((Nf ** 3, 3), dtype=torch.float32, device=gpu)
    assert df.lib.ld_fields_to_rows(df.ctx, fg.data_ptr(), 3, rows.data_ptr(), 2.0) == 0
    assert np.array_equal(rows.cpu().numpy(), (fine.reshape(3, -1).T * 2.0).astype(np.float32))
    assert df.lib.ld_subgrid_compose(df.ctx, net.data_ptr(), 3, bg.data_ptr(), out.data_ptr(), 1.0) == 1
    df.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_frozen_field_droplets_match_rk4_with_oracle_gather(gpu, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    N, npart, h, steps = 16, 300, 0.02, 3
    cfg = Config(N=N, precision=prec, nu=0.02, n_particles=npart, gravity=True)
    rng = np.random.default_rng(4)
    x, y, z = _grid(N)
    u = np.stack([np.sin(y) + 0.3 * np.cos(2 * z), np.cos(x) * np.sin(z), 0.5 * np.sin(x + y)])
    rows = rng.uniform(0, cfg.L, (npart, 7))
    rows[:, 3:6] = rng.standard_normal((npart, 3))
    rows[:, 6] = rng.uniform(0.1, 1.0, npart) ** 1.5 * cfg.factor
    dr = make_drops_oracle(cfg, rows, np.arange(npart))
    st = (rows[:, 6] / dr.factor) ** (2 / 3)
    tau = st * cfg.tf

    def rhs(xx, vv):
        I = dr.interp_cosine(np.mod(xx, cfg.L), u)
        a = (I - vv) / tau[:, None]
        a[:, 2] -= cfg.g
        return vv, a
    X, V = rows[:, :3].copy(), rows[:, 3:6].copy()
    for _ in range(steps):
        k1 = rhs(X, V)
        k2 = rhs(X + 0.5 * h * k1[0], V + 0.5 * h * k1[1])
        k3 = rhs(X + 0.5 * h * k2[0], V + 0.5 * h * k2[1])
        k4 = rhs(X + h * k3[0], V + h * k3[1])
        X = X + h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        V = V + h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
    d = ForcedDNS(cfg)
    rg = _t(rows, gpu, torch.float64)
    ug = _t(u, gpu, d.rdtype)
    d._stream()
    assert d.lib.ld_droplets_frozen_rk4(d.ctx, rg.data_ptr(), npart, ug.data_ptr(), h, steps) == 0
    out = rg.cpu().numpy()
    tol = 1e-12 if prec == "f64" else 1e-6
    assert relerr(out[:, :3], X) < tol and relerr(out[:, 3:6], V) < (1e-11 if prec == "f64" else 1e-5)
    assert np.array_equal(out[:, 6], rows[:, 6])
    d.close()


def test_untrained_model_falls_back_to_the_base(gpu):
    """A network that does no better than none on held-out data is not used:
    the untrained model (last layer zero) validates as a fallback and its
    reconstruction is the base alone, bit for bit."""
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.learned import SubgridModel
    cfg = Config(N=32, precision="f32", nu=0.02, nshells=2, profile="k4")
    m = SubgridModel(cfg)
    uk, _, _ = ic.initial_flow(cfg, np.random.default_rng(9))
    ukf = torch.from_numpy(uk).to(gpu).to(torch.complex64)
    e_plain, e_learned = m.validate(ukf)
    assert e_learned == e_plain and m.fallback
    ukc = m.coarse_spectrum(ukf)
    rec = m.reconstruct(ukc).clone()
    base = m.base(ukc).clone()
    assert torch.equal(rec, base)
    m.close()


def test_learned_form_collision_kernel_against_fine_dns(gpu):
    """128^3 fine DNS against the 64^3 coarse DNS with plain interpolation and
    with the learned reconstruction, 2 x 10^5 droplets (St 0.5): the learned
    form's collision kernel (summed over the separation bins up to 0.1) is
    within 5 % of the fine DNS's and closer than the plain coarse form's; on a
    held-out state the trained network's subgrid velocity has at least 5 %
    less mean squared error than none (the plain reconstruction)."""
    from lucky_droplets_amd.learned import collision_comparison
    cfg = Config(N=128, precision="f32", nu=0.0093, nshells=2, profile="k4")
    r = collision_comparison(cfg, npart=200_000, st=0.5, spinup=300, train_states=8, train_every=10,
                             train_iters=300, relax=200, samples=4, sample_every=20, rmax=0.1, nbins=8)
    print({k: v for k, v in r.items() if k != "forms"})
    for k in ("fine", "plain", "learned"):
        f = r["forms"][k]
        assert np.all(np.isfinite(f["K"])) and np.all(f["counts"] > 0)
    assert r["kernel_err_learned"] < 0.05
    assert r["kernel_err_learned"] < r["kernel_err_plain"]
    assert r["heldout_mse_learned"] < 0.95 * r["heldout_mse_plain"] and r["network_used"]
