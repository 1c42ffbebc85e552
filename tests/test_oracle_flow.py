"""Pins of the CPU oracle's flow functions against closed forms, library
equivalences and invariants (the reference ships no tests or golden data).

Each check is chosen so that a plausible mistake (dropped term, wrong sign,
transposed operand, wrong index) fails it; see oracle/PINNING.md."""
import math

import numpy as np
import pytest
import scipy.fft

import oracle
from helpers import make_flow_oracle, relerr
from lucky_droplets_amd import ic
from lucky_droplets_amd.config import Config

L = 2 * np.pi


def grid_xyz(N):
    X = np.linspace(0, L, N, endpoint=False)
    return np.meshgrid(X, X, X, indexing="ij")


# ------------------------------------------------------------------ transforms
@pytest.mark.parametrize("N", [8, 16, 32])
def test_rfft_mpi_equals_rfftn(N):
    G = oracle.Grid(N)
    u = np.random.default_rng(N).standard_normal((N, N, N))
    assert relerr(G.rfft_mpi(u), scipy.fft.rfftn(u)) < 1e-14
    assert relerr(G.irfft_mpi(G.rfft_mpi(u)), u) < 1e-14


def test_single_mode_known_answer():
    N = 16
    G = oracle.Grid(N)
    x, y, z = grid_xyz(N)
    u = np.cos(2 * x + 3 * y - 1 * z)
    fu = G.rfft_mpi(u)
    # cos(k.x) has N^3/2 at k and at -k; in the half spectrum -k = (-2,-3,+1) appears
    # (kz = 1) and +k = (2,3,-1) does not (kz < 0): exactly one nonzero entry.
    nz = np.argwhere(np.abs(fu) > 1e-9)
    assert nz.shape[0] == 1
    i, j, k = nz[0]
    assert (G.kx[i, j, k], G.ky[i, j, k], G.kz[i, j, k]) == (-2.0, -3.0, 1.0)
    assert abs(fu[i, j, k] - N ** 3 / 2) < 1e-9


def test_parseval_energy():
    """0.5 sum |u|^2 dx^3 == sum of shell energies (forced-dns-sm.py:467-468)."""
    cfg = Config(N=32, nu=0.02, nshells=2, profile="k4")
    fl = make_flow_oracle(cfg)
    uk, n, u = ic.initial_flow(cfg, np.random.default_rng(0))
    G = fl.grid
    e_phys = 0.5 * np.sum(u ** 2) * G.dx ** 3
    e_spec = np.sum(G.e3d_to_e1d(G.energy3d(uk)))
    assert abs(e_phys - e_spec) / e_spec < 1e-12
    assert abs(e_spec - cfg.einit) / cfg.einit < 1e-12


# ------------------------------------------------------------------ operators
def test_operator_counts():
    """Mode counts of the dealias mask and of the low shells (half spectrum)."""
    N = 64
    G = oracle.Grid(N)
    # |k| in [0.5, 1.5): the 6 axis modes and 12 edge modes of the full spectrum;
    # half spectrum keeps kz >= 0: kz = 0 plane (all 4 + 4 in-plane) + kz = 1 (1 + 4).
    assert int(np.sum(G.kint == 1)) == 13
    # shell 2: |k|^2 in {3, 4, 5, 6}
    full = sum(1 for a in range(-3, 4) for b in range(-3, 4) for c in range(0, 4)
               if 1.5 <= math.sqrt(a * a + b * b + c * c) < 2.5)
    assert int(np.sum(G.kint == 2)) == full
    # spherical 2/3 rule: kint <= N/3 (forced-dns-sm.py:129)
    brute = sum(1 for a in range(-N // 2, N // 2) for b in range(-N // 2, N // 2) for c in range(0, N // 2 + 1)
                if round(math.sqrt(a * a + b * b + c * c)) <= N / 3)
    assert int(np.sum(G.dealias)) == brute


# ------------------------------------------------------------------ forcing
def test_forcing_power_and_solenoidality():
    """For solenoidal, dealiased u: total power sum normalize Re(u* f) = nshells f0;
    f is solenoidal and vanishes outside the forced shells."""
    cfg = Config(N=32, nu=0.02, nshells=2, profile="k4")
    fl = make_flow_oracle(cfg)
    uk, _, _ = ic.initial_flow(cfg, np.random.default_rng(1))
    G = fl.grid
    fk = fl.forcing(uk)
    power = np.sum(G.normalize * np.real(np.conj(uk) * fk))
    assert abs(power - cfg.nshells * cfg.f0) / (cfg.nshells * cfg.f0) < 1e-12
    div = G.kx * fk[0] + G.ky * fk[1] + G.kz * fk[2]
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(fk))
    outside = ~np.isin(G.kint, cfg.shell_no)
    assert np.max(np.abs(fk[:, outside])) == 0.0


def test_forcing_zero_energy_shell():
    """A shell with |E| < 1e-10 gets factor 0, not inf (forced-dns-sm.py:250)."""
    cfg = Config(N=16, nu=0.02, nshells=1)
    fl = make_flow_oracle(cfg)
    uk = np.zeros((3, 16, 16, 9), complex)
    assert np.all(fl.forcing(uk) == 0)


# ------------------------------------------------------------------ clip_zero
def test_clip_zero():
    fl = make_flow_oracle(Config(N=8, nu=0.01))
    x = np.random.default_rng(0).standard_normal((8, 8, 8)) + 0.5
    m0 = x.mean()
    x_in = x.copy()
    y = fl.clip_zero(x_in)
    assert np.all(x_in >= 0) and np.all(y >= 0)
    assert abs(y.mean() - m0) < 1e-14
    # identity on non-negative input
    z = np.abs(x)
    assert np.array_equal(fl.clip_zero(z.copy()), z * z.mean() / z.mean())


# ------------------------------------------------------------------ RHS closed forms
def abc_field(N, A=1.0, B=0.7, C=0.4):
    x, y, z = grid_xyz(N)
    return np.stack([A * np.sin(z) + C * np.cos(y), B * np.sin(x) + A * np.cos(z), C * np.sin(y) + B * np.cos(x)])


def test_beltrami_unforced_decay():
    """ABC flow (|k| = 1, curl u = u): u x w = 0, no pressure, so with forcing off
    k_u = -nu k^2 u and k_n = 0 for uniform n; one RK4 step multiplies u_hat by
    1 + z + z^2/2 + z^3/6 + z^4/24, z = -h nu."""
    N = 16
    nu, h = 0.05, 0.1
    fl = oracle.Flow(N, nu, [1], 1.0, tps=0.01, g=0.0, isforcing=False)
    G = fl.grid
    u = abc_field(N)
    uk = np.stack([G.rfft_mpi(u[c]) for c in range(3)])
    n = np.ones((N, N, N))
    ku, kn, _, _ = fl.full_RHS(0, uk, n, np.zeros((0, 7)), np.zeros((0, 7)), None)
    assert relerr(ku, -nu * uk) < 1e-13
    assert np.max(np.abs(kn)) < 1e-12
    _, uknew, nnew = fl.RK4(0, h, None, uk, n.copy())
    zz = -h * nu
    amp = 1 + zz + zz ** 2 / 2 + zz ** 3 / 6 + zz ** 4 / 24
    assert relerr(uknew, amp * uk) < 1e-13
    assert np.max(np.abs(nnew - 1)) < 1e-12


def test_beltrami_forced_recurrence():
    """With shell-1 forcing, k_u = (F - nu) u with F = f0 / (2 E1(stage)): the RK4
    step is a scalar recurrence in the stage amplitude (computed here by hand)."""
    N = 16
    nu, h, f0 = 0.05, 0.05, 0.8
    fl = oracle.Flow(N, nu, [1], f0, tps=0.0, g=0.0, isforcing=True)
    G = fl.grid
    u = abc_field(N)
    uk = np.stack([G.rfft_mpi(u[c]) for c in range(3)])
    E0 = np.sum(G.e3d_to_e1d(G.energy3d(uk))[1])

    def rate(a):  # d a / dt for u = a u0
        return (f0 / (2 * a * a * E0) - nu) * a

    a = 1.0
    k1 = rate(a)
    k2 = rate(a + h / 2 * k1)
    k3 = rate(a + h / 2 * k2)
    k4 = rate(a + h * k3)
    a1 = a + h / 6 * k1 + h / 3 * k2 + h / 3 * k3 + h / 6 * k4
    _, uknew, _ = fl.RK4(0, h, None, uk, np.ones((N, N, N)))
    assert relerr(uknew, a1 * uk) < 1e-12


def test_taylor_green_projection():
    """2-D Taylor-Green: u x w = -grad(sin^2 x sin^2 y) is a pure gradient, so the
    projected nonlinear term vanishes and (forcing off) k_u = -nu k^2 u."""
    N = 16
    nu = 0.03
    fl = oracle.Flow(N, nu, [1], 1.0, tps=0.0, g=0.0, isforcing=False)
    G = fl.grid
    x, y, z = grid_xyz(N)
    u = np.stack([np.sin(x) * np.cos(y), -np.cos(x) * np.sin(y), 0 * x])
    uk = np.stack([G.rfft_mpi(u[c]) for c in range(3)])
    ku, _, _, _ = fl.full_RHS(0, uk, np.ones((N, N, N)), np.zeros((0, 7)), np.zeros((0, 7)), None)
    assert relerr(ku, -nu * 2.0 * uk) < 1e-13
    # the unprojected term is not zero: the pressure must be what removes it
    w = fl.last["omg"]
    nl = np.stack([w[2] * u[1] - w[1] * u[2], w[0] * u[2] - w[2] * u[0], w[1] * u[0] - w[0] * u[1]])
    assert np.max(np.abs(nl)) > 0.1


def test_rhs_divergence_free_and_density_flux():
    """Random solenoidal field: k . k_u = 0 on the mask; with uniform n the
    density RHS is -div(v) with v_hat = (u - tps(-i k p - nu k^2 u)) mask."""
    cfg = Config(N=16, nu=0.05, nshells=2, profile="k4")
    fl = oracle.Flow(16, 0.05, [1, 2], cfg.f0, tps=0.3, g=0.0)
    G = fl.grid
    uk, n, _ = ic.initial_flow(cfg, np.random.default_rng(3))
    ku, kn, _, _ = fl.full_RHS(0, uk, n, np.zeros((0, 7)), np.zeros((0, 7)), None)
    div = G.kx * ku[0] + G.ky * ku[1] + G.kz * ku[2]
    assert np.max(np.abs(div)) < 1e-11 * np.max(np.abs(ku))
    vk = np.stack([G.rfft_mpi(fl.last["v"][c]) for c in range(3)])
    expect = G.irfft_mpi(-1j * (G.kx * vk[0] + G.ky * vk[1] + G.kz * vk[2]) * G.dealias)
    assert relerr(kn, expect) < 1e-11
    # tps = 0 makes v = u (dealiased): the flux of a uniform density vanishes
    fl0 = oracle.Flow(16, 0.05, [1, 2], cfg.f0, tps=0.0, g=0.0)
    _, kn0, _, _ = fl0.full_RHS(0, uk, n, np.zeros((0, 7)), np.zeros((0, 7)), None)
    assert np.max(np.abs(kn0)) < 1e-10


def test_slow_manifold_velocity_beltrami_closed_form():
    """ABC flow (|k| = 1, p = 0): v_hat = (u_hat - tps (0 - nu k^2 u_hat)) mask
    = (1 + tps nu) u_hat, and v_z is shifted by -tps g (forced-dns-sm.py:319-321).
    A flipped sign of the tps term gives (1 - tps nu); a dropped shift leaves
    mean(v_z) = 0."""
    N = 16
    nu, tps, g = 0.05, 0.4, 2.5
    fl = oracle.Flow(N, nu, [1], 1.0, tps=tps, g=g, isforcing=False)
    G = fl.grid
    u = abc_field(N)
    uk = np.stack([G.rfft_mpi(u[c]) for c in range(3)])
    fl.full_RHS(0, uk, np.ones((N, N, N)), np.zeros((0, 7)), np.zeros((0, 7)), None)
    expect = (1 + tps * nu) * u
    expect[2] -= tps * g
    assert relerr(fl.last["v"], expect) < 1e-13
    # forcing on: f_hat = F u_hat is solenoidal, still no pressure; v keeps the same form
    fl1 = oracle.Flow(N, nu, [1], 0.7, tps=tps, g=g, isforcing=True)
    fl1.full_RHS(0, uk, np.ones((N, N, N)), np.zeros((0, 7)), np.zeros((0, 7)), None)
    assert relerr(fl1.last["v"], expect) < 1e-13


def test_slow_manifold_velocity_taylor_green_pressure():
    """2-D Taylor-Green u = (sin x cos y, -cos x sin y, 0): u x w = -grad(phi),
    phi = sin^2 x sin^2 y, so the pressure is p_hat = -phi_hat and
    v = (1 + 2 tps nu) u - tps grad(phi) - tps g e_z (forced-dns-sm.py:309, 319-321).
    Pins the sign of the pressure term inside v (|k|^2 = 2 for u)."""
    N = 16
    nu, tps, g = 0.03, 0.25, 1.5
    fl = oracle.Flow(N, nu, [1], 1.0, tps=tps, g=g, isforcing=False)
    G = fl.grid
    x, y, z = grid_xyz(N)
    u = np.stack([np.sin(x) * np.cos(y), -np.cos(x) * np.sin(y), 0 * x])
    uk = np.stack([G.rfft_mpi(u[c]) for c in range(3)])
    fl.full_RHS(0, uk, np.ones((N, N, N)), np.zeros((0, 7)), np.zeros((0, 7)), None)
    grad_phi = np.stack([np.sin(2 * x) * np.sin(y) ** 2, np.sin(x) ** 2 * np.sin(2 * y), 0 * x])
    expect = (1 + 2 * tps * nu) * u - tps * grad_phi
    expect[2] -= tps * g
    assert relerr(fl.last["v"], expect) < 1e-13
    # the pressure term is not negligible here: dropping it misses by O(tps)
    assert np.max(np.abs(tps * grad_phi)) > 0.1


def _droplet_flow(N, nparts, seed, tps=0.02, g=0.8, u_scale=1.0):
    cfg = Config(N=N, nu=0.05, nshells=2, profile="k4", n_particles=nparts, st_dist=("uniform", 0.05, 1.0),
                 gravity=True)
    fl = oracle.Flow(N, 0.05, [1, 2], cfg.f0, tps=tps, g=g)
    rng = np.random.default_rng(seed)
    uk, n, _ = ic.initial_flow(cfg, rng)
    coord, pid = ic.initial_droplets(cfg, rng)
    n = n + 0.6 * (rng.random(n.shape) - 0.5)
    dr = oracle.Droplets(cfg.L, N, coord.copy(), cfg.sts, g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                         pid.reshape(-1, 1).copy())
    return fl, dr, uk * u_scale, n, coord


def test_density_rhs_mean_equals_minus_deposit_mean():
    """The flux term -i k.(n v)_hat has no k = 0 mode, so the mean of kn is
    -mean(fc) (forced-dns-sm.py:325-330): the droplets' growth is a sink of n."""
    fl, dr, uk, n, coord = _droplet_flow(16, 300, 4)
    ku, kn, kp, _ = fl.full_RHS(0, uk, n, coord.copy(), coord.copy(), dr)
    fc = fl.last["fc"]
    assert np.sum(fc) > 0
    assert abs(np.mean(kn) + np.mean(fc)) < 1e-12 * np.mean(fc)
    # and the sink equals the droplets' mass gain: sum fc dx^3 = sum W = sum kp[:, 6]
    G = fl.grid
    assert abs(np.sum(fc) * G.dx ** 3 - np.sum(kp[:, 6])) < 1e-12 * np.sum(kp[:, 6])


def test_density_rhs_still_flow_single_droplet():
    """u = 0, uniform n = n0, one droplet: v = -tps g e_z is constant, so n v has
    no flux and kn = -irfft(rfft(fc) mask) exactly; W = growth (1 + sqrt(sts/St))^2
    n0 |v_p - v| (particles.py:828), sum fc dx^3 = W (forced-dns-sm.py:325-330)."""
    N = 16
    tps, g, n0 = 0.02, 0.8, 1.7
    fl, dr, uk, n, coord = _droplet_flow(N, 1, 6, tps=tps, g=g, u_scale=0.0)
    n = np.full((N, N, N), n0)
    ku, kn, kp, _ = fl.full_RHS(0, uk, n, coord.copy(), coord.copy(), dr)
    G = fl.grid
    fc = fl.last["fc"]
    st = (coord[0, 6] / dr.factor) ** (2 / 3)
    vrel = coord[0, 3:6] - np.array([0.0, 0.0, -tps * g])
    W = dr.growthfactor * (1 + np.sqrt(dr.st_s / st)) ** 2 * n0 * np.linalg.norm(vrel)
    assert abs(kp[0, 6] - W) < 1e-13 * W
    assert abs(np.sum(fc) * G.dx ** 3 - W) < 1e-12 * W
    expect = -G.irfft_mpi(G.rfft_mpi(fc) * G.dealias)
    assert relerr(kn, expect) < 1e-12
    assert np.max(np.abs(kn)) > 0


def test_reality_round_trip_symmetrises_planes():
    """rfft(irfft(X)) makes the kz = 0 and kz = N/2 planes Hermitian and
    leaves other modes unchanged (forced-dns-sm.py:597-603)."""
    N = 16
    fl = oracle.Flow(N, 0.01, [1], 1.0, tps=0.0)
    G = fl.grid
    rng = np.random.default_rng(5)
    X = rng.standard_normal((3, N, N, N // 2 + 1)) + 1j * rng.standard_normal((3, N, N, N // 2 + 1))
    Y = np.stack([G.rfft_mpi(G.irfft_mpi(X[c])) for c in range(3)])
    neg = (-np.arange(N)) % N
    for kz in (0, N // 2):
        P = X[:, :, :, kz]
        H = 0.5 * (P + np.conj(P[:, neg][:, :, neg]))
        assert relerr(Y[:, :, :, kz], H) < 1e-13
    assert relerr(Y[:, :, :, 1:N // 2], X[:, :, :, 1:N // 2]) < 1e-13


def test_step_projection_and_blowup():
    cfg = Config(N=16, nu=0.05, nshells=2, profile="k4")
    fl = oracle.Flow(16, 0.05, [1, 2], cfg.f0, tps=0.01, g=0.0)
    G = fl.grid
    uk, n, _ = ic.initial_flow(cfg, np.random.default_rng(2))
    uk1, n1 = fl.step(0, 0.01, None, uk, n)
    div = (G.kx * uk1[0] + G.ky * uk1[1] + G.kz * uk1[2]) * G.dealias
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(uk1))
    assert np.all(n1 >= 0)
    big = uk.copy()
    big[1, 0, 0, 1] = 101 * 16 ** 3      # transverse mode k = (0, 0, 1), kz > 0: kept by projection and symmetrisation
    with pytest.raises(oracle.flow.BlowUp):
        fl.step(0, 1e-6, None, big, n)


# ------------------------------------------------------------------ initial condition
def test_initial_flow_properties():
    """The host-side initial condition (lucky_droplets_amd.ic, the recipe of
    forced-dns-sm.py:691-737 and :567-569), checked by properties the recipe
    fixes: solenoidal, energy sum_k E(k) = (2 pi)^3, no mode at k = 0, outside
    the dealiasing mask or at kint >= kinit, Hermitian kz = 0 and N/2 planes,
    n = 1, and the same field for the same seed."""
    for prof in ("k4", "reference"):
        cfg = Config(N=16, nu=0.02, nshells=2, profile=prof, kinit=6)
        fl = make_flow_oracle(cfg)
        G = fl.grid
        uk, n, u = ic.initial_flow(cfg, np.random.default_rng(11))
        div = G.kx * uk[0] + G.ky * uk[1] + G.kz * uk[2]
        assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(uk))
        E = G.e3d_to_e1d(G.energy3d(uk))
        assert abs(E.sum() - (2 * np.pi) ** 3) < 1e-10 * (2 * np.pi) ** 3
        outside = (~G.dealias) | (G.kint == 0) | (G.kint >= cfg.kinit)
        assert np.max(np.abs(uk[:, outside])) < 1e-12 * np.max(np.abs(uk))
        assert np.max(np.abs(uk[:, ~outside])) > 0
        for iz in (0, 8):
            P = uk[:, :, :, iz]
            Q = np.conj(np.roll(np.flip(P, axis=(1, 2)), 1, axis=(1, 2)))
            assert np.max(np.abs(P - Q)) < 1e-12 * np.max(np.abs(uk))
        assert np.array_equal(n, np.ones((16, 16, 16)))
        uk2, _, _ = ic.initial_flow(cfg, np.random.default_rng(11))
        assert np.array_equal(uk, uk2)
        # u is the physical field the loop transforms (uk = rfft(u) * dealias)
        assert relerr(np.stack([G.rfft_mpi(u[c]) for c in range(3)]) * G.dealias, uk) < 1e-14


def test_divergence_closed_forms():
    """diff_x/diff_y/diff_z (forced-dns-sm.py:216-229): the divergence of
    (sin(2x) cos y, sin z, cos(3y)) is 2 cos(2x) cos y; a solenoidal ABC field
    has none; the Nyquist mode cos(N x / 2) has zero derivative (the 1-D irfft
    drops the imaginary Nyquist coefficient)."""
    N = 16
    G = oracle.Grid(N)
    x, y, z = grid_xyz(N)
    u = np.stack([np.sin(2 * x) * np.cos(y), np.sin(z), np.cos(3 * y)])
    assert np.max(np.abs(G.divergence(u) - 2 * np.cos(2 * x) * np.cos(y))) < 1e-13
    # a wrong axis or sign would give -2 cos(2x) cos y or mix in cos(z)
    assert np.max(np.abs(G.divergence(abc_field(N)))) < 1e-13
    ny = np.stack([np.cos(N / 2 * x), np.cos(N / 2 * y), np.cos(N / 2 * z)])
    assert np.max(np.abs(G.divergence(ny))) < 1e-12
    # diff along each axis separately: d/dy of sin(3y) is 3 cos(3y)
    assert np.max(np.abs(G.diff(np.sin(3 * y), 1) - 3 * np.cos(3 * y))) < 1e-12


def test_dissipation_spectrum_closed_form():
    """u = (sin(k0 z), 0, 0): energy sum_k E(k) = L^3/4 and dissipation
    nu int |grad u|^2 = nu k0^2 L^3 / 2, all in shell k0."""
    N, k0, nu = 16, 3, 0.07
    G = oracle.Grid(N)
    x, y, z = grid_xyz(N)
    u = np.stack([np.sin(k0 * z), 0 * z, 0 * z])
    uk = np.stack([G.rfft_mpi(u[c]) for c in range(3)])
    L3 = (2 * np.pi) ** 3
    E = G.e3d_to_e1d(G.energy3d(uk))
    assert abs(E.sum() - L3 / 4) < 1e-12 * L3
    D = G.dissipation_spectrum(uk, nu)
    assert abs(D[k0] - nu * k0 ** 2 * L3 / 2) < 1e-12 * L3
    assert abs(D.sum() - D[k0]) < 1e-14 * L3
