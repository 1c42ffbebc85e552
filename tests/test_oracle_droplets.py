"""Pins of the CPU oracle's droplet functions (particles.py restated): kernel
identities, a brute-force all-nodes interpolation, deposit conservation and
adjointness, the droplet RHS in uniform flow, send() semantics and the RK4
combination on an affine ODE."""
import math

import numpy as np
import pytest

import oracle
from helpers import by_id, make_drops_oracle, relerr
from lucky_droplets_amd.config import Config

L = 2 * np.pi


def drops(N, coord, **kw):
    cfg = Config(N=N, nu=kw.get("nu", 0.01))
    return oracle.Droplets(L, N, coord, kw.get("st_s", 0.001), kw.get("g", 0.0), cfg.nu_eff, kw.get("tau", 0.9),
                           1000.0, 4e-6)


def brute_interp(f, pos, N):
    """Sum over ALL grid nodes of w(dx) w(dy) w(dz) f with the periodic minimum
    image and w(r) = (1 + cos(pi r / 2 dx)) / 4 on |r| < 2 dx (0 outside)."""
    dx = L / N
    X = np.arange(N) * dx
    out = np.zeros(pos.shape[0])
    for p in range(pos.shape[0]):
        w = []
        for a in range(3):
            r = pos[p, a] - X
            r = r - L * np.round(r / L)
            w.append(np.where(np.abs(r) < 2 * dx, 0.25 * (1 + np.cos(np.pi * r / (2 * dx))), 0.0))
        out[p] = np.einsum("i,j,k,ijk->", w[0], w[1], w[2], f)
    return out


def test_partition_of_unity_and_constant_field():
    N = 16
    rng = np.random.default_rng(0)
    c = rng.uniform(0, L, (50, 7))
    d = drops(N, c)
    f = np.full((2, N, N, N), 3.25)
    I = d.interp_cosine(c, f)
    assert np.max(np.abs(I - 3.25)) < 1e-13


def test_node_weights_quarter_half_quarter():
    N = 16
    dx = L / N
    c = np.zeros((1, 7))
    c[0, :3] = [5 * dx, 7 * dx, 2 * dx]
    d = drops(N, c)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((1, N, N, N))
    w = {-1: 0.25, 0: 0.5, 1: 0.25}
    expect = sum(w[a] * w[b] * w[cc] * f[0, 5 + a, 7 + b, 2 + cc] for a in w for b in w for cc in w)
    assert abs(d.interp_cosine(c, f)[0, 0] - expect) < 1e-12


@pytest.mark.parametrize("N", [8, 12, 16])
def test_interp_matches_brute_force(N):
    rng = np.random.default_rng(N)
    c = rng.uniform(0, L, (40, 7))
    c[:5, 0] = [0.0, L - 1e-12, 1e-13, L / N * (N - 1), L / 2]   # near the periodic seam
    d = drops(N, c)
    f = rng.standard_normal((1, N, N, N))
    assert relerr(d.interp_cosine(c, f)[:, 0], brute_interp(f[0], c[:, :3], N)) < 1e-12


def test_deposit_mass_and_adjointness():
    N = 16
    dx = L / N
    rng = np.random.default_rng(2)
    c = rng.uniform(0, L, (60, 7))
    d = drops(N, c)
    W = rng.uniform(0.1, 2.0, 60)
    d.exterpmat = W[:, None].copy()
    fc = d.exterp_cosine_scalar(c, np.zeros((N, N, N)))
    assert abs(np.sum(fc) * dx ** 3 - np.sum(W)) < 1e-12 * np.sum(W)
    g = rng.standard_normal((1, N, N, N))
    lhs = np.sum(fc * g[0]) * dx ** 3
    rhs = np.sum(W * d.interp_cosine(c, g)[:, 0])
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)


def test_prhs_uniform_flow_closed_form():
    """Uniform u = U, v = V, n = n0: interp is exact, so
    rhs = [v_p, (U - v_p)/(St tau), W] with W = growth (1 + sqrt(sts/St))^2 n0 |v_p - V|."""
    N = 8
    rng = np.random.default_rng(3)
    c = rng.uniform(0, L, (20, 7))
    d = drops(N, c, g=0.3, tau=0.7)
    c[:, 6] = rng.uniform(0.05, 1.0, 20) ** 1.5 * d.factor
    d.coord = c.copy()
    U, V, n0 = np.array([0.3, -0.2, 0.5]), np.array([0.1, 0.4, -0.6]), 1.7
    u = np.broadcast_to(U[:, None, None, None], (3, N, N, N)).copy()
    v = np.broadcast_to(V[:, None, None, None], (3, N, N, N)).copy()
    n = np.full((N, N, N), n0)
    sump, rhs, fc = d.pRHS(0, c.copy(), u, v, n, np.zeros((N, N, N)), c.copy())
    cs = by_id(d.coord, d.prtclid)   # pRHS may reorder rows: compare by id
    rs = by_id(rhs, d.prtclid)
    st = (cs[:, 6] / d.factor) ** (2 / 3)
    vp = cs[:, 3:6]
    assert relerr(rs[:, :3], vp) < 1e-14
    acc = (U - vp) / (st[:, None] * 0.7)
    acc[:, 2] -= 0.3
    assert relerr(rs[:, 3:6], acc) < 1e-13
    W = d.growthfactor * (1 + np.sqrt(0.001 / st)) ** 2 * n0 * np.linalg.norm(vp - V, axis=1)
    assert relerr(rs[:, 6], W) < 1e-13
    assert abs(np.sum(fc) * (L / N) ** 3 - np.sum(W)) < 1e-12 * np.sum(W)


def test_send_single_rank_semantics():
    N = 8
    c = np.array([[1.0, 2, 3, 0, 0, 0, 1], [-0.5, 1, 7.0, 0, 0, 0, 1], [L + 0.25, -1, 1, 0, 0, 0, 1],
                  [2.0, 2, 2, 0, 0, 0, 1], [7.0, 3, 3, 0, 0, 0, 1]])
    d = drops(N, c)
    ids = np.arange(5).reshape(-1, 1).astype(float)
    base = c.copy()
    x, (b, i) = d.send(c.copy(), [base.copy(), ids.copy()])
    # in-range rows keep order; out-of-range rows move to the end sorted by x0 mod L
    key = np.mod(c[:, 0], L)
    moved = [k for k in range(5) if not (0 <= c[k, 0] < L)]
    expect = [k for k in range(5) if 0 <= c[k, 0] < L] + sorted(moved, key=lambda k: key[k])
    assert list(i[:, 0].astype(int)) == expect
    assert np.all((x[:, :3] >= 0) & (x[:, :3] < L))
    assert np.allclose(x[:, :3], np.mod(c[expect, :3], L))
    assert np.array_equal(b, base[expect])           # args permuted, not wrapped


def test_rk4_affine_droplet_relaxation():
    """No flow (u = 0), no growth: x' = v, v' = (0 - v)/(St tau) - g e_z; one RK4 step
    of the reference's stage scheme equals RK4 applied by hand to this linear ODE."""
    N = 8
    nu, tau, g, h = 0.01, 0.8, 0.4, 0.05
    fl = oracle.Flow(N, nu, [1], 1.0, tps=0.0, g=0.0, isforcing=False)
    rng = np.random.default_rng(4)
    c = rng.uniform(0, L, (10, 7))
    d = drops(N, c, tau=tau, g=g)
    c[:, 6] = rng.uniform(0.1, 1.0, 10) ** 1.5 * d.factor
    d.coord = c.copy()
    d.growthfactor = 0.0
    uk = np.zeros((3, N, N, N // 2 + 1), complex)
    sump, _, _ = fl.RK4(0, h, d, uk, np.ones((N, N, N)))
    got = by_id(sump, d.prtclid)
    st = (c[:, 6] / d.factor) ** (2 / 3)
    y = c[:, :6].copy()

    def f(y):
        out = np.empty_like(y)
        out[:, :3] = y[:, 3:]
        out[:, 3:] = -y[:, 3:] / (st[:, None] * tau)
        out[:, 5] -= g
        return out

    k1 = f(y); k2 = f(y + h / 2 * k1); k3 = f(y + h / 2 * k2); k4 = f(y + h * k3)
    y1 = y + h / 6 * k1 + h / 3 * k2 + h / 3 * k3 + h / 6 * k4
    assert relerr(got[:, 3:6], y1[:, 3:]) < 1e-13
    assert np.max(np.abs(got[:, :3] - y1[:, :3])) < 1e-13
    assert np.array_equal(got[:, 6], c[:, 6])


def test_factor_operator_precedence():
    """factor uses (nu tau/(2 rho))**3 / 2 (particles.py:74), not **1.5."""
    d = drops(8, np.zeros((1, 7)), nu=0.01, tau=0.9)
    cfg = Config(N=8, nu=0.01)
    x = cfg.nu_eff * 0.9 / 2000.0
    assert d.factor == pytest.approx(x ** 3 / 2 * 36 * math.pi * 1000 / 4e-6, rel=1e-15)
