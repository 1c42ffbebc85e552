"""Pins of the brute-force collision/coalescence oracle (no reference code:
the specification is in oracle/collide.py)."""
import numpy as np

from oracle import collide

L = 2 * np.pi


def test_pairs_across_periodic_seam_and_threshold():
    pos = np.array([[0.01, 1.0, 1.0], [L - 0.01, 1.0, 1.0], [3.0, 3.0, 3.0], [3.0, 3.0, 3.0 + 0.1]])
    r = np.array([0.0105, 0.0105, 0.05, 0.05])
    got = collide.pairs_bruteforce(pos, r, L)
    assert got.tolist() == [[0, 1]]                       # 0.02 < 0.021 across the seam; 0.1 !< 0.1
    r2 = r.copy()
    r2[2] = 0.0500001
    assert collide.pairs_bruteforce(pos, r2, L).tolist() == [[0, 1], [2, 3]]


def test_pairs_against_independent_count():
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, L, (400, 3))
    r = rng.uniform(0.1, 0.3, 400)
    got = collide.pairs_bruteforce(pos, r, L)
    # independent: full distance matrix with np.abs and min(d, L-d)
    d = np.abs(pos[:, None, :] - pos[None, :, :])
    d = np.minimum(d, L - d)
    hit = np.sum(d ** 2, axis=2) < (r[:, None] + r[None, :]) ** 2
    iu = np.triu_indices(400, 1)
    expect = np.stack(iu, axis=1)[hit[iu]]
    assert got.shape[0] > 50
    assert np.array_equal(got, expect)
    assert np.all(np.diff(got[:, 0] * 400 + got[:, 1]) > 0)     # sorted


def test_coalescence_conserves_mass_and_momentum():
    rng = np.random.default_rng(1)
    rows = rng.uniform(0, 1, (50, 7))
    rows[:, :3] *= L
    rows[:, 6] += 0.5
    pairs = np.array([[0, 3], [0, 7], [2, 3], [5, 9], [9, 11], [12, 13]])
    out, ev = collide.coalesce(rows, pairs, L)
    assert ev.tolist() == [[0, 3], [5, 9], [12, 13]]      # later pairs touching 0, 3, 9 are skipped
    assert abs(out[:, 6].sum() - rows[:, 6].sum()) < 1e-12
    mom = (rows[:, 3:6] * rows[:, 6:7]).sum(0)
    assert np.allclose((out[:, 3:6] * out[:, 6:7]).sum(0), mom, rtol=1e-13, atol=1e-13)
    assert out[3, 6] == 0 and out[9, 6] == 0 and out[13, 6] == 0
    assert np.all((out[:, :3] >= 0) & (out[:, :3] < L))


def test_pair_stats_lattice_shells():
    """Simple cubic lattice of spacing a: every point has 6, 12, 8, 6
    neighbours at a, sqrt(2) a, sqrt(3) a, 2 a."""
    m = 6
    a = L / m
    g = (np.arange(m) + 0.25) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    rows = np.zeros((pos.shape[0], 7))
    rows[:, :3] = pos
    nbins = 40
    rmax = 2.05 * a
    c, ws, wn = collide.pair_stats_bruteforce(rows, L, rmax, nbins)
    n = pos.shape[0]
    dr = rmax / nbins
    for dist, per in [(a, 6), (np.sqrt(2) * a, 12), (np.sqrt(3) * a, 8), (2 * a, 6)]:
        assert c[int(dist / dr)] == n * per // 2
    assert c.sum() == n * 32 // 2
    assert np.all(ws == 0) and np.all(wn == 0)          # no relative motion


def test_pair_stats_uniform_expansion():
    """v = alpha x on a cluster far from the seams: w_r = alpha |d| for every
    pair; contraction gives wneg = wsum."""
    rng = np.random.default_rng(3)
    pos = 2.0 + rng.uniform(0, 1.0, (300, 3))
    for alpha in (0.7, -0.4):
        rows = np.zeros((300, 7))
        rows[:, :3] = pos
        rows[:, 3:6] = alpha * pos
        c, ws, wn = collide.pair_stats_bruteforce(rows, L, 2.0, 16)
        d = pos[:, None, :] - pos[None, :, :]
        dist = np.sqrt((d ** 2).sum(-1))[np.triu_indices(300, 1)]
        assert c.sum() == dist.size
        assert np.isclose(ws.sum(), abs(alpha) * dist.sum(), rtol=1e-12)
        assert np.isclose(wn.sum(), (-alpha if alpha < 0 else 0.0) * dist.sum(), rtol=1e-12, atol=1e-12)


def test_pair_stats_matches_pair_test():
    """Cutoff rmax counts exactly the pairs of radius rmax/2 each."""
    rng = np.random.default_rng(4)
    rows = rng.uniform(0, L, (500, 7))
    rmax = 0.7
    c, _, _ = collide.pair_stats_bruteforce(rows, L, rmax, 10)
    assert c.sum() == collide.pairs_bruteforce(rows[:, :3], np.full(500, rmax / 2), L).shape[0]


def test_kdtree_pairs_equal_bruteforce():
    """The periodic k-d tree search (used at 10^5 droplets) finds exactly the
    brute-force pairs, seam crossings included."""
    rng = np.random.default_rng(12)
    L = 2 * np.pi
    n = 2500
    pos = rng.uniform(0, L, (n, 3)) + L * rng.integers(-1, 2, (n, 3))
    pos[: n // 5, 1] = rng.uniform(-2e-3, 2e-3, n // 5)
    r = rng.uniform(0.01, 0.08, n)
    a = collide.pairs_bruteforce(pos, r, L)
    b = collide.pairs_kdtree(pos, r, L)
    assert a.shape[0] > 50
    assert np.array_equal(a, b)
