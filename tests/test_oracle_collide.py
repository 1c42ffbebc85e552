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
