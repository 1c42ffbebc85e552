"""Brute-force CPU oracle for droplet-droplet collision pairs and coalescence.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference has no droplet-droplet collisions (its droplets grow only by
collecting the continuum density, particles.py:828).  The specification these
functions define, and the GPU cell-list path must reproduce bit for bit:

pairs
    all (i, j), i < j (row indices), with
        d = x_j - x_i  per axis,  d -= L * rint(d / L)      (minimum image)
        d2 = (dx*dx + dy*dy) + dz*dz
        collide  iff  d2 < (r_i + r_j) * (r_i + r_j)
    in IEEE binary64 with no fused multiply-add, on wrapped positions
    x mod L; sorted lexicographically by (i, j).
coalescence
    pairs are visited in that order; a pair merges only if neither droplet
    took part in an earlier merge of this call.  The survivor is the smaller
    index i; with m = m_i + m_j:
        x_i <- wrap(x_i + (m_j / m) * d)      (d as above, minimum image)
        v_i <- (m_i v_i + m_j v_j) / m
        m_i <- m
    and droplet j is left with m_j = 0 and v_j = 0 (inactive).
"""
from __future__ import annotations

import numpy as np


def wrap(x, L):
    r = np.fmod(x, L)
    return np.where(r < 0, r + L, r)


def pairs_bruteforce(pos: np.ndarray, radius: np.ndarray, L: float) -> np.ndarray:
    """O(Np^2) enumeration; returns an (npairs, 2) int64 array sorted by (i, j)."""
    x = wrap(np.asarray(pos, dtype=np.float64)[:, :3], L)
    r = np.asarray(radius, dtype=np.float64)
    n = x.shape[0]
    out = []
    for i in range(n - 1):
        d = x[i + 1:] - x[i]
        d = d - L * np.rint(d / L)
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        s = r[i] + r[i + 1:]
        hit = np.nonzero(d2 < s * s)[0]
        for k in hit:
            out.append((i, i + 1 + k))
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def pair_stats_bruteforce(rows: np.ndarray, L: float, rmax: float, nbins: int):
    """Pair statistics within rmax, O(Np^2): for every pair i < j with
    minimum-image d2 < rmax^2 (d2 as in pairs_bruteforce),
        dist = sqrt(d2), k = min(int(dist / (rmax / nbins)), nbins - 1),
        w_r = ((dvx*dx + dvy*dy) + dvz*dz) / dist   (dv = v_j - v_i; 0 if dist == 0),
    returns counts[k] (int64), wsum[k] = sum |w_r|, wneg[k] = sum max(-w_r, 0)."""
    rows = np.asarray(rows, dtype=np.float64)
    x = wrap(rows[:, :3], L)
    v = rows[:, 3:6]
    n = x.shape[0]
    dr = rmax / nbins
    counts = np.zeros(nbins, dtype=np.int64)
    wsum = np.zeros(nbins)
    wneg = np.zeros(nbins)
    for i in range(n - 1):
        d = x[i + 1:] - x[i]
        d = d - L * np.rint(d / L)
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        hit = np.nonzero(d2 < rmax * rmax)[0]
        if hit.size == 0:
            continue
        dist = np.sqrt(d2[hit])
        k = np.minimum((dist / dr).astype(np.int64), nbins - 1)
        dv = v[i + 1:][hit] - v[i]
        dh = d[hit]
        num = (dv[:, 0] * dh[:, 0] + dv[:, 1] * dh[:, 1]) + dv[:, 2] * dh[:, 2]
        wr = np.where(dist > 0, num / np.where(dist > 0, dist, 1.0), 0.0)
        np.add.at(counts, k, 1)
        np.add.at(wsum, k, np.abs(wr))
        np.add.at(wneg, k, np.maximum(-wr, 0.0))
    return counts, wsum, wneg


def coalesce(rows: np.ndarray, pairs: np.ndarray, L: float):
    """Apply the greedy merge rule to (Np, 7) rows; returns (rows, events)."""
    rows = np.array(rows, dtype=np.float64, copy=True)
    used = np.zeros(rows.shape[0], dtype=bool)
    events = []
    for i, j in pairs:
        if used[i] or used[j]:
            continue
        xi = wrap(rows[i, :3], L)
        xj = wrap(rows[j, :3], L)
        d = xj - xi
        d = d - L * np.rint(d / L)
        mi, mj = rows[i, 6], rows[j, 6]
        m = mi + mj
        rows[i, :3] = wrap(xi + (mj / m) * d, L)
        rows[i, 3:6] = (mi * rows[i, 3:6] + mj * rows[j, 3:6]) / m
        rows[i, 6] = m
        rows[j, 3:7] = 0.0
        used[i] = used[j] = True
        events.append((i, j))
    return rows, np.array(events, dtype=np.int64).reshape(-1, 2)


def pairs_kdtree(pos: np.ndarray, radius: np.ndarray, L: float) -> np.ndarray:
    """The pairs of pairs_bruteforce found with an independent periodic
    neighbour search (scipy.spatial.cKDTree with boxsize L) for the candidates
    within 2 max(r) (slightly widened), then the same FP64 test on each
    candidate; for populations too large for the O(Np^2) loop."""
    from scipy.spatial import cKDTree
    x = wrap(np.asarray(pos, dtype=np.float64)[:, :3], L)
    r = np.asarray(radius, dtype=np.float64)
    xt = np.where(x >= L, x - L, x)                # the tree needs [0, L)
    cand = cKDTree(xt, boxsize=L).query_pairs(2.0 * r.max() * (1 + 1e-9), output_type="ndarray")
    if cand.size == 0:
        return np.zeros((0, 2), dtype=np.int64)
    i = np.minimum(cand[:, 0], cand[:, 1]).astype(np.int64)
    j = np.maximum(cand[:, 0], cand[:, 1]).astype(np.int64)
    d = x[j] - x[i]
    d = d - L * np.rint(d / L)
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    s = r[i] + r[j]
    keep = d2 < s * s
    out = np.stack([i[keep], j[keep]], axis=1)
    return out[np.lexsort((out[:, 1], out[:, 0]))]
