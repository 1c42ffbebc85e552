"""The dealiasing mask as the truncated passes use it (CPU, against the oracle's mask).

The GPU passes skip work by two integer tests instead of evaluating the mask
per mode: a mode is kept iff |k|^2 <= kc (kc + 1) with kc = N // 3, and a
(ky, kz) line along kx holds a kept mode iff ky^2 + kz^2 <= kc (kc + 1) (so no
kept mode has kz > kc).  Both are checked here against the mask of
forced-dns-sm.py:127-129 as the oracle forms it (round(|k|) clipped to N/2,
then <= N/3).
"""
import numpy as np
import pytest

from oracle.flow import Grid


@pytest.mark.parametrize("N", [8, 16, 32, 48, 64, 96, 128, 256])
def test_sphere_radius_matches_mask(N):
    G = Grid(N)
    kc = N // 3
    k2 = np.rint(-G.lap).astype(np.int64)
    assert np.array_equal(G.dealias, k2 <= kc * (kc + 1))


@pytest.mark.parametrize("N", [16, 64, 256])
def test_line_and_plane_limits(N):
    G = Grid(N)
    kc = N // 3
    ky = np.rint(G.ky[0]).astype(np.int64)          # (ky, kz) plane at kx = 0
    kz = np.rint(G.kz[0]).astype(np.int64)
    line_kept = G.dealias.any(axis=0)                 # any kept mode along kx
    assert np.array_equal(line_kept, ky ** 2 + kz ** 2 <= kc * (kc + 1))
    assert not G.dealias[:, :, kc + 1:].any()
    assert G.dealias[:, :, kc].any()
