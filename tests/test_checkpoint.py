"""Checkpoint files in the reference's layout: round trip through several
rank slabs, and files written the way the reference writes them."""
import numpy as np

from lucky_droplets_amd import checkpoint, ic
from lucky_droplets_amd.config import Config


def test_round_trip_any_rank_count(tmp_path):
    cfg = Config(N=16, nu=0.02, nshells=2, profile="k4", n_particles=40)
    rng = np.random.default_rng(0)
    uk, n, _ = ic.initial_flow(cfg, rng)
    n = n + rng.random(n.shape)
    coord, pid = ic.initial_droplets(cfg, rng)
    coord[:, :3] += 2 * np.pi          # unwrapped positions are stored wrapped
    for ranks in (1, 4):
        root = tmp_path / f"r{ranks}"
        base = checkpoint.save(root, 1.5, uk, n, ranks=ranks, coord=coord, prtclid=pid,
                               ek=np.arange(9.0), pik=np.arange(9.0) * 2)
        _, ndir, sdir = checkpoint.run_dirs(root, 1.5, "wo_g", 0.001, 0.05, True)
        uk2, n2 = checkpoint.load_fields(base, 16, ndir)
        assert np.array_equal(uk2, uk) and np.array_equal(n2, n)
        rows, ids = checkpoint.load_droplets(sdir)
        order = np.argsort(ids)
        assert np.array_equal(ids[order], pid)
        assert np.allclose(rows[order, :3], np.mod(coord[:, :3], 2 * np.pi), rtol=0, atol=1e-14)
        assert np.array_equal(rows[order, 3:], coord[:, 3:])
        assert np.array_equal(np.load(base / "Energy_spectrum.npz")["ek"], np.arange(9.0))


def test_reads_reference_written_slabs(tmp_path):
    """Files as forced-dns-sm.py:446 writes them with P = 2 ranks (each rank's
    uk slab is (N, N/P, Nf) and n slab (N/P, N, N))."""
    N, P = 8, 2
    rng = np.random.default_rng(1)
    uk = rng.standard_normal((3, N, N, 5)) + 1j * rng.standard_normal((3, N, N, 5))
    n = rng.random((N, N, N))
    d = tmp_path / "time_0.5"
    (d / "wo_g_sts_0.001_stb_0.050").mkdir(parents=True)
    for r in range(P):
        s = slice(r * N // P, (r + 1) * N // P)
        np.savez_compressed(d / f"Fields_k_{r}", uk=uk[0][:, s], vk=uk[1][:, s], wk=uk[2][:, s])
        np.savez_compressed(d / "wo_g_sts_0.001_stb_0.050" / f"n_{r}", n=n[s])
    uk2, n2 = checkpoint.load_fields(d, N, d / "wo_g_sts_0.001_stb_0.050")
    assert np.array_equal(uk2, uk) and np.array_equal(n2, n)


def test_load_fields_mask_removes_modes_outside_the_dealiasing_mask(tmp_path):
    """mask=True zeroes exactly the modes with round|k| > N/3 (the mask of
    forced-dns-sm.py:127-129, as the oracle builds it) and keeps the others."""
    import oracle
    N = 16
    rng = np.random.default_rng(2)
    uk = rng.standard_normal((3, N, N, 9)) + 1j * rng.standard_normal((3, N, N, 9))
    d = tmp_path / "f"
    d.mkdir()
    np.savez_compressed(d / "Fields_k_0", uk=uk[0], vk=uk[1], wk=uk[2])
    um, _ = checkpoint.load_fields(d, N, mask=True)
    keep = oracle.Grid(N).dealias
    assert np.array_equal(um, uk * keep)
    assert np.array_equal(checkpoint.dealias_mask(N), keep)
    assert 0 < keep.sum() < keep.size
