"""Checkpoint files in the reference's format (host side, numpy .npz).

Layout written by the reference's ``save`` (forced-dns-sm.py:439-460) under
``<root>/time_{t:.1f}/``:

  Fields_k_{r}.npz     uk, vk, wk   spectral slab of rank r: (N, N/P, N/2+1) [kx][ky-slab][kz]
  Energy_spectrum.npz  ek           E(k), N/2+1 shells
  Flux_spectrum.npz    Pik          Pi(k), N/2+1 shells
  {wg}_sts_{tps/tf:.3f}_stb_{stb0:.3f}/n_{r}.npz     n      physical slab (N/P, N, N)
  {wg}_stb_{stb0:.3f}_sts_{tps/tf:.3f}/state_{r}.npz pos, vel, mass, prtclid, umat

(without droplets the reference uses ``last/{wg}_sts_{tps/tf:.3f}/n_{r}.npz``).
``load_fields`` re-slabs any number of rank files to one slab as ``load_npz``
does (forced-dns-sm.py:374-407).  Droplet rows of rank r are those whose
wrapped x lies in the rank's x-slab (particles.py:64-70, 235).
"""
from __future__ import annotations

import os
from pathlib import Path
from typing import Optional

import numpy as np


def _np(x):
    if x is None:
        return None
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(x)


def run_dirs(root, t: float, wg: str, sts_ratio: float, stb0: float, with_droplets: bool):
    """(fields dir, n dir, state dir) as the reference names them (forced-dns-sm.py:439-453)."""
    root = Path(root)
    if with_droplets:
        base = root / f"time_{t:.1f}"
        ndir = base / f"{wg}_sts_{sts_ratio:.3f}_stb_{stb0:.3f}"
        sdir = base / f"{wg}_stb_{stb0:.3f}_sts_{sts_ratio:.3f}"
    else:
        base = root / "last"
        ndir = base / f"{wg}_sts_{sts_ratio:.3f}"
        sdir = None
    return base, ndir, sdir


def fluid_velocity_at_droplets(dns, stb, uk):
    """umat of the state files: the fluid velocity u = irfft(uk) interpolated at
    the droplets with the cosine kernel (forced-dns-sm.py:455-460), on the GPU;
    (Np, 3) float64 tensor in the row order of stb.coord."""
    import torch
    u = dns.irfft_mpi(uk, dns.zeros_real(3))
    return stb.interp_cosine(stb.coord, u)[:, :3].clone()


def save(root, t, uk, n, *, ranks: int = 1, coord=None, prtclid=None, umat=None, ek=None, pik=None,
         wg: str = "wo_g", sts_ratio: float = 0.001, stb0: float = 0.05, L: float = 2 * np.pi, dns=None, stb=None):
    """Write one snapshot as `ranks` slabs.  Returns the fields directory.

    With ``dns`` and ``stb`` (and no ``umat``), the state files' umat is the
    fluid velocity at the droplets, as the reference writes it; otherwise umat
    is what the caller passes, or zeros."""
    if umat is None and dns is not None and stb is not None and coord is not None:
        umat = fluid_velocity_at_droplets(dns, stb, uk)
    uk = _np(uk)
    n = _np(n)
    N = n.shape[0]
    if N % ranks:
        raise ValueError("ranks must divide N")
    Np = N // ranks
    with_droplets = coord is not None and _np(coord).shape[0] > 0
    base, ndir, sdir = run_dirs(root, t, wg, sts_ratio, stb0, with_droplets)
    base.mkdir(parents=True, exist_ok=True)
    ndir.mkdir(parents=True, exist_ok=True)
    for r in range(ranks):
        sl = slice(r * Np, (r + 1) * Np)
        np.savez_compressed(base / f"Fields_k_{r}", uk=uk[0][:, sl], vk=uk[1][:, sl], wk=uk[2][:, sl])   # :446
        np.savez_compressed(ndir / f"n_{r}", n=n[sl])                                                      # :449-450
    if ek is not None:
        np.savez_compressed(base / "Energy_spectrum", ek=_np(ek))                                          # :447
    if pik is not None:
        np.savez_compressed(base / "Flux_spectrum", Pik=_np(pik))                                          # :448
    if with_droplets:
        c = _np(coord)
        ids = np.arange(c.shape[0]) if prtclid is None else _np(prtclid).reshape(-1)
        um = np.zeros((c.shape[0], 3)) if umat is None else _np(umat)[:, :3]
        sdir.mkdir(parents=True, exist_ok=True)
        xw = np.mod(c[:, 0], L)
        owner = np.minimum((xw // (L / ranks)).astype(int), ranks - 1)
        for r in range(ranks):
            m = owner == r
            np.savez_compressed(sdir / f"state_{r}.npz", pos=np.mod(c[m, :3], L), vel=c[m, 3:6], mass=c[m, 6],
                                prtclid=ids[m].reshape(-1, 1).astype(np.float64), umat=um[m])              # :460
    return base


def dealias_mask(N: int) -> np.ndarray:
    """The 2/3-rule mask kint <= N/3 on the half spectrum (forced-dns-sm.py:127-129)."""
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    kint = np.clip(np.round(kk).astype(int), None, N // 2)
    return kint <= N / 3


def load_fields(fields_dir, N: int, n_dir: Optional[os.PathLike] = None, mask: bool = False):
    """(uk (3, N, N, N/2+1) complex128, n (N, N, N) or ones) re-slabbed from any
    number of rank files (forced-dns-sm.py:374-407).

    mask=True multiplies uk by the dealiasing mask.  The reference's reality
    round trip leaves round-off (about 1e-16 of the largest mode) in the modes
    outside the mask, so a state it wrote does not qualify for the stepping's
    truncated fast path; the mask removes that round-off (the state then
    differs from the file at that level) and the steps take the fast path."""
    fields_dir = Path(fields_dir)
    nslabs = len([x for x in fields_dir.iterdir() if "Fields" in x.name and x.name.endswith(".npz")])   # :375
    if nslabs == 0:
        raise FileNotFoundError(f"no Fields_k_*.npz in {fields_dir}")
    per = N // nslabs                                                                                       # :376
    uk = np.zeros((3, N, N, N // 2 + 1), dtype=np.complex128)
    n = np.ones((N, N, N))
    for s in range(nslabs):
        F = np.load(fields_dir / f"Fields_k_{s}.npz")
        sl = slice(s * per, (s + 1) * per)
        uk[0][:, sl] = F["uk"]                                                                              # :395-397
        uk[1][:, sl] = F["vk"]
        uk[2][:, sl] = F["wk"]
        if n_dir is not None:
            n[sl] = np.load(Path(n_dir) / f"n_{s}.npz")["n"]                                               # :398
    if mask:
        uk *= dealias_mask(N)
    return uk, n


def load_droplets(state_dir):
    """Rows (Np, 7) [x, y, z, vx, vy, vz, m/M0] and int64 ids from state_{r}.npz
    files (forced-dns-sm.py:675-680)."""
    state_dir = Path(state_dir)
    files = sorted(state_dir.glob("state_*.npz"), key=lambda p: int(p.stem.split("_")[1]))
    rows, ids = [], []
    for f in files:
        d = np.load(f)
        c = np.zeros((d["pos"].shape[0], 7))
        c[:, :3] = d["pos"]
        c[:, 3:6] = d["vel"]
        c[:, 6] = d["mass"]
        rows.append(c)
        ids.append(np.asarray(d["prtclid"]).reshape(-1).astype(np.int64))
    if not rows:
        return np.zeros((0, 7)), np.zeros(0, dtype=np.int64)
    return np.concatenate(rows), np.concatenate(ids)
