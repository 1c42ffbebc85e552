"""Single-GPU drop-in for ``MPI_particles`` (particles.py:52-837), hot-path methods.

The droplet state ``coord`` is a (Np, 7) float64 CUDA tensor
[x, y, z, vx, vy, vz, m/M0]; ``prtclid`` an int64 (Np,) tensor.  On one device
rows are never migrated, so ``pRHS`` keeps row order (the reference reorders
rows in ``send``; results must be compared by ``prtclid``).
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional

import numpy as np
import torch

from . import _lib
from .dns import ForcedDNS, _ptr


class MPI_particles:
    """Constructor arguments follow particles.py:54; ``comm`` and the meshgrid
    arrays are accepted and unused (one slab).  ``dns`` supplies the context,
    ``coord``/``prtclid`` an explicit population, ``rng`` a seeded generator
    for the reference's own initialisation (particles.py:91-96)."""

    def __init__(self, comm, L, N, Nprtcl, st_s, st, g, nu, tau_eta, rho_p, M0, d=3, X=None, Y=None, Z=None,
                 x=None, y=None, z=None, *, dns: ForcedDNS, coord=None, prtclid=None,
                 rng: Optional[np.random.Generator] = None):
        if d != 3:
            raise ValueError("d must be 3")
        self.dns = dns
        self.L, self.N, self.Nprtcl, self.d = L, N, Nprtcl, d
        self.g, self.tau_eta, self.st_s = g, tau_eta, st_s
        self.dx = self.dy = self.dz = L / N
        self.factor = (nu * tau_eta / (2 * rho_p)) ** 3 / 2 * (36 * math.pi * rho_p) / M0   # :74
        self.growthfactor = 9 * math.pi * nu / (2 * rho_p)                                  # :75
        dev = dns.device
        if coord is None:
            rng = np.random.default_rng() if rng is None else rng
            c = rng.uniform(0, L, (Nprtcl, 2 * d + 1))                                        # :91
            c[:, -1] = st ** 1.5 * self.factor                                                # :94
            coord = c
        if isinstance(coord, np.ndarray):
            coord = torch.from_numpy(np.ascontiguousarray(coord, dtype=np.float64)).to(dev)
        self.coord = coord.contiguous()
        n = self.coord.shape[0]
        if prtclid is None:
            prtclid = torch.arange(n, dtype=torch.int64, device=dev)                         # :92
        elif isinstance(prtclid, np.ndarray):
            prtclid = torch.from_numpy(prtclid.astype(np.int64)).to(dev)
        self.prtclid = prtclid
        self.update_intrinsic()
        self.interpmat = torch.zeros((n, 7), dtype=torch.float64, device=dev)
        self.exterpmat = torch.ones((n, 1), dtype=torch.float64, device=dev)

    # particles.py:113, :128
    def to_interp(self, interpdim):
        self.interpdim = interpdim
        self.interpmat = torch.zeros((self.coord.shape[0], interpdim), dtype=torch.float64, device=self.dns.device)

    def to_exterp(self, exterpdim):
        if exterpdim != 1:
            raise ValueError("only scalar deposits (exterpdim = 1) are on the hot path")
        self.exterpdim = exterpdim
        self.exterpmat = torch.ones((self.coord.shape[0], 1), dtype=torch.float64, device=self.dns.device)

    # particles.py:191-195
    def update_intrinsic(self):
        self.st = (self.coord[:, -1] / self.factor) ** (2 / 3.0)
        self.rhs = 0.0 * self.coord

    # particles.py:233-275 (one rank)
    def send(self, x, args):
        """Rows whose x0 left [0, L) move to the end ordered by x0 mod L; x[:, :3]
        is wrapped; args are permuted alike and not wrapped.  Stable order is
        used where the reference's quicksort leaves ties unordered."""
        L = self.L
        out = (x[:, 0] < 0) | (x[:, 0] >= L)
        keep = torch.nonzero(~out).flatten()
        mv = torch.nonzero(out).flatten()
        key = torch.remainder(x[mv, 0], L)
        mv = mv[torch.argsort(key, stable=True)]
        mv = mv[torch.remainder(x[mv, 0], L) < L]            # the reference's send count drops x0 mod L == L
        perm = torch.cat([keep, mv])
        xn = x[perm].clone()
        xn[:, : self.d] = torch.remainder(xn[:, : self.d], L)
        return xn, [a[perm] for a in args]

    # particles.py:435-511
    def interp_cosine(self, coord, u):
        """Cosine-kernel gather of every component of u ((ncomp, N, N, N))."""
        ncomp = u.shape[0]
        self.dns._chk(u, self.dns.rdtype, (ncomp,) + (self.N,) * 3, "u")
        coord, npart = self.dns._rows(coord)
        out = torch.zeros((npart, ncomp), dtype=torch.float64, device=self.dns.device)
        self.dns._stream()
        _lib.check(self.dns.lib.ld_interp_cosine(self.dns.ctx, _ptr(coord), npart, _ptr(u), ncomp, _ptr(out)))
        self.interpmat = out
        return out

    # particles.py:514-577
    def exterp_cosine_scalar(self, coord, c):
        """c = deposit of exterpmat[:, 0] / dx^3 (c zeroed first)."""
        self.dns._chk(c, self.dns.rdtype, (self.N,) * 3, "c")
        coord, npart = self.dns._rows(coord)
        w = self.exterpmat[:, 0].contiguous()
        self.dns._stream()
        _lib.check(self.dns.lib.ld_exterp_cosine_scalar(self.dns.ctx, _ptr(coord), npart, _ptr(w), _ptr(c)))
        return c

    # ------------------------------------------------------- collisions (no reference counterpart)
    def radius(self, nu, rho_p):
        """Droplet radius r = sqrt(9 nu tau_eta St / (2 rho_p)) (Stokes number of
        a sphere; the reference's `factor` is not a radius, see its **3/2)."""
        st = (self.coord[:, -1] / self.factor) ** (2 / 3.0)
        return torch.sqrt(9 * nu * self.tau_eta * st / (2 * rho_p)).contiguous()

    def collision_pairs(self, radius, max_pairs=1 << 20):
        """Sorted (i, j) row pairs with minimum-image distance below r_i + r_j."""
        dns = self.dns
        coord, npart = dns._rows(self.coord)
        radius = radius.to(torch.float64).contiguous()
        rmax = float(radius.max().item()) if npart else 0.0
        out = torch.empty((max_pairs, 2), dtype=torch.int64, device=dns.device)
        cnt = ctypes.c_int64()
        dns._stream()
        _lib.check(dns.lib.ld_collide_pairs(dns.ctx, _ptr(coord), _ptr(radius), npart, rmax, _ptr(out), max_pairs,
                                            ctypes.byref(cnt)))
        return out[: cnt.value].clone()

    def pair_statistics(self, rmax, nbins=64):
        """Pair statistics of the current rows within the cutoff ``rmax``:
        separation histogram, radial relative velocity sums, and from them the
        radial distribution function g(r), the mean |w_r| and the kinematic
        collision kernel K(r) = 2 pi r^2 g(r) <|w_r|>(r) at the bin centres
        (Sundaram & Collins 1997).  Returns a dict of numpy arrays."""
        dns = self.dns
        coord, npart = dns._rows(self.coord)
        counts = torch.empty((nbins,), dtype=torch.int64, device=dns.device)
        wsum = torch.empty((nbins,), dtype=torch.float64, device=dns.device)
        wneg = torch.empty((nbins,), dtype=torch.float64, device=dns.device)
        dns._stream()
        _lib.check(dns.lib.ld_pair_stats(dns.ctx, _ptr(coord), npart, float(rmax), int(nbins), _ptr(counts),
                                         _ptr(wsum), _ptr(wneg)))
        c = counts.cpu().numpy()
        ws = wsum.cpu().numpy()
        wn = wneg.cpu().numpy()
        edges = np.arange(nbins + 1) * (rmax / nbins)
        centre = 0.5 * (edges[1:] + edges[:-1])
        shell = 4.0 / 3.0 * math.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
        pairs_per_volume = npart * (npart - 1) / 2.0 / self.L ** 3
        with np.errstate(invalid="ignore", divide="ignore"):
            g = c / (pairs_per_volume * shell) if npart > 1 else np.zeros(nbins)
            wr = np.where(c > 0, ws / np.maximum(c, 1), np.nan)
            wneg_mean = np.where(c > 0, wn / np.maximum(c, 1), np.nan)
        return {"edges": edges, "counts": c, "wsum": ws, "wneg": wn, "g": g, "wr_abs_mean": wr,
                "wr_in_mean": wneg_mean, "kernel": 2.0 * math.pi * centre ** 2 * g * wr}

    def coalesce(self, pairs):
        """Apply the greedy merge rule in place on coord; returns the (i, j) events."""
        dns = self.dns
        coord, npart = dns._rows(self.coord)
        pairs = pairs.contiguous()
        ev = torch.empty_like(pairs)
        ne = ctypes.c_int64()
        dns._stream()
        _lib.check(dns.lib.ld_coalesce(dns.ctx, _ptr(coord), npart, _ptr(pairs), pairs.shape[0], _ptr(ev),
                                       ctypes.byref(ne)))
        self.update_intrinsic()
        return ev[: ne.value].clone()

    # particles.py:817-837
    def pRHS(self, t, coord, u, us, c, fc, sump):
        """Droplet RHS at the stage rows ``coord``: returns (sump, rhs, fc) and sets
        interpmat, exterpmat, st, rhs."""
        dns = self.dns
        coord, npart = dns._rows(coord)
        dns._chk(u, dns.rdtype, (3,) + (self.N,) * 3, "u")
        dns._chk(us, dns.rdtype, (3,) + (self.N,) * 3, "us")
        dns._chk(c, dns.rdtype, (self.N,) * 3, "c")
        dns._chk(fc, dns.rdtype, (self.N,) * 3, "fc")
        rhs = torch.zeros((npart, 7), dtype=torch.float64, device=dns.device)
        interp = torch.zeros((npart, 7), dtype=torch.float64, device=dns.device)
        ext = torch.zeros((npart,), dtype=torch.float64, device=dns.device)
        dns._stream()
        _lib.check(dns.lib.ld_prhs(dns.ctx, _ptr(coord), npart, _ptr(u), _ptr(us), _ptr(c), _ptr(fc), _ptr(rhs),
                                   _ptr(interp), _ptr(ext)))
        self.interpmat = interp
        self.exterpmat = ext[:, None]
        self.st = (coord[:, -1] / self.factor) ** (2 / 3.0)
        self.rhs = rhs
        return sump, rhs, fc
