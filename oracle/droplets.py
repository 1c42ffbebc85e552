"""Single-rank FP64 restatement of the reference's Lagrangian droplets.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``particles.py`` (class ``MPI_particles``) with one MPI rank:
  constants (factor, growthfactor)     particles.py:54-76
  initial state                        particles.py:91-96 (seeded here)
  to_interp / to_exterp buffers        particles.py:113-142
  update_intrinsic                     particles.py:191-195
  send (wrap + reorder, 1 rank)        particles.py:233-275
  interp_cosine + _calc_usend_numba    particles.py:435-511, 145-158, 6-20
  exterp_cosine_scalar + uadd kernel   particles.py:514-577, 160-172, 23-35
  pRHS                                 particles.py:817-837

Differences that are intended:
  * serial, non-fastmath arithmetic (the reference's numba scatter at
    particles.py:23-35 is a ``prange`` race; any serial order is one of its
    admissible results);
  * the deceleration expression of particles.py:833 is a discarded expression
    statement in the reference (it starts a new line with ``+``), so it is not
    part of the acceleration and is not restated.

Droplet state row layout ``(Np, 7)``: ``[x, y, z, vx, vy, vz, m/M0]``
(particles.py:91).
"""
from __future__ import annotations

import numpy as np


def droplet_constants(nu, tau_eta, rho_p, M0):
    """``factor`` and ``growthfactor`` exactly as particles.py:74-75.

    Note the reference writes ``(nu*tau_eta/(2*rho_p))**3/2`` which Python reads
    as ``x**3 / 2``; that is reproduced.
    """
    factor = (nu * tau_eta / (2 * rho_p)) ** 3 / 2 * (36 * np.pi * rho_p) / M0
    growthfactor = 9 * np.pi * nu / (2 * rho_p)
    return factor, growthfactor


class Droplets:
    """One-rank ``MPI_particles``.

    Parameters mirror ``MPI_particles.__init__`` (particles.py:54) minus the
    communicator and the meshgrid arrays it never uses on the hot path.
    """

    def __init__(self, L, N, coord, st_s, g, nu, tau_eta, rho_p, M0, prtclid=None):
        self.L, self.N, self.d = L, N, 3
        self.g, self.tau_eta, self.st_s = g, tau_eta, st_s
        self.X = np.linspace(0, L, N, endpoint=False)
        self.Y = self.Z = self.X
        self.dx = self.dy = self.dz = self.X[1] - self.X[0]          # :58-60
        self.Np = N                                                   # :64
        self.startdom = self.X[0]                                     # :67
        self.enddom = self.startdom + self.Np * self.dx               # :69
        self.factor, self.growthfactor = droplet_constants(nu, tau_eta, rho_p, M0)
        self.coord = np.array(coord, dtype=np.float64).reshape(-1, 7)
        if prtclid is None:
            prtclid = np.arange(self.coord.shape[0]).reshape((-1, 1))  # :92-93
        self.prtclid = np.asarray(prtclid).reshape((-1, 1))
        self.st = (self.coord[:, -1] / self.factor) ** (2 / 3.0)      # :96
        self.cosorder = np.arange(-1, 3)                              # :109
        self.interpdim = 7
        self.interpmat = np.zeros((self.coord.shape[0], 7))           # :118
        self.exterpmat = np.ones((self.coord.shape[0], 1))            # :132
        self.rhs = 0.0 * self.coord

    # ----------------------------------------------------------- particles.py:191
    def update_intrinsic(self):
        self.st = (self.coord[:, -1] / self.factor) ** (2 / 3.0)
        self.rhs = 0.0 * self.coord

    # ----------------------------------------------------------- particles.py:233
    def send(self, x, args):
        """Single-rank migration: rows whose ``x0`` left ``[0, L)`` move to the end,
        ordered by ``x0 mod L``; only ``x[:, :3]`` is wrapped (particles.py:235-275).
        A row whose ``x0 mod L`` rounds to exactly ``L`` is dropped, as the
        reference's send count (:255) drops it.
        """
        L = self.L
        outcond = (x[:, 0] < self.startdom) + (x[:, 0] >= self.enddom)   # :235
        cond = ~outcond
        sendbuf = x[outcond].copy()                                     # :237
        for a in args:                                                  # :243-244
            sendbuf = np.concatenate((sendbuf, a[outcond]), axis=1)
        cdim = sendbuf.shape[-1]
        sendbuf = sendbuf[(sendbuf[:, 0] % L).argsort()]                 # :251
        key = sendbuf[:, 0] % L
        count = int(np.sum((key >= self.startdom) * (key < self.enddom)))  # :255
        recvbuf = sendbuf[:count].reshape((-1, cdim))                   # :261-266 (self-send)
        x = np.concatenate((x[cond], recvbuf[:, : x.shape[-1]]), axis=0)  # :268
        x[:, : self.d] %= L                                             # :269
        out = []
        col = x.shape[-1]
        for a in args:                                                  # :271-273
            out.append(np.concatenate((a[cond], recvbuf[:, col: col + a.shape[-1]]), axis=0))
            col += a.shape[-1]
        return x, out

    # ------------------------------------------------- particles.py:6-20, 145-158
    def _usend(self, ufield, xidx, yidx, zidx, ypos, zpos):
        """Serial ``_calc_usend_numba``: sum over the 4x4 (y, z) stencil, left to right."""
        N, dy, dz = self.N, self.dy, self.dz
        comp = ufield.shape[0]
        usend = np.zeros((xidx.shape[0], comp))
        for i in self.cosorder:                                         # :10
            ty = (yidx + i) % N
            dyshift = (1 + np.cos((ypos - self.Y[ty]) * (np.pi / (2 * dy)))) / 4.0  # :12
            for j in self.cosorder:                                     # :14
                tz = (zidx + j) % N
                dzshift = (1 + np.cos((zpos - self.Z[tz]) * (np.pi / (2 * dz)))) / 4.0  # :16
                vals = ufield[:, xidx, ty, tz].T                        # (Np, comp)
                usend += vals * dyshift[:, None] * dzshift[:, None]     # :19
        return usend

    # ----------------------------------------------------------- particles.py:435
    def interp_cosine(self, coord, u):
        """4x4x4 cosine-kernel gather of every component of ``u`` (particles.py:435-511).

        One rank: the neighbour requests of :457-504 resolve to the wrapped
        index ``slx % N`` (:480), so every particle is handled identically.
        """
        pos = coord[:, : self.d]
        idx = (pos // self.dx).astype(np.int32)                         # :441
        idx[:, 0] %= self.Np                                            # :442
        out = np.zeros((coord.shape[0], u.shape[0]))                   # :444
        for i in range(4):                                              # :454
            slx = idx[:, 0] - 4 // 2 + 1 + i                            # :455
            dxshift = (1 + np.cos((pos[:, 0] - self.X[slx % self.N]) * (np.pi / (2 * self.dx)))) / 4.0  # :456
            usend = self._usend(u, slx % self.N, idx[:, 1], idx[:, 2], pos[:, 1], pos[:, 2])  # :489, :509
            out += usend * dxshift[:, None]                             # :504, :509
        self.interpmat = out
        return out

    # ----------------------------------------------------------- particles.py:514
    def exterp_cosine_scalar(self, coord, c):
        """Deposit ``exterpmat[:, 0]`` with the same kernel, divided by the cell
        volume (particles.py:514-577, 23-35).  Zeroes ``c`` first (:521).
        """
        N = self.N
        pos = coord[:, : self.d]
        idx = (pos // self.dx).astype(np.int32)                         # :518
        idx[:, 0] %= self.Np                                            # :519
        c *= 0.0                                                        # :521
        vol = self.dx * self.dy * self.dz
        flat = c.reshape(-1)
        W = self.exterpmat[:, 0]
        for i in range(4):                                              # :529
            slx = idx[:, 0] - 4 // 2 + 1 + i                            # :530
            dxshift = (1 + np.cos((pos[:, 0] - self.X[slx % N]) * (np.pi / (2 * self.dx)))) / 4.0  # :531
            udat = W * dxshift                                          # :536, :575
            xi = slx % N
            # serial order: particle-major, then the (y, z) stencil (:26-34)
            idxs = np.empty((coord.shape[0], 16), dtype=np.int64)
            vals = np.empty((coord.shape[0], 16))
            s = 0
            for a in self.cosorder:
                ty = (idx[:, 1] + a) % N
                dyshift = (1 + np.cos((pos[:, 1] - self.Y[ty]) * (np.pi / (2 * self.dy)))) / 4.0  # :29
                for b in self.cosorder:
                    tz = (idx[:, 2] + b) % N
                    dzshift = (1 + np.cos((pos[:, 2] - self.Z[tz]) * (np.pi / (2 * self.dz)))) / 4.0  # :33
                    idxs[:, s] = (xi * N + ty) * N + tz
                    vals[:, s] = udat * dyshift * dzshift / vol        # :34
                    s += 1
            flat += np.bincount(idxs.ravel(), weights=vals.ravel(), minlength=flat.size)
        return c

    # ----------------------------------------------------------- particles.py:817
    def pRHS(self, t, coord, u, us, c, fc, sump):
        """Droplet right-hand side (particles.py:817-837).

        Returns ``(sump, rhs, fc)``; ``self.coord``, ``interpmat``, ``exterpmat``,
        ``prtclid`` and ``sump`` are permuted by ``send`` as in :822.
        """
        d = self.d
        coord, (self.coord, self.interpmat, self.exterpmat, self.prtclid, sump) = self.send(
            coord, [self.coord, self.interpmat, self.exterpmat, self.prtclid, sump])  # :822
        self.st = (coord[:, -1] / self.factor) ** (2 / 3.0)            # :823
        self.rhs = 0.0 * coord                                          # :824
        self.interpmat = self.interp_cosine(coord, np.concatenate((u, us, c[None, ...]), axis=0))  # :826
        self.exterpmat[:, 0] = (self.growthfactor * (1 + (self.st_s / self.st) ** 0.5) ** 2
                                * self.interpmat[:, -1]
                                * np.linalg.norm(coord[:, d: 2 * d] - self.interpmat[:, d: 2 * d], axis=1))  # :828
        fc[:] = self.exterp_cosine_scalar(coord, fc)                    # :829
        self.rhs[:, :d] = coord[:, d: 2 * d]                            # :831
        self.rhs[:, d: 2 * d] = (self.interpmat[:, :d] - coord[:, d: 2 * d]) / (self.st[:, None] * self.tau_eta)  # :832
        self.rhs[:, 2 * d - 1] -= self.g                                # :834
        self.rhs[:, -1] = self.exterpmat[:, 0]                          # :835
        return sump, self.rhs, fc

    # ------------------------------------------------------------- helpers
    def by_id(self, arr=None):
        """Rows of ``arr`` (default ``coord``) ordered by ``prtclid`` (the reference
        reorders rows in ``send``; comparisons must be keyed by id)."""
        if arr is None:
            arr = self.coord
        order = np.argsort(self.prtclid[:, 0].astype(np.int64), kind="stable")
        return arr[order]
