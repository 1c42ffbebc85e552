"""Single-rank FP64 restatement of the reference's spectral flow step.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``forced-dns-sm.py`` with one MPI rank (``Np == N``):
  grid / wavenumbers        forced-dns-sm.py:49-62
  operators                 forced-dns-sm.py:123-141
  rfft_mpi / irfft_mpi      forced-dns-sm.py:200-213
  e3d_to_e1d                forced-dns-sm.py:232-233
  forcing                   forced-dns-sm.py:239-275
  clip_zero                 forced-dns-sm.py:278-285
  full_RHS                  forced-dns-sm.py:287-332
  RK4                       forced-dns-sm.py:335-362
  step epilogue             forced-dns-sm.py:582-620

Layouts (C order): physical ``(3, N, N, N)`` as ``[c][x][y][z]``; spectral
``(3, N, N, N//2+1)`` as ``[c][kx][ky][kz]``.  Forward transforms are
unnormalised, inverse transforms carry ``1/N**3``.
"""
from __future__ import annotations

import numpy as np
from scipy.fft import fft, ifft, irfft, irfft2, rfft, rfft2, fftfreq

# scipy.fft worker count.  The reference passes none (one thread per rank);
# the CPU baseline may raise it.
WORKERS = 1


class BlowUp(RuntimeError):
    """Raised where the reference calls ``comm.Abort()`` (forced-dns-sm.py:618-620)."""


class Grid:
    """Grid, wavenumbers and spectral operators for one slab holding the whole box.

    forced-dns-sm.py:49-62 (grid, ``Kx, Ky, Kz``, meshgrid ``[kx][ky][kz]``) and
    :123-141 (``lap, k, kint, dealias, invlap, vis, normalize, shells``).
    """

    def __init__(self, N: int, L: float = 2 * np.pi, nu: float = 0.0, lp: int = 1):
        self.N = N
        self.L = L
        self.Nf = N // 2 + 1
        TWO_PI = 2 * np.pi
        self.X = np.linspace(0, L, N, endpoint=False)          # :55
        self.dx = self.X[1] - self.X[0]                        # :56
        Kx = fftfreq(N, 1.0 / N) * TWO_PI / L                  # :59
        Kz = np.abs(Kx[: self.Nf])                             # :60
        self.kx, self.ky, self.kz = np.meshgrid(Kx, Kx, Kz, indexing="ij")  # :62
        kx, ky, kz = self.kx, self.ky, self.kz
        self.lap = -(kx ** 2 + ky ** 2 + kz ** 2)              # :125
        self.k = (-self.lap) ** 0.5                            # :126
        self.kint = np.clip(np.round(self.k, 0).astype(int), None, N // 2)  # :127
        self.dealias = self.kint <= N / 3                      # :129
        self.invlap = self.dealias / np.where(self.lap == 0, np.inf, self.lap)  # :131
        self.lp = lp
        self.vis = nu * self.k ** (2 * lp)                     # :134
        self.normalize = np.where((kz == 0) + (kz == N // 2),
                                  1 / (N ** 6 / TWO_PI ** 3),
                                  2 / (N ** 6 / TWO_PI ** 3))  # :136
        self.shells = np.arange(-0.5, self.Nf, 1.0)            # :137
        self.shells[0] = 0.0                                   # :138

    # forced-dns-sm.py:201-206 with num_process == 1: the Alltoall is the identity.
    def rfft_mpi(self, u: np.ndarray) -> np.ndarray:
        fu = rfft2(u, axes=(-2, -1), workers=WORKERS)
        return fft(fu, axis=-3, workers=WORKERS)

    # forced-dns-sm.py:208-213 with num_process == 1.
    def irfft_mpi(self, fu: np.ndarray) -> np.ndarray:
        N = self.N
        a = ifft(fu, axis=-3, workers=WORKERS)
        return irfft2(a, (N, N), axes=(-2, -1), workers=WORKERS)

    # forced-dns-sm.py:232-233
    def e3d_to_e1d(self, x: np.ndarray) -> np.ndarray:
        return np.histogram(self.k.ravel(), bins=self.shells, weights=x.ravel())[0]

    def energy3d(self, uk: np.ndarray) -> np.ndarray:
        """forced-dns-sm.py:243 (also :423, :722, :740)."""
        return 0.5 * (np.abs(uk[0]) ** 2 + np.abs(uk[1]) ** 2 + np.abs(uk[2]) ** 2) * self.normalize

    # forced-dns-sm.py:216-229 (diff_x, diff_y, diff_z) with num_process == 1:
    # a 1-D rfft along the axis, times i k with k = 0..N/2, then irfft.
    def diff(self, u: np.ndarray, axis: int) -> np.ndarray:
        N = self.N
        k = np.arange(N // 2 + 1, dtype=np.float64)           # kx_diff / ky_diff / kz_diff (:69-71)
        shape = [1, 1, 1]
        shape[axis] = N // 2 + 1
        return irfft(1j * k.reshape(shape) * rfft(u, axis=axis), N, axis=axis)

    def divergence(self, u: np.ndarray) -> np.ndarray:
        """diff_x(u) + diff_y(v) + diff_z(w) of a physical velocity (:472)."""
        return self.diff(u[0], 0) + self.diff(u[1], 1) + self.diff(u[2], 2)

    def dissipation_spectrum(self, uk: np.ndarray, nu: float) -> np.ndarray:
        """e3d_to_e1d(nu |k|^2 sum_c |u_c|^2 normalize): its sum is 2 nu sum_k k^2 E(k),
        the dissipation the reference's commented-out line :473 aims at."""
        return self.e3d_to_e1d(nu * self.k ** 2 * 2.0 * self.energy3d(uk))


class Flow:
    """Reference step state and functions, one rank, FP64.

    Parameters mirror the reference's module constants:
      nu, lp, isexplicit   forced-dns-sm.py:21-24, 78-83
      shell_no, f0         forced-dns-sm.py:89-90, 97
      isforcing            forced-dns-sm.py:19
      tps (= sts * tf), g  forced-dns-sm.py:92-94
    """

    def __init__(self, N, nu, shell_no, f0, tps, g=0.0, isforcing=True,
                 isexplicit=1.0, lp=1, L=2 * np.pi):
        self.grid = Grid(N, L, nu, lp)
        self.N = N
        self.nu = nu
        self.lp = lp
        self.isexplicit = isexplicit
        self.shell_no = np.asarray(shell_no, dtype=int)
        self.f0 = f0
        self.tps = tps
        self.g = g
        self.isforcing = isforcing

    # ------------------------------------------------------------------ forcing
    def forcing(self, uk: np.ndarray) -> np.ndarray:
        """Constant-power forcing of the shells ``shell_no`` (forced-dns-sm.py:239-275)."""
        G = self.grid
        ek = G.energy3d(uk)                                              # :243
        ek_arr = G.e3d_to_e1d(ek)                                        # :245
        ek_arr = np.where(np.abs(ek_arr) < 1e-10, np.inf, ek_arr)        # :250
        factor = np.zeros(G.Nf)                                          # :253
        factor[self.shell_no] = self.f0 / (2 * ek_arr[self.shell_no])    # :254
        factor3d = factor[G.kint]                                        # :255
        fk = np.empty_like(uk)
        fk[0] = factor3d * uk[0] * G.dealias                             # :262
        fk[1] = factor3d * uk[1] * G.dealias
        fk[2] = factor3d * uk[2] * G.dealias
        pk = G.invlap * (G.kx * fk[0] + G.ky * fk[1] + G.kz * fk[2]) * G.dealias  # :268
        fk[0] = fk[0] + G.kx * pk                                        # :270
        fk[1] = fk[1] + G.ky * pk
        fk[2] = fk[2] + G.kz * pk
        return fk * self.isforcing * G.dealias                           # :275

    def shell_factors(self, uk: np.ndarray) -> np.ndarray:
        """The ``factor`` array of forced-dns-sm.py:253-254 (exposed for tests)."""
        G = self.grid
        ek_arr = G.e3d_to_e1d(G.energy3d(uk))
        ek_arr = np.where(np.abs(ek_arr) < 1e-10, np.inf, ek_arr)
        factor = np.zeros(G.Nf)
        factor[self.shell_no] = self.f0 / (2 * ek_arr[self.shell_no])
        return factor

    # ---------------------------------------------------------------- clip_zero
    def clip_zero(self, x: np.ndarray) -> np.ndarray:
        """forced-dns-sm.py:278-285.  Clips ``x`` IN PLACE, returns the rescaled copy."""
        N = self.N
        oldmean = np.sum(x) / N ** 3
        x[:] = np.clip(x, 0, None)
        newmean = np.sum(x) / N ** 3
        return x * oldmean / newmean

    # ----------------------------------------------------------------- full_RHS
    def full_RHS(self, t, uk, n, sump, coord, stb, visc=1, forc=1):
        """One right-hand-side evaluation (forced-dns-sm.py:287-332).

        Returns ``(ku, kn, kp, sump)`` like the reference; ``stb`` is an
        ``oracle.Droplets`` (or ``None`` for a particle-free flow, which is the
        reference's ``Nprtcl = 0`` case).  Side outputs of the last call are kept
        on ``self.last`` for tests: ``u, v, omg, fk, fc, pk``.
        """
        G = self.grid
        kx, ky, kz = G.kx, G.ky, G.kz
        nu, lp, isexp, tps = self.nu, self.lp, self.isexplicit, self.tps
        fk = self.forcing(uk) * forc                                     # :289
        u = np.empty((3,) + (self.N,) * 3)
        u[0] = G.irfft_mpi(uk[0])                                        # :291-293
        u[1] = G.irfft_mpi(uk[1])
        u[2] = G.irfft_mpi(uk[2])
        omg = np.empty_like(u)
        omg[0] = G.irfft_mpi(1j * (ky * uk[2] - kz * uk[1]))             # :295-297
        omg[1] = G.irfft_mpi(1j * (kz * uk[0] - kx * uk[2]))
        omg[2] = G.irfft_mpi(1j * (kx * uk[1] - ky * uk[0]))
        rhsu = omg[2] * u[1] - omg[1] * u[2]                             # :299-301
        rhsv = omg[0] * u[2] - omg[2] * u[0]
        rhsw = omg[1] * u[0] - omg[0] * u[1]
        rhsuk = (G.rfft_mpi(rhsu) + fk[0]) * G.dealias                   # :304-306
        rhsvk = (G.rfft_mpi(rhsv) + fk[1]) * G.dealias
        rhswk = (G.rfft_mpi(rhsw) + fk[2]) * G.dealias
        pk = 1j * G.invlap * (kx * rhsuk + ky * rhsvk + kz * rhswk)      # :309
        klap = (-G.lap) ** lp
        ku = np.empty_like(uk)
        ku[0] = rhsuk - 1j * kx * pk - nu * klap * uk[0] * isexp * visc  # :314-316
        ku[1] = rhsvk - 1j * ky * pk - nu * klap * uk[1] * isexp * visc
        ku[2] = rhswk - 1j * kz * pk - nu * klap * uk[2] * isexp * visc
        v = np.empty_like(u)
        v[0] = G.irfft_mpi((uk[0] - tps * (-1j * kx * pk - nu * klap * uk[0] * isexp * visc)) * G.dealias)  # :319
        v[1] = G.irfft_mpi((uk[1] - tps * (-1j * ky * pk - nu * klap * uk[1] * isexp * visc)) * G.dealias)  # :320
        v[2] = G.irfft_mpi((uk[2] - tps * (-1j * kz * pk - nu * klap * uk[2] * isexp * visc)) * G.dealias) - tps * self.g  # :321
        fc = np.zeros_like(n)
        if stb is not None:
            sump, kp, fc = stb.pRHS(t, coord, u, v, n, fc, sump)        # :324
        else:
            kp = np.zeros((0, 7))
        fck = G.rfft_mpi(fc) * G.dealias                                 # :325
        vn = n[None, ...] * v                                            # :326
        vnk0 = G.rfft_mpi(vn[0]) * G.dealias                             # :327-329
        vnk1 = G.rfft_mpi(vn[1]) * G.dealias
        vnk2 = G.rfft_mpi(vn[2]) * G.dealias
        kn = G.irfft_mpi((-1j * (kx * vnk0 + ky * vnk1 + kz * vnk2) - fck) * G.dealias)  # :330
        self.last = dict(u=u, v=v, omg=omg, fk=fk, fc=fc, pk=pk)
        return ku, kn, kp, sump

    # ---------------------------------------------------------------------- RK4
    def RK4(self, t, h, stb, uk, n):
        """Classical RK4 on (uk, n, droplets) (forced-dns-sm.py:335-362).

        ``n`` is clipped in place by the first stage's ``clip_zero`` exactly as
        in the reference (:339, :341).  Returns ``(sump, uknew, nnew)``.
        """
        empty = np.zeros((0, 7))
        sump = 1.0 * (stb.coord if stb is not None else empty)          # :337
        uknew = 1.0 * uk                                                 # :338
        nnew = 1.0 * n                                                   # :339
        c0 = stb.coord if stb is not None else empty
        ku, kn, kp, sump = self.full_RHS(t, uk, self.clip_zero(n), sump, c0, stb)  # :341
        sump += h / 6.0 * kp                                             # :342
        uknew += h / 6.0 * ku
        nnew += h / 6.0 * kn
        c0 = stb.coord if stb is not None else empty
        ku, kn, kp, sump = self.full_RHS(t + h / 2, uk + ku * h / 2, self.clip_zero(n + kn * h / 2),
                                         sump, c0 + h / 2 * kp, stb)     # :347
        sump += h / 3.0 * kp
        uknew += h / 3.0 * ku
        nnew += h / 3.0 * kn
        c0 = stb.coord if stb is not None else empty
        ku, kn, kp, sump = self.full_RHS(t + h / 2, uk + ku * h / 2, self.clip_zero(n + kn * h / 2),
                                         sump, c0 + h / 2 * kp, stb)     # :352
        sump += h / 3.0 * kp
        uknew += h / 3.0 * ku
        nnew += h / 3.0 * kn
        c0 = stb.coord if stb is not None else empty
        ku, kn, kp, sump = self.full_RHS(t + h, uk + ku * h, self.clip_zero(n + kn * h),
                                         sump, c0 + h * kp, stb)         # :357
        sump += h / 6.0 * kp
        uknew += h / 6.0 * ku
        nnew += h / 6.0 * kn
        return sump * 1.0, uknew, nnew                                   # :362

    # --------------------------------------------------------------- epilogue
    def reality_and_projection(self, uknew: np.ndarray) -> np.ndarray:
        """forced-dns-sm.py:597-610: irfft/rfft round trip, then ``uk += k (invlap k.uk)``."""
        G = self.grid
        uk = np.empty_like(uknew)
        for c in range(3):
            uk[c] = G.rfft_mpi(G.irfft_mpi(uknew[c]))                    # :597-603
        pk = G.invlap * (G.kx * uk[0] + G.ky * uk[1] + G.kz * uk[2])     # :607
        uk[0] = uk[0] + G.kx * pk                                        # :608-610
        uk[1] = uk[1] + G.ky * pk
        uk[2] = uk[2] + G.kz * pk
        return uk

    def step(self, t, h, stb, uk, n, hypervisc=1.0):
        """One pass of the hot loop body (forced-dns-sm.py:582-620).

        Returns ``(uk, n)``; droplet state is updated on ``stb``.
        """
        sump, uknew, nnew = self.RK4(t, h, stb, uk, n)                   # :582
        if stb is not None:
            stb.coord = sump
            stb.update_intrinsic()                                       # :583
        n = self.clip_zero(nnew)                                         # :584
        uknew = uknew * hypervisc                                        # :585
        uk = self.reality_and_projection(uknew)                          # :597-610
        if uk.max() > 100 * self.N ** 3:                                 # :618
            raise BlowUp("Threshold exceeded")
        return uk, n
