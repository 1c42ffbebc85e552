"""Single-GPU drop-in for the hot path of forced-dns-sm.py.

``ForcedDNS`` carries what the reference keeps in module globals (grid size,
viscosity, forcing, particle-coupling constants) and exposes the reference's
functions with their signatures and in-place semantics:

  rfft_mpi(u, fu)                  forced-dns-sm.py:201
  irfft_mpi(fu, u)                 forced-dns-sm.py:208
  forcing(uk, fk)                  forced-dns-sm.py:239
  clip_zero(x)                     forced-dns-sm.py:278
  full_RHS(t, uk, n, sump, coord, stb, ku, kn, fc, visc, forc)   :287
  RK4(t, h, stb, uk, n, uknew, nnew)                             :335
  step(t, h, stb, uk, n)           the loop body :582-620

Arrays are torch tensors on the GPU (complex64/float32 for precision "f32",
complex128/float64 for "f64"; droplet rows always float64).  Every call goes
through libld.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from .config import Config


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class ForcedDNS:
    def __init__(self, cfg: Config, max_particles: Optional[int] = None, device: str = "cuda", batch: int = 1,
                 graph: bool = False, autotune: bool = False):
        """batch > 1: an ensemble stepped with one launch per kernel (only the
        resident-state methods accept it); graph: advance() replays a captured
        HIP graph of one step; autotune: pick the y-pass launch geometry by
        timing at creation.  None of the three changes the results."""
        if not torch.cuda.is_available():
            raise RuntimeError("ForcedDNS needs a GPU (libld.so has no CPU path)")
        self.cfg = cfg
        self.N = cfg.N
        self.Nf = cfg.N // 2 + 1
        self.device = torch.device(device)
        self.rdtype = torch.float32 if cfg.precision == "f32" else torch.float64
        self.cdtype = torch.complex64 if cfg.precision == "f32" else torch.complex128
        self.lib = _lib.load()
        p = _lib.LdParams()
        p.N = cfg.N
        p.precision = _lib.LD_F32 if cfg.precision == "f32" else _lib.LD_F64
        p.L = cfg.L
        p.nu = cfg.nu_eff
        p.lp = cfg.lp
        p.isexplicit = 1.0 if cfg.explicit else 0.0
        p.nshells = cfg.nshells
        for i, s in enumerate(cfg.shell_no):
            p.shell_no[i] = s
        p.f0 = cfg.f0
        p.isforcing = 1.0 if cfg.isforcing else 0.0
        p.tps = cfg.tps
        p.g = cfg.g
        p.st_s = cfg.sts
        p.tau_eta = cfg.tf
        p.factor = cfg.factor
        p.growthfactor = cfg.growthfactor
        p.max_particles = int(cfg.n_particles if max_particles is None else max_particles)
        p.batch = int(batch)
        p.use_graph = 1 if graph else 0
        p.autotune = 1 if autotune else 0
        self.batch = max(1, int(batch))
        self.np = 0
        self.params = p
        self.max_particles = p.max_particles
        ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.ld_create(ctypes.byref(p), ctypes.byref(ctx)))
        self.ctx = ctx

    # ------------------------------------------------------------ plumbing
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.ld_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        self.lib.ld_set_stream(self.ctx, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))

    def _chk(self, t: torch.Tensor, dtype, shape=None, name="array"):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise TypeError(f"{name} must be a CUDA tensor")
        if t.dtype != dtype:
            raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
        return t

    def rshape(self, nf=None):
        return (self.N,) * 3 if nf is None else (nf,) + (self.N,) * 3

    def cshape(self, nf=None):
        s = (self.N, self.N, self.Nf)
        return s if nf is None else (nf,) + s

    def zeros_real(self, nf=None):
        return torch.zeros(self.rshape(nf), dtype=self.rdtype, device=self.device)

    def zeros_spec(self, nf=None):
        return torch.zeros(self.cshape(nf), dtype=self.cdtype, device=self.device)

    @staticmethod
    def _rows(coord):
        if coord is None:
            return None, 0
        if coord.dtype != torch.float64 or coord.dim() != 2 or coord.shape[1] != 7 or not coord.is_contiguous():
            raise ValueError("droplet rows must be a contiguous (Np, 7) float64 CUDA tensor")
        return coord, coord.shape[0]

    # ------------------------------------------------------------ transforms
    def rfft_mpi(self, u: torch.Tensor, fu: torch.Tensor) -> torch.Tensor:
        """3-D r2c of one field or a stack of fields (forced-dns-sm.py:201-206)."""
        nf = 1 if u.dim() == 3 else u.shape[0]
        self._chk(u, self.rdtype, self.rshape(None if u.dim() == 3 else nf), "u")
        self._chk(fu, self.cdtype, self.cshape(None if u.dim() == 3 else nf), "fu")
        self._stream()
        _lib.check(self.lib.ld_rfft(self.ctx, _ptr(u), _ptr(fu), nf))
        return fu

    def irfft_mpi(self, fu: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
        """3-D c2r with 1/N^3 (forced-dns-sm.py:208-213)."""
        nf = 1 if fu.dim() == 3 else fu.shape[0]
        self._chk(fu, self.cdtype, self.cshape(None if fu.dim() == 3 else nf), "fu")
        self._chk(u, self.rdtype, self.rshape(None if fu.dim() == 3 else nf), "u")
        self._stream()
        _lib.check(self.lib.ld_irfft(self.ctx, _ptr(fu), _ptr(u), nf))
        return u

    # ------------------------------------------------------------ physics
    def forcing(self, uk: torch.Tensor, fk: torch.Tensor, ek: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Constant-power forcing of the shells shell_no (forced-dns-sm.py:239-275).

        ek (optional, float64 (N/2+1,)): receives the forced shells' energies."""
        self._chk(uk, self.cdtype, self.cshape(3), "uk")
        self._chk(fk, self.cdtype, self.cshape(3), "fk")
        if ek is not None:
            self._chk(ek, torch.float64, (self.Nf,), "ek")
        self._stream()
        _lib.check(self.lib.ld_forcing(self.ctx, _ptr(uk), _ptr(fk), _ptr(ek)))
        return fk

    def clip_zero(self, x: torch.Tensor) -> torch.Tensor:
        """x is clipped in place; returns x * oldmean / newmean (forced-dns-sm.py:278-285)."""
        self._chk(x, self.rdtype, self.rshape(), "x")
        out = torch.empty_like(x)
        self._stream()
        _lib.check(self.lib.ld_clip_zero(self.ctx, _ptr(x), _ptr(out)))
        return out

    def full_RHS(self, t, uk, n, sump, coord, stb=None, ku=None, kn=None, fc=None, visc=1, forc=1,
                 u_out=None, v_out=None):
        """One RHS evaluation (forced-dns-sm.py:287-332): returns (ku, kn, kp, sump).

        coord are the droplet stage rows (Np, 7) float64 (None for no droplets);
        sump is returned unchanged (rows keep their order on one device)."""
        self._chk(uk, self.cdtype, self.cshape(3), "uk")
        self._chk(n, self.rdtype, self.rshape(), "n")
        ku = self.zeros_spec(3) if ku is None else self._chk(ku, self.cdtype, self.cshape(3), "ku")
        kn = self.zeros_real() if kn is None else self._chk(kn, self.rdtype, self.rshape(), "kn")
        fc = self.zeros_real() if fc is None else self._chk(fc, self.rdtype, self.rshape(), "fc")
        coord, npart = self._rows(coord)
        kp = torch.zeros((npart, 7), dtype=torch.float64, device=self.device)
        if u_out is not None:
            self._chk(u_out, self.rdtype, self.rshape(3), "u_out")
        if v_out is not None:
            self._chk(v_out, self.rdtype, self.rshape(3), "v_out")
        self._stream()
        _lib.check(self.lib.ld_full_rhs(self.ctx, _ptr(uk), _ptr(n), _ptr(coord), npart, _ptr(ku), _ptr(kn),
                                        _ptr(kp), _ptr(fc), float(visc), float(forc), _ptr(u_out), _ptr(v_out)))
        if stb is not None:
            stb.rhs = kp
        return ku, kn, kp, sump

    def RK4(self, t, h, stb, uk, n, uknew=None, nnew=None):
        """Classical RK4 of (uk, n, droplets) (forced-dns-sm.py:335-362).

        Returns (sump, uknew, nnew); n is clipped in place like the reference's
        first clip_zero.  stb may be None (no droplets) or an MPI_particles."""
        self._chk(uk, self.cdtype, self.cshape(3), "uk")
        self._chk(n, self.rdtype, self.rshape(), "n")
        uknew = self.zeros_spec(3) if uknew is None else self._chk(uknew, self.cdtype, self.cshape(3), "uknew")
        nnew = self.zeros_real() if nnew is None else self._chk(nnew, self.rdtype, self.rshape(), "nnew")
        coord, npart = self._rows(stb.coord if stb is not None else None)
        sump = torch.empty_like(coord) if coord is not None else None
        self._stream()
        _lib.check(self.lib.ld_rk4(self.ctx, float(h), _ptr(uk), _ptr(n), _ptr(coord), npart, _ptr(uknew),
                                   _ptr(nnew), _ptr(sump)))
        return sump, uknew, nnew

    def step(self, t, h, stb, uk, n, nsteps: int = 1):
        """nsteps passes of the loop body (forced-dns-sm.py:582-620), in place on
        uk, n and stb.coord.  Raises _lib.BlowUp where the reference aborts."""
        self._chk(uk, self.cdtype, self.cshape(3), "uk")
        self._chk(n, self.rdtype, self.rshape(), "n")
        coord, npart = self._rows(stb.coord if stb is not None else None)
        self._stream()
        rc = self.lib.ld_step(self.ctx, float(h), _ptr(uk), _ptr(n), _ptr(coord), npart, int(nsteps))
        if stb is not None:
            stb.update_intrinsic()
        _lib.check(rc)
        return uk, n

    # ------------------------------------------------------------ resident state (fast path)
    def _bshape(self, shape):
        return shape if self.batch == 1 else (self.batch,) + tuple(shape)

    def set_state(self, uk, n, coord=None):
        """Load the resident state.  With batch B > 1 the arrays carry a leading
        realization axis: uk (B, 3, N, N, Nf), n (B, N, N, N), coord (B, np, 7)."""
        self._chk(uk, self.cdtype, self._bshape(self.cshape(3)), "uk")
        self._chk(n, self.rdtype, self._bshape(self.rshape()), "n")
        if self.batch > 1 and coord is not None:
            if coord.dim() != 3 or coord.shape[0] != self.batch:
                raise ValueError("ensemble droplet rows must be (B, np, 7)")
            coord, npart = self._rows(coord.reshape(-1, 7))
            npart //= self.batch
        else:
            coord, npart = self._rows(coord)
        self._stream()
        _lib.check(self.lib.ld_state_set(self.ctx, _ptr(uk), _ptr(n), _ptr(coord), npart))
        self.np = npart

    def get_state(self, uk=None, n=None, coord=None):
        if uk is not None:
            self._chk(uk, self.cdtype, self._bshape(self.cshape(3)), "uk")
        if n is not None:
            self._chk(n, self.rdtype, self._bshape(self.rshape()), "n")
        if coord is not None:
            self._chk(coord, torch.float64, self._bshape((self.np, 7)), "coord")
        self._stream()
        _lib.check(self.lib.ld_state_get(self.ctx, _ptr(uk), _ptr(n), _ptr(coord)))
        return uk, n, coord

    def advance(self, h, nsteps: int = 1):
        self._stream()
        _lib.check(self.lib.ld_advance(self.ctx, float(h), int(nsteps)))

    def tuning(self):
        """(y-pass lines per block, [median us with M, M/2 and 2M lines])."""
        yl = ctypes.c_int32()
        us = (ctypes.c_double * 3)()
        _lib.check(self.lib.ld_tuning(self.ctx, ctypes.byref(yl), us))
        return yl.value, [us[0], us[1], us[2]]

    def tuning_all(self):
        """Every launch-geometry choice: (dict of choices, dict of timings in us)."""
        ch = (ctypes.c_int32 * 7)()
        us = (ctypes.c_double * 19)()
        _lib.check(self.lib.ld_tuning_all(self.ctx, ch, us))
        names = ["y_lines", "z_cross_tile", "z_vel_tile", "x_uw_lines", "x_rhs_lines", "x_kn_lines", "y_split_n1"]
        times = {"y_lines": list(us[0:3])}
        for k, nm in enumerate(names[1:6]):
            times[nm] = [us[3 + 3 * k], us[4 + 3 * k], us[5 + 3 * k]]
        times["y_half_split_at_chosen_lines"] = [us[18]]
        return dict(zip(names, ch[:])), times

    def status_all(self):
        """(last max Re(u_hat), blown-up flag) per realization, as numpy arrays."""
        import numpy as np
        lm = (ctypes.c_double * self.batch)()
        bl = (ctypes.c_int32 * self.batch)()
        self._stream()
        _lib.check(self.lib.ld_status_all(self.ctx, lm, bl))
        return np.array(lm[:]), np.array(bl[:], dtype=bool)

    # ------------------------------------------------------------ diagnostics (save(), plot.py)
    def spectrum(self, fk: torch.Tensor, weight: float = 0.5) -> torch.Tensor:
        """e3d_to_e1d(weight * sum_c |f_c|^2 * normalize): E(k) for u_hat with weight 0.5
        (forced-dns-sm.py:423, 431), the n spectrum with weight 1 (plot.py:37-38)."""
        nf = 1 if fk.dim() == 3 else fk.shape[0]
        self._chk(fk, self.cdtype, self.cshape(None if fk.dim() == 3 else nf), "fk")
        out = torch.zeros(self.Nf, dtype=torch.float64, device=self.device)
        self._stream()
        _lib.check(self.lib.ld_spectrum(self.ctx, _ptr(fk), nf, float(weight), _ptr(out)))
        return out

    def flux(self, uk: torch.Tensor, n: torch.Tensor) -> torch.Tensor:
        """Pi(k) of save() (forced-dns-sm.py:425-428): full_RHS with visc = forc = 0
        (no droplets), Re(conj(uk) . ku) dealiased, shell-binned, cumulated from high k."""
        ku, _, _, _ = self.full_RHS(0.0, uk, n, None, None, visc=0, forc=0)
        out = torch.zeros(self.Nf, dtype=torch.float64, device=self.device)
        self._stream()
        _lib.check(self.lib.ld_flux(self.ctx, _ptr(uk), _ptr(ku), _ptr(out)))
        return out

    def field_stats(self, x: torch.Tensor):
        """(sum, min, max, sum of squares) of one or several real fields."""
        nf = 1 if x.dim() == 3 else x.shape[0]
        self._chk(x, self.rdtype, self.rshape(None if x.dim() == 3 else nf), "x")
        out = torch.zeros(4, dtype=torch.float64, device=self.device)
        self._stream()
        _lib.check(self.lib.ld_field_stats(self.ctx, _ptr(x), nf, _ptr(out)))
        return tuple(out.cpu().tolist())

    def dissipation(self, uk: torch.Tensor) -> torch.Tensor:
        """Dissipation spectrum nu |k|^2 sum_c |u_c|^2 normalize by shell (N/2+1,);
        its sum is the dissipation rate 2 nu sum_k k^2 E(k)."""
        self._chk(uk, self.cdtype, self.cshape(3), "uk")
        out = torch.zeros(self.Nf, dtype=torch.float64, device=self.device)
        self._stream()
        _lib.check(self.lib.ld_dissipation(self.ctx, _ptr(uk), _ptr(out)))
        return out

    def divergence(self, uk: torch.Tensor, div: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Physical divergence diff_x(u) + diff_y(v) + diff_z(w) (forced-dns-sm.py:216-229)."""
        self._chk(uk, self.cdtype, self.cshape(3), "uk")
        div = self.zeros_real() if div is None else self._chk(div, self.rdtype, self.rshape(), "div")
        self._stream()
        _lib.check(self.lib.ld_divergence(self.ctx, _ptr(uk), _ptr(div)))
        return div

    def divergence_max(self, uk: torch.Tensor) -> float:
        """max |div u| as save() prints it (forced-dns-sm.py:472)."""
        _, mn, mx, _ = self.field_stats(self.divergence(uk))
        return max(abs(mn), abs(mx))

    def droplet_stats(self):
        """(sum, min, max, sum of squares) of the resident droplets' mass m/M0."""
        out = torch.zeros(4, dtype=torch.float64, device=self.device)
        self._stream()
        _lib.check(self.lib.ld_droplet_stats(self.ctx, _ptr(out)))
        return tuple(out.cpu().tolist())

    def advance_plain(self, h, nsteps: int = 1):
        """nsteps with the reference's unfused dataflow (ld_advance_plain)."""
        self._stream()
        _lib.check(self.lib.ld_advance_plain(self.ctx, float(h), int(nsteps)))

    def profile(self, kernel: str, enable: bool = True):
        """Time every launch of one kernel class with HIP events (see ld_profile)."""
        _lib.check(self.lib.ld_profile(self.ctx, _lib.KERNELS[kernel], 1 if enable else 0))

    def profile_read(self):
        """(total_ms, launches) since profile() was enabled."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _lib.check(self.lib.ld_profile_read(self.ctx, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def status(self):
        m = ctypes.c_double()
        b = ctypes.c_int32()
        _lib.check(self.lib.ld_status(self.ctx, ctypes.byref(m), ctypes.byref(b)))
        return m.value, bool(b.value)
