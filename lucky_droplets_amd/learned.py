"""Learned subgrid form: a coarse DNS whose droplets see a fine-grid velocity
reconstructed by the CNN (no counterpart in the reference).

The coarse grid has Nc^3 nodes and the fine grid Nf^3 = (r Nc)^3 (r = 2).
From a coarse spectral velocity u_hat_c:

  base   = the coarse velocity on the fine grid (spectral zero padding,
           ld_spectral_resample + ld_irfft on the fine context);
  x      = the coarse physical velocity and its gradient as CNN rows (3 + 9
           channels per coarse node; ld_velocity_gradient_spec + ld_irfft);
  net(x) = 3 r^3 channels per coarse node: the subgrid velocity at the node's
           r^3 fine sub-nodes;
  u_rec  = base + net(x) on the fine grid (ld_subgrid_compose).

Training pairs come from fine DNS states: the coarse input is the fine
spectrum truncated to the coarse grid (ld_spectral_resample), the target the
fine velocity minus base (ld_subgrid_residual), i.e. the modes the coarse grid
cannot carry.  The droplets of every form are advanced one-way coupled in a
frozen velocity field per step (ld_droplets_frozen_rk4, the drag and settling
of particles.py:831-835): the fine DNS field (the reference form), the coarse field on
the coarse grid (plain interpolation), or u_rec (the learned form, which
falls back to the base alone when the trained network does worse than none
on a held-out state).  Pair statistics (ld_pair_stats) give g(r) and the kinematic collision
kernel of each form.
"""
from __future__ import annotations

import math
import time
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib, ic
from .cnn import SubgridCNN
from .config import Config
from .dns import ForcedDNS, _ptr


class SubgridModel:
    """CNN reconstruction of the fine-grid velocity from a coarse one."""

    def __init__(self, cfg_fine: Config, ratio: int = 2, hidden=(32, 32, 32), device="cuda", seed: int = 0):
        if cfg_fine.precision != "f32":
            raise ValueError("the learned form runs with FP32 fields")
        self.r = int(ratio)
        self.cf = cfg_fine
        self.Nf = cfg_fine.N
        self.Nc = cfg_fine.N // self.r
        self.cc = cfg_fine.replace(N=self.Nc)
        self.device = torch.device(device)
        self.fine = ForcedDNS(cfg_fine, max_particles=0, device=device)
        self.coarse = ForcedDNS(self.cc, max_particles=0, device=device)
        self.lib = self.fine.lib
        self.channels = (12,) + tuple(hidden) + (3 * self.r ** 3,)
        self.net = SubgridCNN(self.Nc, self.channels, device=device)
        # last layer zero: training starts from the plain reconstruction (base)
        self.net.init_params(seed=seed, gain=0.0)
        self.sx = self.sg = self.sy = 1.0
        self.fallback = False       # set by validate(): the network is not used
        self._ukf = self.fine.zeros_spec(3)
        self._base = self.fine.zeros_real(3)
        self._f12 = self.coarse.zeros_real(12)
        self._gk = self.coarse.zeros_spec(9)
        self._x = torch.empty((self.Nc ** 3, 12), dtype=torch.float32, device=self.device)

    def close(self):
        for o in (self.fine, self.coarse, self.net):
            o.close()

    # ------------------------------------------------------------ field plumbing
    def coarse_spectrum(self, uk_fine: torch.Tensor) -> torch.Tensor:
        """A fine spectral velocity truncated to the coarse grid."""
        out = self.coarse.zeros_spec(3)
        self.coarse._stream()
        _lib.check(self.lib.ld_spectral_resample(self.coarse.ctx, _ptr(uk_fine), self.Nf, _ptr(out), 3))
        return out

    def base(self, uk_c: torch.Tensor) -> torch.Tensor:
        """The coarse velocity on the fine grid (zero padding)."""
        self.fine._stream()
        _lib.check(self.lib.ld_spectral_resample(self.fine.ctx, _ptr(uk_c), self.Nc, _ptr(self._ukf), 3))
        return self.fine.irfft_mpi(self._ukf, self._base)

    def inputs(self, uk_c: torch.Tensor) -> torch.Tensor:
        """CNN rows of the coarse physical velocity (scaled by 1/sx) and its
        gradient (scaled by 1/sg)."""
        self.coarse.irfft_mpi(uk_c, self._f12[:3])
        self.coarse._stream()
        _lib.check(self.lib.ld_velocity_gradient_spec(self.coarse.ctx, _ptr(uk_c), _ptr(self._gk), self.sx / self.sg))
        self.coarse.irfft_mpi(self._gk, self._f12[3:])
        self.coarse._stream()
        _lib.check(self.lib.ld_fields_to_rows(self.coarse.ctx, _ptr(self._f12), 12, _ptr(self._x), 1.0 / self.sx))
        return self._x

    def pair(self, uk_fine: torch.Tensor):
        """(CNN input rows, scaled subgrid target rows) of one fine state."""
        uk_c = self.coarse_spectrum(uk_fine)
        x = self.inputs(uk_c).clone()
        base = self.base(uk_c)
        uf = self.fine.irfft_mpi(uk_fine, self.fine.zeros_real(3))
        t = torch.empty((self.Nc ** 3, 3 * self.r ** 3), dtype=torch.float32, device=self.device)
        self.fine._stream()
        _lib.check(self.lib.ld_subgrid_residual(self.fine.ctx, _ptr(uf), _ptr(base), self.r, _ptr(t), 1.0 / self.sy))
        return x, t

    def reconstruct(self, uk_c: torch.Tensor, out: Optional[torch.Tensor] = None, learned: bool = True):
        """u_rec = base + sy net(x) on the fine grid (learned=False, or a model
        whose network did worse than none on held-out data: base only)."""
        out = self.fine.zeros_real(3) if out is None else out
        base = self.base(uk_c)
        if not learned or self.fallback:
            out.copy_(base)
            return out
        y = self.net.forward(self.inputs(uk_c))
        self.fine._stream()
        _lib.check(self.lib.ld_subgrid_compose(self.fine.ctx, _ptr(y), self.r, _ptr(base), _ptr(out), self.sy))
        return out

    # ------------------------------------------------------------ training
    def fit_scales(self, uk_fines: List[torch.Tensor]):
        """sx, sg, sy = rms of the coarse velocity, of its gradient and of the
        subgrid target."""
        self.sx = self.sg = self.sy = 1.0
        xs, gs, ts = 0.0, 0.0, 0.0
        for uk in uk_fines:
            x, t = self.pair(uk)
            xs += float((x[:, :3].double() ** 2).mean())
            gs += float((x[:, 3:].double() ** 2).mean())
            ts += float((t.double() ** 2).mean())
        self.sx = math.sqrt(xs / len(uk_fines))
        self.sg = math.sqrt(gs / len(uk_fines))
        self.sy = math.sqrt(ts / len(uk_fines))

    def train(self, pairs, iters: int, lr: float = 3e-4, lr_end: float = 3e-5) -> List[float]:
        """Adam steps over the pairs in turn, the step size decaying
        geometrically from lr to lr_end.  The first Adam steps move every
        weight by about the step size; at 3e-3 that switched off every hidden
        ReLU of this network (fan-in 27 x 32) for these targets, which left it
        at the plain reconstruction."""
        losses = []
        for it in range(iters):
            x, t = pairs[it % len(pairs)]
            a = lr * (lr_end / lr) ** (it / max(1, iters - 1))
            losses.append(self.net.train_step(x, t, lr=a))
        return losses

    def validate(self, uk_fine: torch.Tensor):
        """Held-out errors (plain, learned) of one fine state; the model falls
        back to the plain reconstruction when the network's is not smaller."""
        e_plain, e_learned = self.heldout_error(uk_fine)
        self.fallback = not e_learned < e_plain
        return e_plain, e_learned

    def heldout_error(self, uk_fine: torch.Tensor):
        """Mean squared subgrid error at the fine nodes, (plain, learned): plain
        is the coarse velocity alone (zero subgrid part)."""
        x, t = self.pair(uk_fine)
        y = self.net.forward(x)
        plain = float((t.double() ** 2).mean()) * self.sy ** 2
        learned = float(((y - t).double() ** 2).mean()) * self.sy ** 2
        return plain, learned


def _spin_up(d: ForcedDNS, uk: np.ndarray, n: np.ndarray, dt: float, steps: int):
    dev = d.device
    d.set_state(torch.from_numpy(uk).to(dev).to(d.cdtype), torch.from_numpy(n).to(dev).to(d.rdtype))
    if steps:
        d.advance(dt, steps)


def _state_uk(d: ForcedDNS) -> torch.Tensor:
    uk = d.zeros_spec(3)
    d.get_state(uk, None, None)
    return uk


def sort_rows(rows: torch.Tensor, L: float, n: int) -> torch.Tensor:
    """Droplet rows reordered by cell of an n^3 grid (x-major), so that the
    gathers of neighbouring threads share cache lines; the statistics do not
    depend on the row order."""
    c = torch.remainder(rows[:, :3], L).mul(n / L).long().clamp_(0, n - 1)
    key = (c[:, 0] * n + c[:, 1]) * n + c[:, 2]
    return rows[torch.argsort(key)].contiguous()


def pair_stats(d: ForcedDNS, rows: torch.Tensor, L: float, rmax: float, nbins: int):
    counts = torch.empty(nbins, dtype=torch.int64, device=d.device)
    wsum = torch.empty(nbins, dtype=torch.float64, device=d.device)
    wneg = torch.empty(nbins, dtype=torch.float64, device=d.device)
    d._stream()
    _lib.check(d.lib.ld_pair_stats(d.ctx, _ptr(rows), rows.shape[0], float(rmax), int(nbins), _ptr(counts),
                                   _ptr(wsum), _ptr(wneg)))
    return counts.cpu().numpy(), wsum.cpu().numpy(), wneg.cpu().numpy()


def kernel_from_stats(counts, wsum, wneg, npart, L, rmax):
    """g(r), the mean |w_r| and the kinematic collision kernel K(r) =
    2 pi r^2 g(r) <|w_r|> at the bin centres (Sundaram & Collins 1997), and the
    cumulative inward flux sum_{r' < r} wneg / (pairs per volume), the rate at
    which pairs cross a sphere of radius r inward."""
    nb = len(counts)
    edges = np.arange(nb + 1) * (rmax / nb)
    centre = 0.5 * (edges[1:] + edges[:-1])
    shell = 4.0 / 3.0 * math.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    ppv = npart * (npart - 1) / 2.0 / L ** 3
    with np.errstate(invalid="ignore", divide="ignore"):
        g = counts / (ppv * shell)
        wr = np.where(counts > 0, wsum / np.maximum(counts, 1), np.nan)
    return {"r": centre, "g": g, "wr": wr, "K": 2 * math.pi * centre ** 2 * g * wr}


def collision_comparison(cfg_fine: Config, npart: int, st: float, spinup: int, train_states: int,
                         train_every: int, train_iters: int, relax: int, samples: int, sample_every: int,
                         rmax: float, nbins: int, seed_train: int = 100, seed_eval: int = 200, log=None) -> Dict:
    """Train the subgrid model on states of a fine DNS (seed_train), then run
    from one initial state (seed_eval) the fine DNS, the plain coarse DNS and
    the learned coarse DNS, each with the same droplets (Stokes number st,
    one-way coupled), and compare their pair statistics after `relax` steps,
    averaged over `samples` snapshots `sample_every` steps apart."""
    log = log or (lambda m: None)
    dev = torch.device("cuda")
    model = SubgridModel(cfg_fine)
    fine, coarse = model.fine, model.coarse
    rng = np.random.default_rng(seed_train)
    uk0, n0, u0 = ic.initial_flow(cfg_fine, rng)
    dt = ic.cfl_dt(cfg_fine, u0)
    # training states from the fine DNS
    t0 = time.time()
    _spin_up(fine, uk0, n0, dt, spinup)
    states = []
    for k in range(train_states + 1):
        states.append(_state_uk(fine))
        fine.advance(dt, train_every)
    heldout = states.pop()
    model.fit_scales(states[:4])
    pairs = [model.pair(s) for s in states]
    pairs = [(x.clone(), t.clone()) for x, t in pairs]
    losses = model.train(pairs, train_iters)
    e_plain, e_learned = model.validate(heldout)
    log(f"[learned] trained on {len(states)} states, {train_iters} steps in {time.time() - t0:.1f}s; "
        f"held-out subgrid mse plain {e_plain:.3e} learned {e_learned:.3e}")
    del pairs, states
    # evaluation: one initial state, three forms
    rng = np.random.default_rng(seed_eval)
    uk0, n0, u0 = ic.initial_flow(cfg_fine, rng)
    _spin_up(fine, uk0, n0, dt, spinup)
    ukf = _state_uk(fine)
    ukc = model.coarse_spectrum(ukf)
    coarse.set_state(ukc, coarse.zeros_real().fill_(1.0))
    drng = np.random.default_rng(seed_eval + 1)
    rows0 = drng.uniform(0, cfg_fine.L, (npart, 7))
    rows0[:, 3:6] = 0.0
    rows0[:, 6] = st ** 1.5 * cfg_fine.factor
    forms = {k: torch.from_numpy(rows0).to(dev) for k in ("fine", "plain", "learned")}
    uf, uc, ur = fine.zeros_real(3), coarse.zeros_real(3), fine.zeros_real(3)
    stats = {k: [] for k in forms}
    times = {k: 0.0 for k in forms}
    nsteps = relax + (samples - 1) * sample_every + 1
    for it in range(nsteps):
        if it % 50 == 0:
            for k in forms:
                forms[k] = sort_rows(forms[k], cfg_fine.L, 64)
        # field at the step start, droplets one RK4 step in it, then the flow steps
        torch.cuda.synchronize()
        a = time.time()
        fine.get_state(ukf, None, None)
        fine.irfft_mpi(ukf, uf)
        _lib.check(fine.lib.ld_droplets_frozen_rk4(fine.ctx, _ptr(forms["fine"]), npart, _ptr(uf), dt, 1))
        fine.advance(dt, 1)
        torch.cuda.synchronize()
        b = time.time()
        coarse.get_state(ukc, None, None)
        coarse.irfft_mpi(ukc, uc)
        _lib.check(coarse.lib.ld_droplets_frozen_rk4(coarse.ctx, _ptr(forms["plain"]), npart, _ptr(uc), dt, 1))
        coarse.advance(dt, 1)
        torch.cuda.synchronize()
        c = time.time()
        model.reconstruct(ukc, ur)
        _lib.check(fine.lib.ld_droplets_frozen_rk4(fine.ctx, _ptr(forms["learned"]), npart, _ptr(ur), dt, 1))
        torch.cuda.synchronize()
        e = time.time()
        times["fine"] += b - a
        times["plain"] += c - b
        times["learned"] += (c - b) + (e - c)
        if it >= relax and (it - relax) % sample_every == 0:
            for k, rws in forms.items():
                stats[k].append(pair_stats(fine, rws, cfg_fine.L, rmax, nbins))
    out = {"dt": dt, "steps": nsteps, "heldout_mse_plain": e_plain, "heldout_mse_learned": e_learned,
           "network_used": not model.fallback,
           "train_loss_first": float(np.mean(losses[:10])), "train_loss_last": float(np.mean(losses[-10:])),
           "ms_per_step": {k: 1e3 * v / nsteps for k, v in times.items()}}
    res = {}
    for k, lst in stats.items():
        c = sum(s[0] for s in lst).astype(np.float64)
        w = sum(s[1] for s in lst)
        wn = sum(s[2] for s in lst)
        res[k] = kernel_from_stats(c / len(lst), w / len(lst), wn / len(lst), npart, cfg_fine.L, rmax)
        res[k]["counts"] = c
        res[k]["wsum"] = w
    out["forms"] = res
    # collision kernel over the separations where every bin has pairs: K summed over the bins
    ok = np.all([res[k]["counts"] > 0 for k in res], axis=0)
    kf = float(np.sum(res["fine"]["K"][ok]))
    for k in ("plain", "learned"):
        out[f"kernel_err_{k}"] = float(abs(np.sum(res[k]["K"][ok]) - kf) / kf)
        out[f"g_err_{k}"] = float(np.nanmean(np.abs(res[k]["g"][ok] - res["fine"]["g"][ok]) / res["fine"]["g"][ok]))
    out["throughput_ratio_learned_vs_fine"] = out["ms_per_step"]["fine"] / out["ms_per_step"]["learned"]
    model.close()
    return out
