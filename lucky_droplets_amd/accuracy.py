"""Accuracy of the FP32 form against the FP64 form of the same engine.

Both forms start from the same state and run the same steps; the report
compares what the reference's save() prints and writes (forced-dns-sm.py:
419-481): the energy spectrum E(k), the energy and the dissipation rate, the
small-droplet density statistics, and the droplet population (positions,
velocities, masses, compared row by row).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .config import Config
from .dns import ForcedDNS


def _state_summary(d: ForcedDNS, steps_done: int):
    uk, n = d.zeros_spec(3), d.zeros_real()
    coord = torch.empty((d.np, 7), dtype=torch.float64, device=d.device) if d.np else None
    d.get_state(uk, n, coord)
    ek = d.spectrum(uk, 0.5).cpu().numpy()
    dk = d.dissipation(uk).cpu().numpy()
    s, mn, mx, q = d.field_stats(n)
    out = {"ek": ek, "dk": dk, "n_sum": s, "n_min": mn, "n_max": mx, "n_sq": q, "coord": coord,
           "steps": steps_done}
    return out


def compare_precisions(cfg: Config, uk: np.ndarray, n: np.ndarray, coord: np.ndarray, dt: float, steps: int,
                       device="cuda"):
    """Run `steps` steps of cfg in FP32 and in FP64 from the same inputs and
    compare.  Returns a dict of scalar errors (relative unless named otherwise)
    and the two spectra."""
    res = {}
    for prec in ("f64", "f32"):
        c = cfg.replace(precision=prec)
        d = ForcedDNS(c, device=device)
        d.set_state(torch.from_numpy(uk).to(device).to(d.cdtype), torch.from_numpy(n).to(device).to(d.rdtype),
                    torch.from_numpy(coord).to(device) if coord is not None and len(coord) else None)
        d.advance(dt, steps)
        res[prec] = _state_summary(d, steps)
        d.close()
    a, b = res["f32"], res["f64"]
    kc = cfg.N // 3
    shells = np.arange(1, kc + 1)
    e32, e64 = a["ek"][shells], b["ek"][shells]
    d32, d64 = a["dk"][shells], b["dk"][shells]
    carry = e64 > 1e-12 * e64.sum()             # shells that hold energy at this time
    rel = np.abs(e32 - e64) / np.where(e64 > 0, e64, 1.0)
    out = {
        "steps": steps,
        "spectrum_err_max": float(rel[carry].max()) if carry.any() else 0.0,
        "spectrum_err_max_shell": int(shells[carry][np.argmax(rel[carry])]) if carry.any() else 0,
        "spectrum_err_l1": float(np.abs(e32 - e64).sum() / e64.sum()),
        "energy": float(e64.sum() + b["ek"][0]),
        "energy_err": float(abs(a["ek"].sum() - b["ek"].sum()) / b["ek"].sum()),
        "dissipation": float(b["dk"].sum()),
        "dissipation_err": float(abs(a["dk"].sum() - b["dk"].sum()) / b["dk"].sum()),
        "dissipation_spectrum_err_l1": float(np.abs(d32 - d64).sum() / d64.sum()),
        "n_mean_err": float(abs(a["n_sum"] - b["n_sum"]) / abs(b["n_sum"])),
        "n_min": (a["n_min"], b["n_min"]),
        "n_max_err": float(abs(a["n_max"] - b["n_max"]) / abs(b["n_max"])),
        "n_var_err": float(abs(a["n_sq"] - b["n_sq"]) / abs(b["n_sq"])),
        "shells_compared": int(carry.sum()),
    }
    if a["coord"] is not None:
        ca, cb = a["coord"], b["coord"]
        dpos = ca[:, :3] - cb[:, :3]
        dpos = dpos - cfg.L * torch.round(dpos / cfg.L)
        va, vb = ca[:, 3:6], cb[:, 3:6]
        ma, mb = ca[:, 6], cb[:, 6]
        out.update({
            "droplet_pos_err_max_abs": float(dpos.abs().max()),
            "droplet_pos_err_rms_over_dx": float(dpos.pow(2).mean().sqrt()) / cfg.dx,
            "droplet_vel_err_rms": float((va - vb).pow(2).mean().sqrt() / vb.pow(2).mean().sqrt()),
            "droplet_speed_mean_err": float(abs(va.norm(dim=1).mean() - vb.norm(dim=1).mean())
                                            / vb.norm(dim=1).mean()),
            "droplet_mass_total_err": float(abs(ma.sum() - mb.sum()) / mb.sum()),
            "droplet_mass_err_max": float(((ma - mb).abs() / mb.abs()).max()),
        })
    out["ek_f64"] = b["ek"].tolist()
    out["ek_f32"] = a["ek"].tolist()
    return out


def worst(report: dict) -> float:
    """The largest of the relative errors in a report (spectrum, energy,
    dissipation, density and droplet statistics)."""
    keys = ["spectrum_err_max", "energy_err", "dissipation_err", "n_mean_err", "n_max_err", "n_var_err",
            "droplet_vel_err_rms", "droplet_speed_mean_err", "droplet_mass_total_err"]
    return max(report[k] for k in keys if k in report and not math.isnan(report[k]))
