"""Mass tracking of a resident run: the droplets' total mass, the density
field's total and their sum.

The step conserves the sum: the density RHS kn has mean -mean(fc) and the
deposit puts sum(fc) dx^3 = sum(W), the droplets' mass growth
(forced-dns-sm.py:325-330, particles.py:828-835), and clip_zero keeps the
density's mean (forced-dns-sm.py:278-285).  The record is taken on the device
(ld_field_stats, ld_droplet_stats) every `every` steps.
"""
from __future__ import annotations

import numpy as np


def mass_history(dns, dt: float, steps: int, every: int, log=None):
    """Advance `dns` (a ForcedDNS with a resident state) `steps` steps of size
    dt and return an array of rows (step, droplet mass, density mass
    sum(n) dx^3, their sum), the first row before the first step; `log`
    (optional) receives a line per row."""
    dx3 = dns.cfg.dx ** 3
    n = dns.zeros_real()
    rows = []

    def record(step):
        dns.get_state(None, n, None)
        sn = dns.field_stats(n)[0] * dx3
        sm = dns.droplet_stats()[0]
        rows.append((step, sm, sn, sm + sn))
        if log:
            log(f"[tracking] step {step}: droplet mass {sm:.9e}, density mass {sn:.9e}, sum {sm + sn:.12e}")

    t = 0
    record(t)
    while t < steps:
        k = min(every, steps - t)
        dns.advance(dt, k)
        t += k
        record(t)
    return np.array(rows)


def summary(hist: np.ndarray) -> dict:
    """Droplet mass gained, and the largest drift of the conserved sum relative
    to the mass the droplets gained and to the sum itself."""
    m0, total0 = hist[0, 1], hist[0, 3]
    gained = float(hist[-1, 1] - m0)
    drift = float(np.max(np.abs(hist[:, 3] - total0)))
    return {"steps": int(hist[-1, 0]), "droplet_mass_start": float(m0), "droplet_mass_end": float(hist[-1, 1]),
            "droplet_mass_gained": gained, "total_start": float(total0),
            "max_total_drift": drift, "max_total_drift_per_gained": drift / gained if gained else None,
            "max_total_drift_rel": drift / abs(total0) if total0 else None}
