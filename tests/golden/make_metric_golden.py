"""Golden fixture: the CPU FP64 oracle on the 256^3 + 10^6-droplet workload.

Workload: N = 256, nu = 2.5e-3, forced shells |k| <= 2,
E(k) ~ k^4 exp(-2 (k/3)^2) (seed 0), dt from CFL 0.5, 10^6 droplets with St
lognormal (median 0.3, sigma 0.6), no gravity; computed in FP64.
Saves a fixed random subset of
  * one full_RHS at the initial state: u, v, ku, kn, fc and the droplet RHS;
  * one full step (RK4 + epilogue): uk, n and the droplet rows;
together with each field's largest magnitude (the normalisation of the
relative error), so the GPU engine can be checked at this size (both
precisions) without rerunning the oracle.

Run:  python tests/golden/make_metric_golden.py  (a few minutes on 8 cores,
about 12 GB of host memory)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle.flow as oflow  # noqa: E402
from helpers import make_drops_oracle, make_flow_oracle  # noqa: E402
from lucky_droplets_amd import ic  # noqa: E402
from lucky_droplets_amd.config import preset  # noqa: E402

NS = 4096          # samples per field
NDROP = 2000       # droplets by id


def main():
    oflow.WORKERS = min(16, os.cpu_count() or 1)
    cfg = preset("c2").replace(precision="f64")
    rng = np.random.default_rng(cfg.seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    del u
    coord, pid = ic.initial_droplets(cfg, rng)
    fl = make_flow_oracle(cfg)
    N, Nf = cfg.N, cfg.N // 2 + 1
    sel = np.random.default_rng(4321)
    spec_ix = np.stack([sel.integers(0, 3, NS), sel.integers(0, N, NS), sel.integers(0, N, NS),
                        sel.integers(0, Nf, NS)])
    phys_ix = np.stack([sel.integers(0, 3, NS), sel.integers(0, N, NS), sel.integers(0, N, NS),
                        sel.integers(0, N, NS)])
    ids = np.sort(sel.choice(cfg.n_particles, NDROP, replace=False))
    out = dict(dt=dt, spec_ix=spec_ix, phys_ix=phys_ix, droplet_ids=ids)

    def spec(name, a):
        out[name] = a[tuple(spec_ix)]
        out[name + "_absmax"] = np.abs(a).max()

    def phys(name, a):
        ix = tuple(phys_ix) if a.ndim == 4 else tuple(phys_ix[1:])
        out[name] = a[ix]
        out[name + "_absmax"] = np.abs(a).max()

    t0 = time.time()
    dr = make_drops_oracle(cfg, coord, pid)
    ku, kn, kp, _ = fl.full_RHS(0.0, uk, n.copy(), coord.copy(), coord.copy(), dr)
    print(f"full_RHS {time.time() - t0:.1f}s", flush=True)
    spec("rhs_ku", ku)
    phys("rhs_kn", kn)
    phys("rhs_u", fl.last["u"])
    phys("rhs_v", fl.last["v"])
    phys("rhs_fc", fl.last["fc"])
    kp_id = dr.by_id(kp)
    out["rhs_kp"] = kp_id[ids]
    out["rhs_kp_absmax"] = np.abs(kp_id).max(axis=0)
    del ku, kn, kp, kp_id
    fl.last.clear()

    dr = make_drops_oracle(cfg, coord, pid)
    uk1, n1 = fl.step(0.0, dt, dr, uk.copy(), n.copy())
    print(f"step {time.time() - t0:.1f}s", flush=True)
    spec("step_uk", uk1)
    phys("step_n", n1)
    rows = dr.by_id()
    out["step_rows"] = rows[ids]
    out["step_n_mean"] = n1.mean()
    out["step_mass"] = rows[:, 6].sum()
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "metric_oracle_256.npz"), **out)
    print("done", time.time() - t0)


if __name__ == "__main__":
    main()
