"""Golden fixture: the CPU FP64 oracle run of the 64^3 workload for 200 steps.

Workload: N = 64, FP64, nu = 0.01, forced shells |k| <= 2, E(k) ~ k^4 exp(-2 (k/3)^2)
(seed 0), dt from CFL 0.5, 10^4 droplets, St ~ U[0.05, 1], gravity on (settling).
Saves the inputs' seed and a fixed subset of the final state (3000 spectral
modes per component, 3000 density points, 1000 droplets by id, E(k)), so the
GPU run can be compared at 1e-10 without rerunning the oracle.

Run:  python tests/golden/make_c1_golden.py  (about two minutes on 8 cores)
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

STEPS = int(os.environ.get("C1_STEPS", "200"))


def main():
    oflow.WORKERS = os.cpu_count() or 1
    cfg = preset("c1")
    rng = np.random.default_rng(cfg.seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, pid = ic.initial_droplets(cfg, rng)
    fl = make_flow_oracle(cfg)
    dr = make_drops_oracle(cfg, coord, pid)
    t0 = time.time()
    for i in range(STEPS):
        uk, n = fl.step(i * dt, dt, dr, uk, n)
        if i % 20 == 0:
            print(f"step {i} {time.time() - t0:.1f}s", flush=True)
    sel = np.random.default_rng(1234)
    N, Nf = cfg.N, cfg.N // 2 + 1
    mi = sel.integers(0, N, 3000), sel.integers(0, N, 3000), sel.integers(0, Nf, 3000)
    ni = sel.integers(0, N, 3000), sel.integers(0, N, 3000), sel.integers(0, N, 3000)
    ids = np.sort(sel.choice(cfg.n_particles, 1000, replace=False))
    rows = dr.by_id()[ids]
    G = fl.grid
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "c1_oracle_200.npz"),
                        steps=STEPS, dt=dt, mode_ix=np.stack(mi), uk_modes=uk[:, mi[0], mi[1], mi[2]],
                        n_ix=np.stack(ni), n_pts=n[ni], droplet_ids=ids, droplet_rows=rows,
                        ek=G.e3d_to_e1d(G.energy3d(uk)), n_mean=n.mean(), uk_absmax=np.abs(uk).max(),
                        n_absmax=np.abs(n).max())
    print("done", time.time() - t0)


if __name__ == "__main__":
    main()
