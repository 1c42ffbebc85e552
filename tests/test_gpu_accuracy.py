"""The FP32 form against the FP64 form of the engine on the 256^3 +
10^6-droplet workload over 100 steps: energy spectrum within 1e-4 relative
per shell (every shell up to N/3 that holds energy), energy, dissipation,
density statistics and droplet statistics far inside that
(forced-dns-sm.py:419-481 lists what the reference reports)."""
import numpy as np
import pytest

from lucky_droplets_amd.config import preset

pytestmark = pytest.mark.gpu


def test_fp32_form_against_fp64_over_100_steps(gpu):
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.accuracy import compare_precisions, worst
    cfg = preset("c2")
    rng = np.random.default_rng(cfg.seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    del u
    coord, _ = ic.initial_droplets(cfg, rng)
    r = compare_precisions(cfg, uk, n, coord, dt, 100)
    print({k: v for k, v in r.items() if not k.startswith("ek_")})
    assert r["shells_compared"] >= 40
    assert r["spectrum_err_max"] < 1e-4
    assert r["spectrum_err_l1"] < 1e-5
    assert r["energy_err"] < 1e-5 and r["dissipation_err"] < 1e-5
    assert r["n_mean_err"] < 1e-6 and r["n_var_err"] < 1e-5 and r["n_max_err"] < 1e-4
    assert r["droplet_vel_err_rms"] < 1e-4 and r["droplet_mass_total_err"] < 1e-6
    assert r["droplet_pos_err_max_abs"] < 1e-3 * cfg.dx * 10
    assert worst(r) < 1e-4
