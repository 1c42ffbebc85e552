"""The fallback deposit: droplets that move more than one spatial bin within
a step (a large step, fast droplets) are deposited outside the brick pass,
also in 64-bit fixed point, so the step stays bitwise reproducible and still
agrees with the oracle (particles.py:514-577, 817-837 inside
forced-dns-sm.py:335-362)."""
import numpy as np
import pytest

from helpers import by_id, inputs, make_drops_oracle, make_flow_oracle, min_image, relerr
from lucky_droplets_amd.config import Config

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-10, "f32": 1e-4}


def _t(x, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fast_droplets_take_the_fallback_and_stay_reproducible(gpu, prec):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    N, nparts, h = 32, 2000, 0.04
    cfg = Config(N=N, precision=prec, nu=0.02, nshells=2, profile="k4", n_particles=nparts,
                 st_dist=("uniform", 0.05, 1.0), gravity=True)
    uk, n, coord, pid = inputs(cfg, seed=21)
    coord[:, 3:6] *= 3.0            # up to 6 pi: several cells per stage at this step
    # the stage displacements h/2 |v| (stages 2, 3) and h |v| (stage 4) exceed one bin
    # (two cells) for most droplets
    bins = 2 * cfg.dx
    assert np.mean(h * np.abs(coord[:, 3:6]).max(axis=1) > 1.5 * bins) > 0.3
    fl = make_flow_oracle(cfg)
    dr = make_drops_oracle(cfg, coord, pid)
    uk_r, n_r = fl.step(0.0, h, dr, uk.copy(), n.copy())
    outs = []
    for run in range(2):
        d = ForcedDNS(cfg)
        t = (_t(uk, gpu, d.cdtype), _t(n, gpu, d.rdtype), _t(coord, gpu, torch.float64))
        d.set_state(*t)
        d.advance(h, 1)
        d.get_state(*t)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in t])
        d.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    ukg, ng, cg = outs[0]
    tol = TOL[prec]
    assert relerr(ukg, uk_r) < tol
    assert relerr(ng, n_r) < tol
    b = by_id(dr.coord, dr.prtclid)
    assert min_image(cg, b, cfg.L) < tol * cfg.L
    assert relerr(cg[:, 3:6], b[:, 3:6]) < tol
    assert relerr(cg[:, 6:], b[:, 6:]) < tol
