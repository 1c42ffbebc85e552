"""ld_spectrum_noalloc (the allocation-free variant of ld_spectrum) against
ld_spectrum through the C-ABI: bit-equal results on two calls in a row, the
second on a smaller input (another context, smaller grid and fewer fields),
plus its argument errors."""
import ctypes

import numpy as np
import pytest

from helpers import inputs
from lucky_droplets_amd import _lib
from lucky_droplets_amd.config import Config

pytestmark = pytest.mark.gpu


def _spectra(gpu, N, prec, nfields, seed):
    import torch
    from lucky_droplets_amd.dns import ForcedDNS, _ptr
    cfg = Config(N=N, precision=prec, nu=0.02, nshells=2, profile="k4")
    uk, _, _, _ = inputs(cfg, seed=seed)
    d = ForcedDNS(cfg, max_particles=0)
    f = torch.from_numpy(np.ascontiguousarray(uk[:nfields])).to(gpu).to(d.cdtype).contiguous()
    a = torch.full((N // 2 + 1,), -1.0, dtype=torch.float64, device=gpu)
    b = torch.full((N // 2 + 1,), -2.0, dtype=torch.float64, device=gpu)
    d._stream()
    _lib.check(d.lib.ld_spectrum(d.ctx, _ptr(f), nfields, 0.5, _ptr(a)))
    _lib.check(d.lib.ld_spectrum_noalloc(d.ctx, _ptr(f), nfields, 0.5, _ptr(b)))
    torch.cuda.synchronize()
    # argument errors are reported, not run
    assert d.lib.ld_spectrum_noalloc(d.ctx, _ptr(f), 0, 0.5, _ptr(b)) == _lib.LD_ERR_ARG
    assert d.lib.ld_spectrum_noalloc(d.ctx, _ptr(f), 1, 0.5, None) == _lib.LD_ERR_ARG
    assert d.lib.ld_spectrum_noalloc(None, _ptr(f), 1, 0.5, _ptr(b)) == _lib.LD_ERR_ARG
    out = (a.cpu().numpy(), b.cpu().numpy())
    d.close()
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_spectrum_noalloc_equals_spectrum(gpu, prec):
    a, b = _spectra(gpu, 128, prec, 3, seed=1)
    assert np.array_equal(a, b)
    assert a[1] > 0 and np.all(a >= 0)
    # second call: smaller grid, one field; the shared scratch keeps the first call's
    # larger contents beyond what this call writes, which must not leak into it
    a2, b2 = _spectra(gpu, 32, prec, 1, seed=2)
    assert np.array_equal(a2, b2)
    assert a2.shape == (17,)
