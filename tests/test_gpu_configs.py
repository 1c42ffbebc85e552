"""The batched-ensemble and learned-subgrid workloads at their own sizes.

* Ensemble: 64 independent 64^3 FP32 realizations (seeds 0-63, nu = 0.01),
  10^5 droplets each (St ~ U[0.05, 1]), stepped together with HIP-graph
  replay; realizations 0, 31 and 63 equal a context holding that realization
  alone, bit for bit.
* Learned subgrid CNN at 128^3 (channels 3-32-32-32-3): the forward pass at
  sampled nodes of a random field against the oracle, and the weight and bias
  gradients of a localised input against the oracle on a 32^3 box that holds
  all of it (outside the box every activation and gradient is exactly zero,
  so the 128^3 loss is the 32^3 loss over 64).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(x, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


def test_c3_ensemble_64x64_matches_single_runs(gpu):
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    from lucky_droplets_amd.dns import ForcedDNS
    cfg = preset("c3")
    B, steps = 64, 3
    uks, ns, cs, dts = [], [], [], []
    for b in range(B):
        rng = np.random.default_rng(cfg.seed + b)
        uk, n, u = ic.initial_flow(cfg, rng)
        dts.append(ic.cfl_dt(cfg, u))
        coord, _ = ic.initial_droplets(cfg, rng)
        uks.append(uk.astype(np.complex64)); ns.append(n.astype(np.float32)); cs.append(coord)
    dt = min(dts)
    e = ForcedDNS(cfg, batch=B, graph=True)
    uke = _t(np.stack(uks), gpu, e.cdtype)
    ne = _t(np.stack(ns), gpu, e.rdtype)
    ce = _t(np.stack(cs), gpu, torch.float64)
    e.set_state(uke, ne, ce)
    e.advance(dt, steps)
    e.get_state(uke, ne, ce)
    lm, bl = e.status_all()
    assert lm.shape == (B,) and not bl.any() and np.all(lm > 0)
    e.close()
    del e
    for b in (0, 31, 63):
        d = ForcedDNS(cfg)
        ukg, ng, cg = _t(uks[b], gpu, d.cdtype), _t(ns[b], gpu, d.rdtype), _t(cs[b], gpu, torch.float64)
        d.set_state(ukg, ng, cg)
        d.advance(dt, steps)
        d.get_state(ukg, ng, cg)
        torch.cuda.synchronize()
        for a, ref in ((uke[b], ukg), (ne[b], ng), (ce[b], cg)):
            assert np.array_equal(a.cpu().numpy().view(np.uint8), ref.cpu().numpy().view(np.uint8)), b
        d.close()


def _layers(ch, seed, bias=True):
    rng = np.random.default_rng(seed)
    return [(rng.normal(0, np.sqrt(2.0 / (27 * ci)), (27, co, ci)).astype(np.float32),
             (rng.normal(0, 0.1, co) if bias else np.zeros(co)).astype(np.float32))
            for ci, co in zip(ch[:-1], ch[1:])]


def test_c5_cnn_128_forward_at_sampled_nodes(gpu):
    import torch
    from lucky_droplets_amd.cnn import SubgridCNN
    from oracle import cnn as ocnn
    N, ch = 128, (3, 32, 32, 32, 3)
    layers = _layers(ch, 7)
    net = SubgridCNN(N, ch)
    net.set_params(layers)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(N, N, N, ch[0])).astype(np.float32)
    y = net.forward(torch.from_numpy(x).to(gpu)).cpu().numpy().reshape(N, N, N, ch[-1])
    nodes = np.concatenate([rng.integers(0, N, (300, 3)), [[0, 0, 0], [N - 1, N - 1, N - 1], [0, N - 1, 64]]])
    yr = ocnn.forward_at(layers, x, nodes, bf16=True)
    yg = y[nodes[:, 0], nodes[:, 1], nodes[:, 2]]
    assert np.linalg.norm(yg - yr) / np.linalg.norm(yr) < 1e-2
    net.close()


def test_c5_cnn_128_gradients_of_localised_input(gpu):
    import torch
    from lucky_droplets_amd.cnn import SubgridCNN
    from oracle import cnn as ocnn
    N, M, ch = 128, 32, (3, 32, 32, 32, 3)
    layers = _layers(ch, 9, bias=False)
    rng = np.random.default_rng(10)
    # a 12^3 box of data straddling the periodic seam in x, zero elsewhere
    xs = np.zeros((N, N, N, ch[0]), np.float32)
    ts = np.zeros((N, N, N, ch[-1]), np.float32)
    ix = (np.arange(-6, 6) % N)
    xb = rng.normal(size=(12, 12, 12, ch[0])).astype(np.float32)
    tb = rng.normal(size=(12, 12, 12, ch[-1])).astype(np.float32)
    xs[np.ix_(ix, np.arange(40, 52), np.arange(70, 82))] = xb
    ts[np.ix_(ix, np.arange(40, 52), np.arange(70, 82))] = tb
    net = SubgridCNN(N, ch)
    net.set_params(layers)
    loss, g = net.grad(torch.from_numpy(xs).to(gpu), torch.from_numpy(ts).to(gpu))
    gl = net.unpack(g.cpu().numpy())
    # the same data centred in a 32^3 periodic box
    xm = np.zeros((M, M, M, ch[0]), np.float32)
    tm = np.zeros((M, M, M, ch[-1]), np.float32)
    xm[10:22, 10:22, 10:22] = xb
    tm[10:22, 10:22, 10:22] = tb
    lr, gr = ocnn.loss_and_grads(layers, xm, tm, bf16=True)
    scale = (M / N) ** 3            # mean over N^3 nodes instead of M^3: a power of two, exact in bf16
    assert abs(loss - lr * scale) < 1e-2 * lr * scale
    for l in range(len(layers)):
        ref_w, ref_b = gr[l][0] * scale, gr[l][1] * scale
        assert np.linalg.norm(gl[l][0] - ref_w) / np.linalg.norm(ref_w) < 3e-2, l
        assert np.linalg.norm(gl[l][1] - ref_b) / max(np.linalg.norm(ref_b), 1e-30) < 3e-2, l
    net.close()
