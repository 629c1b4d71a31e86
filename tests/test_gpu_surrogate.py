"""Surrogate network on the GPU: every kernel against the CPU oracle
(oracle/nn_oracle.py) on BF16-rounded inputs, the whole network's forward
and weight gradients, and what a trained network gains as a warm start."""
import os
import sys

import numpy as np
import pytest

from gpu_helpers import log

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', 'oracle'))
import nn_oracle as NN  # noqa: E402

pytestmark = pytest.mark.gpu


def bf(a):
    import torch
    return torch.as_tensor(np.asarray(a, dtype=np.float32)).to('cuda').to(torch.bfloat16).contiguous()


def host(t):
    import torch
    return t.detach().to(torch.float32).cpu().numpy().astype(np.float64)


def L():
    from adaptivegriddiscretizations_amd import _lib
    return _lib


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


def test_conv_forward_and_gradients():
    import torch
    lib = L()
    rng = np.random.default_rng(0)
    x = bf(rng.normal(size=(2, 8, 32, 32)))
    w = bf(rng.normal(scale=0.2, size=(9, 64, 32)))
    y = torch.empty((2, 8, 32, 64), dtype=torch.bfloat16, device='cuda')
    lib.check(lib.lib.agd_nn_conv3x3(x.data_ptr(), w.data_ptr(), y.data_ptr(), 2, 8, 32, 32, 64, 1, None))
    ref = NN.conv3x3(host(x), host(w), relu=True)
    e = relerr(host(y), ref)
    log(f"conv forward: rel err {e:.2e}")
    assert e < 1e-2
    gy = bf(rng.normal(size=(2, 8, 32, 64)))
    dw = torch.zeros((9, 64, 32), dtype=torch.float32, device='cuda')
    lib.check(lib.lib.agd_nn_conv3x3_wgrad(x.data_ptr(), gy.data_ptr(), dw.data_ptr(), 2, 8, 32, 32, 64, None))
    e = relerr(host(dw), NN.conv3x3_wgrad(host(x), host(gy)))
    log(f"conv weight gradient: rel err {e:.2e}")
    assert e < 1e-4
    wt = torch.empty((9, 32, 64), dtype=torch.bfloat16, device='cuda')
    lib.check(lib.lib.agd_nn_conv3x3_flip(w.data_ptr(), wt.data_ptr(), 32, 64, None))
    gx = torch.empty_like(x)
    lib.check(lib.lib.agd_nn_conv3x3(gy.data_ptr(), wt.data_ptr(), gx.data_ptr(), 2, 8, 32, 64, 32, 0, None))
    e = relerr(host(gx), NN.conv3x3_dgrad(host(gy), host(w)))
    log(f"conv data gradient: rel err {e:.2e}")
    assert e < 1e-2


def test_pool_up_loss_adam():
    import torch
    lib = L()
    rng = np.random.default_rng(1)
    x = bf(rng.normal(size=(2, 16, 32, 32)))
    y = torch.empty((2, 8, 16, 32), dtype=torch.bfloat16, device='cuda')
    lib.check(lib.lib.agd_nn_pool2(x.data_ptr(), y.data_ptr(), 2, 8, 16, 32, 0, None))
    assert relerr(host(y), NN.pool2(host(x))) < 1e-2
    g = torch.empty_like(x)
    lib.check(lib.lib.agd_nn_pool2(y.data_ptr(), g.data_ptr(), 2, 8, 16, 32, 1, None))
    assert relerr(host(g), NN.pool2_bwd(host(y))) < 1e-2
    u = torch.empty_like(x)
    lib.check(lib.lib.agd_nn_up2(y.data_ptr(), u.data_ptr(), 2, 8, 16, 32, 0, None))
    assert np.array_equal(host(u), NN.up2(host(y)))
    gd = torch.empty_like(y)
    lib.check(lib.lib.agd_nn_up2(x.data_ptr(), gd.data_ptr(), 2, 8, 16, 32, 1, None))
    assert relerr(host(gd), NN.up2_bwd(host(x))) < 1e-2
    t = torch.as_tensor(rng.normal(size=(2, 16, 32)).astype(np.float32), device='cuda')
    gp = torch.empty_like(x)
    loss = torch.zeros(1, dtype=torch.float32, device='cuda')
    lib.check(lib.lib.agd_nn_mse(x.data_ptr(), t.data_ptr(), None, gp.data_ptr(), 2 * 16 * 32, 32, 1.0,
                                 float(2 * 16 * 32), loss.data_ptr(), None))
    lref, gref = NN.mse(host(x), host(t))
    assert float(loss.item()) == pytest.approx(lref, rel=1e-4)
    assert relerr(host(gp), gref) < 1e-2
    w = torch.as_tensor(rng.normal(size=1000).astype(np.float32), device='cuda')
    gw = torch.as_tensor(rng.normal(size=1000).astype(np.float32), device='cuda')
    m, v = torch.zeros_like(w), torch.zeros_like(w)
    wb = torch.empty(1000, dtype=torch.bfloat16, device='cuda')
    w0 = host(w)
    lib.check(lib.lib.agd_nn_adam(w.data_ptr(), gw.data_ptr(), m.data_ptr(), v.data_ptr(), wb.data_ptr(),
                                  1000, 1e-2, 0.9, 0.999, 1e-8, 1, None))
    wr, _, _ = NN.adam(w0, host(gw), 0., 0., 1, 1e-2)
    assert relerr(host(w), wr) < 1e-5


def test_network_forward_and_weight_gradients():
    """Surrogate.forward / backward against the float64 restatement of the
    whole encoder-decoder on the same BF16 weights and input."""
    import torch
    from adaptivegriddiscretizations_amd.Eikonal import surrogate as S
    rng = np.random.default_rng(2)
    net = S.Surrogate(seed=3)
    x = bf(np.concatenate([rng.normal(size=(1, 128, 128, S.NFEAT)),
                           np.zeros((1, 128, 128, S.CP - S.NFEAT))], axis=-1))
    t = torch.as_tensor(rng.normal(size=(1, 128, 128)).astype(np.float32), device='cuda')
    cache = {}
    pred = net.forward(x, cache=cache)
    _, gy = net.loss_and_grad(pred, t)
    net.backward(gy, cache)
    torch.cuda.synchronize()
    layers = [host(l.wb) for l in net.layers]
    ref, c = NN.network_forward(host(x), layers)
    e = relerr(host(pred)[..., 0], ref[..., 0])
    log(f"network forward: rel err {e:.2e}")
    assert e < 5e-2
    _, gref = NN.mse(ref, host(t))
    grads = NN.network_backward(gref, layers, c)
    for k in (0, 3, 7, 10, 14):
        g = host(net.layers[k].g)
        cos = np.sum(g * grads[k]) / (np.linalg.norm(g) * np.linalg.norm(grads[k]))
        e = np.linalg.norm(g - grads[k]) / np.linalg.norm(grads[k])
        log(f"layer {k}: weight gradient rel err {e:.2e} (cos {cos:.5f})")
        assert e < 0.1 and cos > 0.995


def test_trained_network_gain_and_solver_judged_warm_start():
    """The learned warm start on the 1024^2 Riemann2 workload (fields as
    riemann2_aniso, one seed each), at a training size that fits a test:
    1024 fields labelled by the solver, 12 epochs, network on the 256^2
    coarse grid.  (i) On held-out fields the trained network predicts
    r = u / d_s better than the method without it (r = 1: the distance in
    the metric frozen at the seed): relative L2 error below 0.8 of it.
    (ii) Every warm start is checked against the scheme before solving
    (agd_solver_repair, one pass): a supersolution start is solved warm,
    any other start cold.  The solve then ends on the cold solution (same
    fixed point, 2e-5 of max u) and costs, per field, no more passes than
    the cold solve plus the one check pass."""
    import torch
    from adaptivegriddiscretizations_amd.Eikonal import surrogate as S
    n, atol, rtol = 1024, 1e-6, 1e-6
    feats, targ = S.build_dataset(1024, n=n, seed=1, chunk=256, atol=atol, rtol=rtol, log=log)
    net = S.Surrogate(seed=0)
    hist = net.train(feats, targ, epochs=12, batch=16, lr=2e-3, log=log)
    assert np.isfinite(hist).all() and hist[-1] < 0.5 * hist[0]
    del feats, targ
    ftest, ttest = S.build_dataset(16, n=n, seed=7, chunk=16, atol=atol, rtol=rtol)
    pred = net.predict(ftest)[..., 0].float()
    e_net = float(torch.linalg.norm(pred - ttest) / torch.linalg.norm(ttest))
    e_base = float(torch.linalg.norm(torch.ones_like(ttest) - ttest) / torch.linalg.norm(ttest))
    log(f"held-out relative L2 error of r: network {e_net:.3f}, frozen-metric distance {e_base:.3f}")
    assert e_net < 0.8 * e_base
    metric, seeds = S.riemann_fields(8, n, seed=999)
    cold, ms_c, p_c = S.solve_batched(metric, seeds, atol, rtol, batch=1)
    cold_passes = list(S.solve_batched.last_batch_passes)
    warm, rep = S.solve_warm(net, metric, seeds, atol, rtol, batch=1)
    fin = torch.isfinite(cold)          # cells no stencil reaches stay +inf in both
    assert bool((torch.isfinite(warm) == fin).all())
    diff = float((warm - cold).abs()[fin].max() / cold[fin].abs().max())
    log(f"held-out solves: cold {cold_passes} passes {ms_c:.1f} ms; warm {rep['batch_passes']} passes "
        f"+ check rounds {rep['rounds']} (violations {rep['violations']}), {rep['total_ms']:.1f} ms "
        f"(prediction {rep['pred_ms']:.1f}, check {rep['repair_ms']:.1f}, solve {rep['solve_ms']:.1f}); "
        f"fallbacks {rep['fallbacks']}; "
        f"max rel diff {diff:.2e}")
    assert diff < 2e-5
    for pw, r, pc in zip(rep['batch_passes'], rep['rounds'], cold_passes):
        assert pw + r <= pc + 1, (pw, r, pc)
