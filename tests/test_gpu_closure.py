"""Learned flux-correction closure on the GPU.

The HIP convolutions, the weight gradients, the loss adjoint and a full
forward + backward pass are checked against the CPU oracle (torch CPU in
float64, oracle/cnn.py) on the same BF16-rounded weights and inputs.
Tolerances: FP32 outputs 1e-4 relative to the output scale (FP32
accumulation of exact BF16 products); BF16 outputs 1/128 relative (two BF16
roundings); parameter gradients through the BF16 activations 3e-2 relative
in norm.

The closure path is checked against the plain path: with an all-zero network
(and with no network) it is bit for bit the plain SSPRK3 run; with outputs
forced huge from the host (out_scale 1e8) every stage falls back to the plain
scheme and the run again equals the plain one, with every stage counted.
"""

import math

import numpy as np
import pytest
import torch

from oracle import cnn as ocnn

pytestmark = pytest.mark.gpu


def _bf16(t):
    """Round FP32 to BF16 (nearest even) and back, on the CPU."""
    u = t.float().contiguous().view(torch.int32)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.view(torch.float32)


def _to_dev_bf16(t):
    return (_bf16(t).view(torch.int32) >> 16).to(torch.int16).cuda()


def _from_dev_bf16(t):
    return (t.cpu().to(torch.int32) << 16).view(torch.float32)


def _net(seed=0, scale=1.0):
    from mfvm.closure import FluxCorrectionCNN
    net = FluxCorrectionCNN(seed=seed, last_scale=1.0)
    g = torch.Generator().manual_seed(seed + 7)
    for b in net.b:
        b.copy_((torch.randn(b.shape, generator=g) * 0.1).cuda())
    for w in net.w:
        w.mul_(scale)
    net.pack()
    return net


def _params64(net):
    out = []
    for i in range(3):
        out.append(_bf16(net.w[i].cpu()).double())
        out.append(net.b[i].cpu().double())
    return out


def test_conv_forward_matches_oracle():
    net = _net()
    B, n2, n1 = 2, 6, 32
    g = torch.Generator().manual_seed(3)
    x = torch.randn((B, n2, n1, 4), generator=g)
    xd = _to_dev_bf16(x)
    y, a1, a2 = net.forward(xd)
    torch.cuda.synchronize()
    p = _params64(net)
    xr = _bf16(x).double()
    ref_a1 = torch.relu(ocnn.conv_periodic(xr.permute(0, 3, 1, 2), p[0], p[1])).permute(0, 2, 3, 1)
    got_a1 = _from_dev_bf16(a1).double()
    assert torch.allclose(got_a1, ref_a1, rtol=1 / 128, atol=1e-3 * float(ref_a1.abs().max()))
    # layer 2 and 3 from the GPU's own (BF16) inputs: isolates each layer
    a1r = got_a1
    ref_a2 = torch.relu(ocnn.conv_periodic(a1r.permute(0, 3, 1, 2), p[2], p[3])).permute(0, 2, 3, 1)
    got_a2 = _from_dev_bf16(a2).double()
    assert torch.allclose(got_a2, ref_a2, rtol=1 / 128, atol=1e-3 * float(ref_a2.abs().max()))
    ref_y = ocnn.conv_periodic(got_a2.permute(0, 3, 1, 2), p[4], p[5]).permute(0, 2, 3, 1)
    got_y = y.cpu().double()
    scale = float(ref_y.abs().max())
    assert float((got_y - ref_y).abs().max()) <= 1e-4 * scale


def test_backward_data_and_weight_gradients_match_oracle():
    net = _net(seed=1)
    B, n2, n1 = 2, 5, 16
    g = torch.Generator().manual_seed(4)
    x = _bf16(torch.randn((B, n2, n1, 4), generator=g))
    xd = _to_dev_bf16(x)
    y, a1, a2 = net.forward(xd)
    gy = _bf16(torch.randn((B, n2, n1, 8), generator=g))
    net.zero_grad()
    net.backward(xd, a1, a2, _to_dev_bf16(gy))
    torch.cuda.synchronize()
    # CPU: layer by layer on the GPU's own BF16 activations
    p = [t.requires_grad_() for t in _params64(net)]
    a1r = _from_dev_bf16(a1).double()
    a2r = _from_dev_bf16(a2).double().requires_grad_()
    a1q = a1r.clone().requires_grad_()
    y3 = ocnn.conv_periodic(a2r.permute(0, 3, 1, 2), p[4], p[5]).permute(0, 2, 3, 1)
    gw3, gb3, ga2 = torch.autograd.grad(y3, [p[4], p[5], a2r], gy.double())
    for got, ref in ((net.gw[2], gw3), (net.gb[2], gb3)):
        assert torch.allclose(got.cpu().double(), ref, rtol=1e-4, atol=1e-4 * float(ref.abs().max()))
    ga2m = _bf16((ga2 * (a2r > 0)).float()).double()
    z2 = ocnn.conv_periodic(a1q.permute(0, 3, 1, 2), p[2], p[3]).permute(0, 2, 3, 1)
    gw2, gb2, ga1 = torch.autograd.grad(z2, [p[2], p[3], a1q], ga2m)
    # the GPU's backward-data output (ga2) is not exposed; its effect is in gw2
    for got, ref in ((net.gw[1], gw2), (net.gb[1], gb2)):
        err = float((got.cpu().double() - ref).norm() / ref.norm())
        assert err < 1e-2, err
    ga1m = (ga1 * (a1r > 0))
    z1 = ocnn.conv_periodic(x.double().permute(0, 3, 1, 2), p[0], p[1]).permute(0, 2, 3, 1)
    gw1, gb1 = torch.autograd.grad(z1, [p[0], p[1]], ga1m)
    for got, ref in ((net.gw[0], gw1), (net.gb[0], gb1)):
        err = float((got.cpu().double() - ref).norm() / ref.norm())
        assert err < 3e-2, err


def test_loss_and_its_gradient_match_oracle():
    net = _net(seed=2)
    B, n2, n1 = 2, 8, 16
    g = torch.Generator().manual_seed(5)
    y = torch.randn((B, n2, n1, 8), generator=g)
    tgt = torch.randn((B, n2, n1, 4), generator=g)
    net.out_scale.copy_(torch.tensor([0.5, 1.0, 1.5, 2.0]))
    wf = torch.tensor([1.0, 0.5, 0.25, 2.0]).cuda()
    L, gy = net.loss(y.cuda(), tgt.cuda(), wf, 0.25, 0.5)
    torch.cuda.synchronize()
    yr = y.double().requires_grad_()
    ref = ocnn.loss(yr, tgt.double(), net.out_scale.cpu().double(), wf.cpu().double(), 0.25, 0.5)
    (gref,) = torch.autograd.grad(ref, [yr])
    assert abs(float(L) - float(ref.detach())) <= 1e-5 * float(ref.detach())
    got = _from_dev_bf16(gy).double()
    assert torch.allclose(got, gref, rtol=1 / 128, atol=1e-3 * float(gref.abs().max()))


def test_network_gradient_matches_oracle_end_to_end():
    net = _net(seed=3)
    B, n2, n1 = 2, 8, 16
    g = torch.Generator().manual_seed(6)
    x = _bf16(torch.randn((B, n2, n1, 4), generator=g))
    tgt = torch.randn((B, n2, n1, 4), generator=g)
    net.out_scale.copy_(torch.tensor([0.5, 1.0, 1.5, 2.0]))
    wf = torch.tensor([1.0, 0.5, 0.25, 2.0]).cuda()
    xd = _to_dev_bf16(x)
    y, a1, a2 = net.forward(xd)
    L, gy = net.loss(y, tgt.cuda(), wf, 0.25, 0.5)
    net.zero_grad()
    net.backward(xd, a1, a2, gy)
    torch.cuda.synchronize()
    p = [t.requires_grad_() for t in _params64(net)]
    ref = ocnn.loss(ocnn.forward(x.double(), p), tgt.double(), net.out_scale.cpu().double(),
                    wf.cpu().double(), 0.25, 0.5)
    grads = torch.autograd.grad(ref, p)
    assert abs(float(L) - float(ref)) <= 2e-2 * float(ref)
    total = float(torch.cat([gw.reshape(-1) for gw in grads]).norm())
    for got, want in zip(net.grads(), grads):
        if float(want.norm()) < 1e-9 * total:
            # the last bias shifts every face flux equally: div(dF) and hence
            # the loss do not depend on it (exact zero gradient)
            err = float(got.cpu().double().norm()) / total
            print("absolute gradient error (zero gradient)", tuple(want.shape), f"{err:.2e}", flush=True)
            assert err < 1e-3
            continue
        err = float((got.cpu().double() - want).norm() / want.norm())
        print("relative gradient error", tuple(want.shape), f"{err:.2e}", flush=True)
        assert err < 3e-2


def _kh(n, seed=0):
    from mfvm.closure import kh_initial
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": n, "x2_min": 0.0, "x2_max": 1.0, "nx2": n,
           "x3_min": 0.0, "x3_max": 1.0, "nx3": 1}
    U = ns.allocate_u(cfg)
    for k, a in zip(ns.VARS, kh_initial(n, seed)):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(a).cuda()
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.minmod, "diffusive_solver": diffusive_solvers.centered}
    bcs = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    return ns, U, build_grid("cartesian", cfg), params, schemes, bcs


@pytest.mark.parametrize("mode", ["zero", "none", "huge"])
@pytest.mark.parametrize("graph", [0, 6])
def test_closure_path_reduces_to_plain(mode, graph):
    from mfvm import closure
    from mfvm.numerics import time_integrators as ti
    ns, U, grid, params, schemes, bcs = _kh(64)
    V = ns.allocate_u({"nx1": 64, "nx2": 64, "nx3": 1})
    V.buf.copy_(U.buf)
    steps = 15
    ref = ns.advance(V, grid, params, schemes, bcs, ti.rk3_ssp._mfvm_code, 0.4, steps)
    net = None
    if mode == "zero":
        net = closure.FluxCorrectionCNN()
        net.zero_()
    elif mode == "huge":
        net = _net(seed=4)
        net.out_scale.fill_(1e8)
    print(f"closure path {mode} graph={graph}", flush=True)
    dts, counts, _ = closure.advance(U, grid, params, schemes, bcs, net, 0.4, steps, graph=graph)
    torch.cuda.synchronize()
    assert torch.equal(dts, ref)
    assert torch.equal(U.buf, V.buf)
    want = [steps] * 3 if mode == "huge" else [0, 0, 0]
    assert counts.cpu().tolist() == want


def test_small_correction_changes_the_solution_without_fallback():
    from mfvm import closure
    ns, U, grid, params, schemes, bcs = _kh(64, seed=1)
    V = ns.allocate_u({"nx1": 64, "nx2": 64, "nx3": 1})
    V.buf.copy_(U.buf)
    net = _net(seed=5)
    net.out_scale.fill_(1e-4)
    _, counts, _ = closure.advance(U, grid, params, schemes, bcs, net, 0.4, 10)
    closure.advance(V, grid, params, schemes, bcs, None, 0.4, 10)
    torch.cuda.synchronize()
    assert counts.cpu().tolist() == [0, 0, 0]
    d = (U.buf - V.buf).abs().max()
    assert 0 < float(d) < 1e-2


@pytest.mark.parametrize("n,graph", [(64, 0), (128, 4)])
def test_fused_inference_matches_the_layer_kernels(n, graph):
    """The fused stage inference (features, three layers and the corrected
    state in one launch, gfx950 K=32 MFMA) against the same network run one
    launch per layer (K=16 MFMA): the same BF16 activations up to the order of
    the FP32 sums inside a 32-wide block, so after 8 steps the states agree to
    1e-9 and no stage falls back in either."""
    from mfvm import closure
    ns, U, grid, params, schemes, bcs = _kh(n, seed=2)
    V = ns.allocate_u({"nx1": n, "nx2": n, "nx3": 1})
    V.buf.copy_(U.buf)
    net = _net(seed=7)
    net.out_scale.fill_(2e-3)
    steps = 8
    print(f"fused inference n={n} graph={graph}", flush=True)
    d1, c1, _ = closure.advance(U, grid, params, schemes, bcs, net, 0.4, steps, graph=graph)
    d2, c2, _ = closure.advance(V, grid, params, schemes, bcs, net, 0.4, steps, graph=graph, layered=True)
    torch.cuda.synchronize()
    assert c1.cpu().tolist() == [0, 0, 0] and c2.cpu().tolist() == [0, 0, 0]
    diff = float((U.buf - V.buf).abs().max())
    assert diff < 1e-9, diff
    assert float((d1 - d2).abs().max()) < 1e-12


def test_t_end_clips_the_last_step():
    from mfvm import closure
    ns, U, grid, params, schemes, bcs = _kh(32)
    dts, _, t = closure.advance(U, grid, params, schemes, bcs, None, 0.4, 40, t_end=0.05)
    torch.cuda.synchronize()
    d = dts.cpu().numpy()
    assert abs(float(t) - 0.05) < 1e-15
    assert abs(d[:-1].sum() - 0.05) < 1e-15
    assert d[-1] == 0.0 and d[-2] == 0.0 and np.all(d >= 0)


def test_training_reduces_the_loss():
    from mfvm import closure
    from mfvm.grid import boundary_conditions as bc
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.minmod, "diffusive_solver": diffusive_solvers.centered}
    bcs = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    net = closure.FluxCorrectionCNN(seed=0)
    tr = closure.Trainer(net, params, schemes, bcs, n_coarse=32, factor=4)
    X, T = tr.generate([11, 12, 13, 14], t_end=0.3, pairs=6, batch=4, compensated=False)
    assert X.shape == (24, 32, 32, 4) and T.shape == (24, 32, 32, 4)
    assert bool(torch.isfinite(T).all())
    tr.set_scales(T)
    hist = tr.train(X, T, epochs=25, batch=4, lr=3e-3)
    print("loss history", [round(h, 4) for h in hist], flush=True)
    assert math.isfinite(hist[-1]) and hist[-1] < 0.8 * hist[0]


def test_trained_closure_beats_the_plain_scheme_on_held_out_seeds():
    """Gain test: the committed network (trained on seeds 1000-1255) against
    the plain 128^2 scheme on two held-out seeds, L1 at t = 2 against the 4x
    block average of the 512^2 plain run."""
    import os
    from mfvm import closure
    from mfvm.grid import boundary_conditions as bc
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    path = os.path.join(os.path.dirname(closure.__file__), "data", "closure_kh128.npz")
    net = closure.FluxCorrectionCNN.load(path)
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.minmod, "diffusive_solver": diffusive_solvers.centered}
    bcs = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    rep = closure.evaluate(net, [5016, 5017], params, schemes, bcs, t_end=2.0)
    print("gain test", rep, flush=True)
    plain, corr = np.array(rep["plain_c"]), np.array(rep["closure_c"])
    assert rep["closure_fallback_stages"] == 0
    assert np.all(corr < plain)              # every field improves
    assert corr.mean() < 0.75 * plain.mean()
