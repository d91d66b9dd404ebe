"""Pins of the CPU oracle of the learned closure (oracle/cnn.py).

* the periodic convolution equals the explicit sum over taps (its
  definition), and a one-hot kernel shifts the input by the tap offset;
* the correction kc of a flux that is constant in space is zero, and the
  correction of a single face flux moves mass from one cell to its neighbour
  (conservation: sum of kc is zero);
* torch autograd of the float64 network + loss agrees with central finite
  differences on every parameter tensor.
"""

import torch

from oracle import cnn


def _params(seed=0, scale=0.3):
    g = torch.Generator().manual_seed(seed)
    shapes = [(32, 4, 3, 3), (32,), (32, 32, 3, 3), (32,), (8, 32, 3, 3), (8,)]
    return [(torch.randn(s, generator=g, dtype=torch.float64) * scale) for s in shapes]


def test_conv_matches_loop_and_shift():
    g = torch.Generator().manual_seed(1)
    x = torch.randn((2, 4, 6, 16), generator=g, dtype=torch.float64)
    w = torch.randn((5, 4, 3, 3), generator=g, dtype=torch.float64)
    b = torch.randn(5, generator=g, dtype=torch.float64)
    assert torch.allclose(cnn.conv_periodic(x, w, b), cnn.conv_loop(x, w, b), atol=1e-12)
    # one-hot tap (ky, kx) = (2, 0): out[j, i] = x[j + 1, i - 1]
    w1 = torch.zeros((1, 1, 3, 3), dtype=torch.float64)
    w1[0, 0, 2, 0] = 1.0
    out = cnn.conv_periodic(x[:, :1], w1)
    want = torch.roll(x[:, :1], shifts=(-1, 1), dims=(2, 3))
    assert torch.equal(out, want)


def test_correction_conserves_and_vanishes_for_uniform_flux():
    y = torch.zeros((1, 4, 16, 8), dtype=torch.float64)
    y[..., :] = 0.7
    osc = torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64)
    assert torch.count_nonzero(cnn.correction(y, osc, 0.1, 0.2)) == 0
    y = torch.zeros((1, 4, 16, 8), dtype=torch.float64)
    y[0, 2, 5, 0] = 1.0        # x1 face i-1/2 of cell (i=5, j=2), mass
    kc = cnn.correction(y, osc, 0.5, 1.0)
    assert kc[0, 2, 5, 0] == 2.0 and kc[0, 2, 4, 0] == -2.0   # +dF/dx into cell 5, out of cell 4
    assert abs(float(kc.sum())) < 1e-15


def test_autograd_matches_finite_differences():
    torch.manual_seed(0)
    params = [p.requires_grad_() for p in _params()]
    g = torch.Generator().manual_seed(2)
    x = torch.randn((2, 5, 16, 4), generator=g, dtype=torch.float64)
    tgt = torch.randn((2, 5, 16, 4), generator=g, dtype=torch.float64)
    osc = torch.tensor([0.5, 1.0, 1.5, 2.0], dtype=torch.float64)
    wf = torch.tensor([1.0, 0.5, 0.25, 2.0], dtype=torch.float64)
    L = cnn.loss(cnn.forward(x, params), tgt, osc, wf, 0.25, 0.5)
    grads = torch.autograd.grad(L, params)
    eps = 1e-6
    for p, gp in zip(params, grads):
        flat = p.detach().view(-1)
        for idx in torch.randperm(flat.numel(), generator=g)[:6]:
            old = float(flat[idx])
            with torch.no_grad():
                flat[idx] = old + eps
                lp = float(cnn.loss(cnn.forward(x, params), tgt, osc, wf, 0.25, 0.5))
                flat[idx] = old - eps
                lm = float(cnn.loss(cnn.forward(x, params), tgt, osc, wf, 0.25, 0.5))
                flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = float(gp.view(-1)[idx])
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (p.shape, int(idx), fd, an)
