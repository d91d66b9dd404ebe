"""CPU oracle of the learned flux-correction network and its training loss.

TEST INFRASTRUCTURE ONLY (see the package docstring).  No reference
counterpart: this is the plain torch-CPU statement of what the HIP kernels
of the closure compute, used to check their forward pass, their gradients
and the loss adjoint.

Network: x (channels-last [B, n2, n1, 4]) -> conv3x3(4->32) -> relu ->
conv3x3(32->32) -> relu -> conv3x3(32->8) = y, each convolution a
cross-correlation with periodic padding in x1 and x2 and weights
[cout, cin, 3, 3] (torch's conv2d layout; tap (ky, kx) reads pixel
(j + ky - 1, i + kx - 1)).

Correction of the right-hand side from y (dF = y * out_scale, x1 faces in
channels 0-3, x2 faces in 4-7, face i-1/2 / j-1/2 stored at cell (i, j)):
    kc(i,j) = -((dF1(i+1,j) - dF1(i,j)) / dx1 + (dF2(i,j+1) - dF2(i,j)) / dx2).
Loss: mean over cells of sum_f wf[f] (kc_f - target_f)^2.
"""

import torch
import torch.nn.functional as F


def conv_periodic(x, w, b=None):
    """x [B, C, n2, n1], w [cout, cin, 3, 3] -> [B, cout, n2, n1] (periodic)."""
    xp = F.pad(x, (1, 1, 1, 1), mode="circular")
    return F.conv2d(xp, w, b)


def conv_loop(x, w, b=None):
    """The same convolution as an explicit sum over taps (the definition)."""
    B, C, n2, n1 = x.shape
    out = torch.zeros((B, w.shape[0], n2, n1), dtype=x.dtype)
    for ky in range(3):
        for kx in range(3):
            sh = torch.roll(x, shifts=(1 - ky, 1 - kx), dims=(2, 3))  # sh[j,i] = x[j+ky-1, i+kx-1]
            out += torch.einsum("oc,bcji->boji", w[:, :, ky, kx], sh)
    if b is not None:
        out += b[None, :, None, None]
    return out


def forward(x_cl, params, keep=False):
    """x_cl [B, n2, n1, 4] -> y [B, n2, n1, 8]; params = (w1, b1, w2, b2, w3, b3)."""
    w1, b1, w2, b2, w3, b3 = params
    x = x_cl.permute(0, 3, 1, 2)
    a1 = torch.relu(conv_periodic(x, w1, b1))
    a2 = torch.relu(conv_periodic(a1, w2, b2))
    y = conv_periodic(a2, w3, b3)
    y_cl = y.permute(0, 2, 3, 1)
    if keep:
        return y_cl, a1.permute(0, 2, 3, 1), a2.permute(0, 2, 3, 1)
    return y_cl


def correction(y_cl, out_scale, dx1, dx2):
    """kc [B, n2, n1, 4] from the network output y [B, n2, n1, 8]."""
    d1 = y_cl[..., 0:4] * out_scale
    d2 = y_cl[..., 4:8] * out_scale
    return -((torch.roll(d1, -1, dims=2) - d1) / dx1 + (torch.roll(d2, -1, dims=1) - d2) / dx2)


def loss(y_cl, target, out_scale, wf, dx1, dx2):
    kc = correction(y_cl, out_scale, dx1, dx2)
    return ((kc - target) ** 2 * wf).sum(-1).mean()
