"""Conserved-variable state: named views of one SoA device buffer.

The reference keeps ``U`` as a dict of float64 arrays of shape
(nx1+2, nx2+2, nx3+2) (``navier_stokes.py:34-39``).  Here the dict values are
views with that shape and indexing, backed by one buffer laid out
[var][x3][x2][x1] (x1 contiguous) so that kernels read coalesced rows.
"""

import torch


class State(dict):
    """dict of field views into ``self.buf`` (shape [nvar, nx3+2, nx2+2, nx1+2])."""

    def __init__(self, names, grid_config, dtype=torch.float64, device="cuda"):
        n = tuple(int(grid_config[f"nx{a}"]) for a in (1, 2, 3))
        self.names = tuple(names)
        self.n = n
        self.buf = torch.zeros((len(names), n[2] + 2, n[1] + 2, n[0] + 2), dtype=dtype,
                               device=device)
        for v, name in enumerate(self.names):
            dict.__setitem__(self, name, self.buf[v].permute(2, 1, 0))
        self._work = {}

    def __setitem__(self, key, value):
        # Assignment copies into the shared buffer (the kernels read it).
        self[key].copy_(torch.as_tensor(value, dtype=self.buf.dtype).to(self.buf.device))

    @property
    def dtype(self):
        return self.buf.dtype

    @property
    def device(self):
        return self.buf.device

    def scratch(self, key, shape, zero=False):
        """Per-state device scratch buffer, allocated once per key/shape."""
        t = self._work.get(key)
        if t is None or tuple(t.shape) != tuple(shape):
            t = (torch.zeros if zero else torch.empty)(shape, dtype=self.buf.dtype, device=self.buf.device)
            self._work[key] = t
        return t

    def scratch64(self, key, n):
        t = self._work.get(key)
        if t is None or t.numel() < n:
            t = torch.zeros(n, dtype=torch.float64, device=self.buf.device)
            self._work[key] = t
        return t

    def interior_views(self, t):
        """Reference-shaped (nx1, nx2, nx3) views of an interior buffer [nvar, nx3, nx2, nx1]."""
        return {name: t[v].permute(2, 1, 0) for v, name in enumerate(self.names)}


def is_state(U):
    return isinstance(U, State)
