"""Learned flux-correction closure for 2D periodic Navier-Stokes.

No counterpart in the reference.  A 3-layer 3x3 convolutional network
(4 -> 32 -> 32 -> 8 channels, ReLU, periodic padding; BF16 weights and
activations, FP32 accumulation on MFMA, all in libmfvm.so) maps the
normalised primitives (rho, v1, v2, p) of a coarse state to corrections dF of
the x1- and x2-face fluxes of mass, momentum 1, momentum 2 and energy.  Each
SSPRK3 stage uses k - div(dF) instead of the plain right-hand side k; if the
corrected stage state has rho <= 0 or p <= 0 anywhere, that stage keeps the
plain MUSCL/HLL result (the positivity fallback).  The receptive field of
three stacked 3x3 layers is 7x7 cells around each cell (it contains the 5x5
stencil around the cell's two faces).

Training (on the GPU, with the library's own forward and backward kernels):
coarse states are 4x block averages of fine plain-method runs; the target
of a coarse state Uc(t) is the coarse-grained tendency the plain coarse
scheme misses, R = (Uc(t + tau) - Uc(t)) / tau - k(Uc(t)), with tau a few
fine steps; the loss is the per-field weighted mean square of -div(dF) - R.
"""

import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .physics import _common as c
from .physics import navier_stokes as ns
from .state import State

LAYERS = ((4, 32), (32, 32), (32, 8))
FIELDS = (0, 1, 2, 4)  # mass, momentum 1, momentum 2, energy

# Fixed normalisation of the primitives (Kelvin-Helmholtz: rho in [1, 3],
# |v| <= 0.7, p ~ 2.5): x = (q - shift) * scale.
IN_SHIFT = (1.5, 0.0, 0.0, 2.5)
IN_SCALE = (2.0, 2.0, 2.0, 5.0)


def _round16(n):
    return (n + 15) // 16 * 16


class FluxCorrectionCNN:
    """Parameters, optimiser state and the packed BF16 weights of the network."""

    def __init__(self, device="cuda", seed=0, last_scale=0.1):
        gen = torch.Generator().manual_seed(seed)
        self.device = torch.device(device)
        self.w, self.b = [], []
        for layer, (cin, cout) in enumerate(LAYERS):
            std = math.sqrt(2.0 / (9 * cin)) * (last_scale if layer == 2 else 1.0)
            self.w.append((torch.randn((cout, cin, 3, 3), generator=gen) * std).float().to(self.device))
            self.b.append(torch.zeros(cout, dtype=torch.float32, device=self.device))
        self.in_shift = torch.tensor(IN_SHIFT, dtype=torch.float32, device=self.device)
        self.in_scale = torch.tensor(IN_SCALE, dtype=torch.float32, device=self.device)
        self.out_scale = torch.ones(4, dtype=torch.float32, device=self.device)
        self.gw = [torch.zeros_like(w) for w in self.w]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.m = [torch.zeros_like(p) for p in self.params()]
        self.v = [torch.zeros_like(p) for p in self.params()]
        self.t = 0
        self.wf = [torch.empty(9 * _round16(co) * ci, dtype=torch.int16, device=self.device) for ci, co in LAYERS]
        self.wb = [torch.empty(9 * _round16(ci) * co, dtype=torch.int16, device=self.device) for ci, co in LAYERS]
        self.pack()

    # ---- parameters ------------------------------------------------------------
    def params(self):
        return [self.w[0], self.b[0], self.w[1], self.b[1], self.w[2], self.b[2]]

    def grads(self):
        return [self.gw[0], self.gb[0], self.gw[1], self.gb[1], self.gw[2], self.gb[2]]

    def pack(self):
        for layer, (cin, cout) in enumerate(LAYERS):
            _lib.check(_lib.lib.mfvm_cnn_pack(cin, cout, c.ptr(self.w[layer]), c.ptr(self.wf[layer]),
                                              c.ptr(self.wb[layer]), c.stream()))

    def zero_(self):
        """All weights and biases zero: the network outputs exactly 0."""
        for p in self.params():
            p.zero_()
        self.pack()

    def state_dict(self):
        d = {f"w{i}": self.w[i].cpu().numpy() for i in range(3)}
        d.update({f"b{i}": self.b[i].cpu().numpy() for i in range(3)})
        d["in_shift"] = self.in_shift.cpu().numpy()
        d["in_scale"] = self.in_scale.cpu().numpy()
        d["out_scale"] = self.out_scale.cpu().numpy()
        return d

    def load_state_dict(self, d):
        for i in range(3):
            self.w[i].copy_(torch.from_numpy(np.asarray(d[f"w{i}"], dtype=np.float32)))
            self.b[i].copy_(torch.from_numpy(np.asarray(d[f"b{i}"], dtype=np.float32)))
        for k in ("in_shift", "in_scale", "out_scale"):
            getattr(self, k).copy_(torch.from_numpy(np.asarray(d[k], dtype=np.float32)))
        self.pack()

    def save(self, path):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path, device="cuda"):
        net = cls(device=device)
        with np.load(path, allow_pickle=False) as d:
            net.load_state_dict({k: d[k] for k in d.files})
        return net

    def c_struct(self):
        s = _lib.CNN()
        for i in range(3):
            s.wf[i] = self.wf[i].data_ptr()
            s.b[i] = self.b[i].data_ptr()
        s.in_shift = self.in_shift.data_ptr()
        s.in_scale = self.in_scale.data_ptr()
        s.out_scale = self.out_scale.data_ptr()
        return s

    # ---- passes ---------------------------------------------------------------
    def features(self, U, grid, gamma, batch=1, batch_stride=0):
        """BF16 network input [batch, n2, n1, 4] from ghosted 2D state(s)."""
        g = grid["mfvm"] if isinstance(grid, dict) else grid
        n1, n2 = g.n[0], g.n[1]
        buf = U.buf if isinstance(U, State) else U
        x = torch.empty((batch, n2, n1, 4), dtype=torch.int16, device=self.device)
        code = _lib.F64 if buf.dtype == torch.float64 else _lib.F32
        _lib.check(_lib.lib.mfvm_cnn_features(g.ptr(), code, float(gamma), batch, int(batch_stride),
                                              c.ptr(buf), c.ptr(self.in_shift), c.ptr(self.in_scale),
                                              c.ptr(x), c.stream()))
        return x

    def forward(self, x):
        """x [B, n2, n1, 4] BF16 -> (y [B, n2, n1, 8] FP32, a1, a2 BF16)."""
        B, n2, n1, _ = x.shape
        a1 = torch.empty((B, n2, n1, 32), dtype=torch.int16, device=self.device)
        a2 = torch.empty_like(a1)
        y = torch.empty((B, n2, n1, 8), dtype=torch.float32, device=self.device)
        for layer, (inp, out, mode) in enumerate(((x, a1, _lib.CNN_FWD_RELU), (a1, a2, _lib.CNN_FWD_RELU),
                                                  (a2, y, _lib.CNN_FWD_LINEAR))):
            cin, cout = LAYERS[layer]
            _lib.check(_lib.lib.mfvm_cnn_conv(mode, cin, cout, B, n1, n2, c.ptr(inp), c.ptr(self.wf[layer]),
                                              c.ptr(self.b[layer]), None, c.ptr(out), c.stream()))
        return y, a1, a2

    def zero_grad(self):
        for gt in self.grads():
            gt.zero_()

    def backward(self, x, a1, a2, gy):
        """Accumulate parameter gradients from gy = dL/dy (BF16 [B, n2, n1, 8])."""
        B, n2, n1, _ = x.shape
        s = c.stream()
        lib = _lib.lib
        _lib.check(lib.mfvm_cnn_wgrad(32, 8, B, n1, n2, c.ptr(a2), c.ptr(gy), c.ptr(self.gw[2]),
                                      c.ptr(self.gb[2]), s))
        ga2 = torch.empty_like(a2)
        _lib.check(lib.mfvm_cnn_conv(_lib.CNN_BWD_DATA, 8, 32, B, n1, n2, c.ptr(gy), c.ptr(self.wb[2]), None,
                                     c.ptr(a2), c.ptr(ga2), s))
        _lib.check(lib.mfvm_cnn_wgrad(32, 32, B, n1, n2, c.ptr(a1), c.ptr(ga2), c.ptr(self.gw[1]),
                                      c.ptr(self.gb[1]), s))
        ga1 = torch.empty_like(a1)
        _lib.check(lib.mfvm_cnn_conv(_lib.CNN_BWD_DATA, 32, 32, B, n1, n2, c.ptr(ga2), c.ptr(self.wb[1]), None,
                                     c.ptr(a1), c.ptr(ga1), s))
        _lib.check(lib.mfvm_cnn_wgrad(4, 32, B, n1, n2, c.ptr(x), c.ptr(ga1), c.ptr(self.gw[0]),
                                      c.ptr(self.gb[0]), s))

    def loss(self, y, target, wf, dx1, dx2):
        """(loss tensor [1], gy) of the weighted mean-square correction error."""
        B, n2, n1, _ = y.shape
        gy = torch.empty((B, n2, n1, 8), dtype=torch.int16, device=self.device)
        out = torch.zeros(1, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.mfvm_cnn_loss(B, n1, n2, float(dx1), float(dx2), c.ptr(y), c.ptr(target),
                                          c.ptr(self.out_scale), c.ptr(wf), c.ptr(gy), c.ptr(out), c.stream()))
        return out, gy

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        for p, g, m, v in zip(self.params(), self.grads(), self.m, self.v):
            _lib.check(_lib.lib.mfvm_cnn_adam(p.numel(), c.ptr(p), c.ptr(g), c.ptr(m), c.ptr(v), lr, b1, b2, eps,
                                              self.t, c.stream()))
        self.pack()


# ----------------------------------------------------------------------------
# Time stepping with the closure
# ----------------------------------------------------------------------------

def _closure_work(U, layered=False):
    n1, n2, _ = U.n
    w = _lib.ClosureWork()
    w.layered = int(bool(layered))
    w.a1 = _int16(U, "cnn_a1", (n2, n1, 32)).data_ptr()
    w.a2 = _int16(U, "cnn_a2", (n2, n1, 32)).data_ptr()
    w.y = _typed(U, "cnn_y", (n2, n1, 8), torch.float32).data_ptr()
    w.T = U.scratch("cnn_T", tuple(U.buf.shape)).data_ptr()
    w.flag = _typed(U, "cnn_flag", (3,), torch.int32).data_ptr()
    w.counts = _typed(U, "cnn_counts", (3,), torch.int32).data_ptr()
    w.dt_cur = _typed(U, "cnn_dt", (1,), torch.float64).data_ptr()
    w.t_cur = _typed(U, "cnn_t", (1,), torch.float64).data_ptr()
    w.step = _typed(U, "cnn_step", (1,), torch.int32).data_ptr()
    return w


def _typed(U, key, shape, dtype):
    t = U._work.get(key)
    if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
        t = torch.zeros(shape, dtype=dtype, device=U.device)
        U._work[key] = t
    return t


def _int16(U, key, shape):
    return _typed(U, key, shape, torch.int16)


_SIDE = {}


def advance(U, grid, params, space_schemes, bcs, net, cfl, nsteps, t_end=math.inf, graph=0, layered=False):
    """``nsteps`` SSPRK3 steps with the learned correction (net=None: the
    plain scheme).  dt follows the CFL rule, clipped at ``t_end``.  The
    network and the correction of a stage run as one fused launch
    (layered=True: one launch per layer, the training kernels).  Returns
    (dts [nsteps+1] device tensor, fallback counts per stage [3] device tensor,
    time reached as a device scalar)."""
    c.require_state(U, ns.VARS)
    g = c.require_grid(grid, U)
    p = ns.make_params(U, params, space_schemes, bcs)
    w = ns.work(U)
    nparts = _lib.lib.mfvm_ns_partials_size(g.ptr(), ctypes.byref(p))
    w.partials = U.scratch64("partials", max(int(nparts), 1)).data_ptr()
    dts = torch.zeros(nsteps + 1, dtype=torch.float64, device=U.device)
    w.dts = dts.data_ptr()
    cw = _closure_work(U, layered)
    cnn = ctypes.byref(net.c_struct()) if net is not None else None
    cur = torch.cuda.current_stream()
    if graph > 0 and cur.cuda_stream == 0:
        side = _SIDE.setdefault(U.device, torch.cuda.Stream(device=U.device))
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            rc = _lib.lib.mfvm_closure_advance(g.ptr(), ctypes.byref(p), ctypes.byref(w), cnn, ctypes.byref(cw),
                                               float(cfl), float(t_end), int(nsteps), int(graph),
                                               ctypes.c_void_p(side.cuda_stream))
        cur.wait_stream(side)
    else:
        rc = _lib.lib.mfvm_closure_advance(g.ptr(), ctypes.byref(p), ctypes.byref(w), cnn, ctypes.byref(cw),
                                           float(cfl), float(t_end), int(nsteps), int(graph), c.stream())
    _lib.check(rc)
    # time reached: the dts of the steps taken (dts[nsteps] is the next step's)
    return dts, U._work["cnn_counts"], dts[:nsteps].sum()


# ----------------------------------------------------------------------------
# Training data and training
# ----------------------------------------------------------------------------

def kh_initial(n, seed, amplitude=0.01, noise=1e-3):
    """Kelvin-Helmholtz state on an n x n grid (rho 2 / 1, shear +-0.5,
    p 2.5) with a seeded perturbation: vy = amplitude sin(4 pi x + phase) +
    uniform noise, phase and noise from ``seed``."""
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0.0, 2 * np.pi)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    vy = amplitude * np.sin(4 * np.pi * X + phase) + rng.uniform(-noise, noise, (n, n))
    E = 2.5 / 0.4 + 0.5 * rho * (vx * vx + vy * vy)
    z = np.zeros_like(rho)
    return [a[:, :, None] for a in (rho, rho * vx, rho * vy, z, E)]


def coarse_grain(U, factor):
    """Block averages of the conserved fields over factor x factor cells
    (interior, float64) -> tensor [5, n2/f, n1/f]."""
    n1, n2, _ = U.n
    inner = U.buf[:, 1, 1:-1, 1:-1]
    return inner.reshape(5, n2 // factor, factor, n1 // factor, factor).mean(dim=(2, 4))


def state_from(fields_2d, grid_config):
    """Ghost-free coarse fields [5, n2, n1] -> float64 State (ghosts unset)."""
    U = ns.allocate_u(grid_config)
    U.buf[:, 1, 1:-1, 1:-1] = fields_2d.to(U.buf.dtype)
    return U


class Trainer:
    """Training data from fine plain-method runs and the training loop."""

    def __init__(self, net, params, space_schemes, bcs, n_coarse=128, factor=4, cfl=0.4):
        self.net = net
        self.params = params
        self.schemes = space_schemes
        self.bcs = bcs
        self.nc = n_coarse
        self.factor = factor
        self.cfl = cfl
        from .grid.geometry import build_grid
        self.cfg_c = {"x1_min": 0.0, "x1_max": 1.0, "nx1": n_coarse, "x2_min": 0.0, "x2_max": 1.0,
                      "nx2": n_coarse, "x3_min": 0.0, "x3_max": 1.0, "nx3": 1}
        nf = n_coarse * factor
        self.cfg_f = dict(self.cfg_c, nx1=nf, nx2=nf)
        self.grid_c = build_grid("cartesian", self.cfg_c)
        self.grid_f = build_grid("cartesian", self.cfg_f)
        self.dx = 1.0 / n_coarse

    def fine_states(self, seeds):
        out = []
        for s in seeds:
            U = ns.allocate_u(self.cfg_f)
            for k, a in zip(ns.VARS, kh_initial(self.cfg_f["nx1"], s)):
                U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(a).to(U.device)
            out.append(U)
        return out

    def samples_from(self, A, B, tau):
        """Features and targets of coarse snapshot pairs A -> B (lists of
        [5, nc, nc] tensors) over intervals tau."""
        from . import solver
        xs, ts = [], []
        for a, b, dt in zip(A, B, tau):
            Uc = state_from(a, self.cfg_c)
            solver.apply_bcs(Uc, self.bcs)
            k = ns.compute_rhs(Uc, self.grid_c, self.params, self.schemes)
            kk = torch.stack([k[ns.VARS[f]] for f in FIELDS], dim=-1)  # [n1, n2, 1, 4]
            kk = kk[:, :, 0, :].permute(1, 0, 2)                        # [n2, n1, 4]
            R = ((b - a)[list(FIELDS)] / dt).permute(1, 2, 0) - kk       # [n2, n1, 4]
            xs.append(self.net.features(Uc, self.grid_c, self.params["heat_capacity_ratio"])[0])
            ts.append(R.float())
        return torch.stack(xs), torch.stack(ts)

    def generate(self, seeds, t_end=2.0, pairs=40, gap=4, batch=32, compensated=True, log=print):
        """Run the fine plain method for ``seeds`` (in batched ensembles) and
        collect ``pairs`` coarse snapshot pairs per run, evenly spaced in
        steps up to about ``t_end``."""
        from .ensemble import Ensemble
        from .numerics import time_integrators as ti
        X, T = [], []
        for b0 in range(0, len(seeds), batch):
            chunk = seeds[b0:b0 + batch]
            t0 = time.perf_counter()
            S = self.fine_states(chunk)
            if compensated:
                S = [ns.to_compensated(U) for U in S]
            ens = Ensemble(S, [self.grid_f] * len(S), self.params, self.schemes, self.bcs, ti.rk3_ssp,
                           compensated=compensated)
            # step count to t_end from the first dt (the KH dt changes little)
            d0 = ens.advance(self.cfl, 1, {"graph": 0})
            steps = int(math.ceil(t_end / float(d0[:, 0].min()))) - 1
            every = max(gap + 1, steps // pairs)
            for _ in range(pairs):
                ens.advance(self.cfl, every - gap)
                A = [coarse_grain(self._f64(U), self.factor) for U in S]
                d = ens.advance(self.cfl, gap)
                Bv = [coarse_grain(self._f64(U), self.factor) for U in S]
                tau = d[:, :gap].sum(dim=1).tolist()
                x, tt = self.samples_from(A, Bv, tau)
                X.append(x)
                T.append(tt)
            torch.cuda.synchronize()
            log(f"generated runs {b0}..{b0 + len(chunk) - 1}: {pairs} pairs each, "
                f"{steps} steps, {time.perf_counter() - t0:.1f} s")
            del ens, S
        return torch.cat(X), torch.cat(T)

    @staticmethod
    def _f64(U):
        if U.dtype == torch.float64:
            return U
        V = State(ns.VARS, {"nx1": U.n[0], "nx2": U.n[1], "nx3": U.n[2]}, device=U.device)
        V.buf.copy_(ns.compensated_value(U))
        return V

    def set_scales(self, T):
        """Output scales and loss weights from the target statistics."""
        std = T.reshape(-1, 4).double().std(dim=0).clamp_min(1e-12)
        self.net.out_scale.copy_((std * self.dx).float())
        self.wf = (1.0 / std ** 2).float().to(self.net.device)
        return std

    def train(self, X, T, epochs=10, batch=16, lr=1e-3, cosine=False, log=print):
        """Adam over shuffled minibatches; ``cosine``: the step size follows
        lr * (1 + cos(pi * epoch / epochs)) / 2 down to zero."""
        n = X.shape[0]
        hist = []
        base = lr
        for ep in range(epochs):
            if cosine:
                lr = base * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))
            perm = torch.randperm(n, device=X.device)
            tot = torch.zeros(1, dtype=torch.float32, device=X.device)
            nb = 0
            for i in range(0, n - batch + 1, batch):
                idx = perm[i:i + batch]
                x = X[idx].contiguous()
                tg = T[idx].contiguous()
                y, a1, a2 = self.net.forward(x)
                L, gy = self.net.loss(y, tg, self.wf, self.dx, self.dx)
                self.net.zero_grad()
                self.net.backward(x, a1, a2, gy)
                self.net.adam_step(lr)
                tot += L
                nb += 1
            hist.append(float(tot) / max(nb, 1))
            log(f"epoch {ep}: loss {hist[-1]:.4f}")
        return hist


# ----------------------------------------------------------------------------
# Evaluation: error against the coarse-grained fine solution, and wall time
# ----------------------------------------------------------------------------

def _run_to(U, grid, params, schemes, bcs, net, cfl, t_end, graph=50):
    """Advance U to exactly t_end in chunks sized from the current dt;
    returns (wall ms, steps taken, stages that fell back)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, steps, fb, dt = 0.0, 0, 0, None
    while t < t_end * (1.0 - 1e-13):
        n = 1 if dt is None else max(1, int(math.ceil((t_end - t) / dt)))
        dts, counts, _ = advance(U, grid, params, schemes, bcs, net, cfl, n, t_end=t_end - t,
                                 graph=graph if n > graph else 0)
        d = dts.cpu()
        t += float(d[:n].sum())
        steps += int((d[:n] > 0).sum())
        fb += int(counts.sum())
        dt = float(d[n]) if float(d[n]) > 0 else float(d[:n].max())
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3, steps, fb


def evaluate(net, seeds, params, schemes, bcs, n_coarse=128, factor=4, t_end=2.0, cfl=0.4, log=print):
    """Per-field L1 error at t_end against the 4x block average of the fine
    plain run, for plain and corrected runs at n_coarse and the plain run at
    2 n_coarse (block-averaged to n_coarse); plus wall times (graph replay)."""
    from .grid.geometry import build_grid
    nf = n_coarse * factor

    def cfg(n):
        return {"x1_min": 0.0, "x1_max": 1.0, "nx1": n, "x2_min": 0.0, "x2_max": 1.0, "nx2": n,
                "x3_min": 0.0, "x3_max": 1.0, "nx3": 1}

    grids = {n: build_grid("cartesian", cfg(n)) for n in (n_coarse, 2 * n_coarse, nf)}
    res = {"plain_c": [], "closure_c": [], "plain_2c": []}
    wall = {"plain_c": 0.0, "closure_c": 0.0, "plain_2c": 0.0, "plain_f": 0.0}
    fallbacks = 0
    for s in seeds:
        fine = ns.allocate_u(cfg(nf))
        for k, a in zip(ns.VARS, kh_initial(nf, s)):
            fine[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(a).to(fine.device)
        ic = {n: coarse_grain(fine, nf // n) for n in (n_coarse, 2 * n_coarse)}
        ms, _, _ = _run_to(fine, grids[nf], params, schemes, bcs, None, cfl, t_end)
        wall["plain_f"] += ms
        truth = coarse_grain(fine, factor)
        for key, n, nn in (("plain_c", n_coarse, None), ("closure_c", n_coarse, net), ("plain_2c", 2 * n_coarse, None)):
            U = state_from(ic[n], cfg(n))
            ms, _, fb = _run_to(U, grids[n], params, schemes, bcs, nn, cfl, t_end)
            wall[key] += ms
            if key == "closure_c":
                fallbacks += fb
            got = coarse_grain(U, n // n_coarse)
            res[key].append((got - truth)[list(FIELDS)].abs().mean(dim=(1, 2)).cpu().numpy())
        log(f"seed {s}: L1 mass plain {res['plain_c'][-1][0]:.3e} closure {res['closure_c'][-1][0]:.3e} "
            f"plain2x {res['plain_2c'][-1][0]:.3e}")
    out = {k: np.mean(np.stack(v), axis=0).tolist() for k, v in res.items()}
    out["wall_ms_per_run"] = {k: v / len(seeds) for k, v in wall.items()}
    out["closure_fallback_stages"] = fallbacks
    out["seeds"] = list(seeds)
    return out
