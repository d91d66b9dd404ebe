"""Batched ensembles of Navier-Stokes instances (mfvm_ens_advance).

The reference advances one grid per process (``src/solver.py:88-151``).  An
:class:`Ensemble` advances many independent instances of different sizes at
once: every phase of an RK stage (ghost fill, closure, fused stage) is one
kernel launch over the tiles of all members, the steps are replayed from a
captured hipGraph, and each member still takes its own CFL time step.  Every
member ends bit for bit where :func:`mfvm.physics.navier_stokes.advance` (or
``advance_compensated``) would have taken it alone.

:meth:`Ensemble.autotune` picks the launch shape -- the stage tile of each
size class, the work-list order, persistent or one-workgroup-per-tile
launches and the XCD-aware order -- by timing candidates on copies of the
members' states.
"""

import ctypes
import time

import torch

from . import _lib
from .physics import _common as c
from .physics import navier_stokes as ns
from .state import State

ORDERS = {"member": _lib.ENS_ORDER_MEMBER, "cost": _lib.ENS_ORDER_COST,
          "interleave": _lib.ENS_ORDER_INTERLEAVE}

DEFAULT_OPTIONS = {"tiles": {}, "order": "cost", "persist": 0, "graph": 50, "remap": 1, "split": 1}

# Stage tile candidates (x1 threads, x2 threads, x3 planes); x1*x2 == 256.
TILES_2D = ((64, 4, 1), (32, 8, 1), (16, 16, 1), (128, 2, 1))
TILES_3D = ((64, 4, 32), (64, 4, 16), (64, 4, 64), (32, 8, 32), (16, 16, 32))


def default_tile(n):
    return (64, 4, min(n[2], 32))


class Ensemble:
    """Independent Navier-Stokes instances advanced together.

    states: mfvm States of the members: float64 (``allocate_u``), or the
    float32 states of ``to_compensated`` with ``compensated=True``.
    grids: the members' ``build_grid`` dicts.  params / space_schemes: shared
    by all members (reference dict formats).  bcs: one ``bcs`` dict for all
    members or a list with one per member.  integrator: an mfvm integrator
    (rk3, rk3_ssp or rk4) or its integer code.
    """

    def __init__(self, states, grids, params, space_schemes, bcs, integrator, compensated=False):
        if len(states) == 0 or len(states) != len(grids):
            raise ValueError("need one grid per state and at least one member")
        self.states = list(states)
        self.grids = [c.require_grid(g, U) for g, U in zip(grids, self.states)]
        for U in self.states:
            c.require_state(U, ns.VARS)
        self.compensated = bool(compensated)
        want = torch.float32 if self.compensated else torch.float64
        if any(U.dtype != want for U in self.states):
            raise TypeError("compensated ensembles take to_compensated states; plain ones float64 states")
        self.params = params
        self.space_schemes = space_schemes
        self.bcs = list(bcs) if isinstance(bcs, (list, tuple)) else [bcs] * len(self.states)
        self.integrator = int(getattr(integrator, "_mfvm_code", integrator))
        self.p = ns.make_params(self.states[0], params, space_schemes)
        self.options = dict(DEFAULT_OPTIONS)
        self._table = None
        self._stream = None

    # ---- members -------------------------------------------------------------
    def _member(self, U, g, bcs, tile, split):
        m = _lib.EnsMember()
        m.grid = g.desc
        rows, _ = c.bc_table(ns.VARS, bcs)
        for v in range(5):
            m.bc[v][:] = rows[v]
        m.tile[:] = list(tile)
        n1, n2, n3 = U.n
        m.U = U.buf.data_ptr()
        m.S1 = U.scratch("S1", tuple(U.buf.shape)).data_ptr()
        m.S2 = U.scratch("S2", tuple(U.buf.shape)).data_ptr()
        m.C = U.scratch("C", (ns.NCLOSURE, n3 + 2, n2 + 2, n1 + 2), zero=True).data_ptr()
        m.kacc = U.scratch("kacc", (5, n3, n2, n1)).data_ptr()
        if self.compensated:
            m.lo = U.scratch("lo", tuple(U.buf.shape)).data_ptr()
            m.dinc = U.scratch("dinc", (5, n3, n2, n1)).data_ptr()
        ntiles = _lib.lib.mfvm_ens_member_tiles(ctypes.byref(m))
        if ntiles < 0:
            _lib.check(_lib.E_ARG)
        m.partials = U.scratch64("ens_partials", 2 * int(ntiles)).data_ptr()
        # split stage: x1/x2 pass into this interior scratch, then the x3 pass
        m.xy = U.scratch("xy", (5, n3, n2, n1)).data_ptr() if split else None
        return m

    def _tile_of(self, U, tiles):
        return tuple(tiles.get(tuple(U.n), default_tile(U.n)))

    def _run(self, states, options, cfl, nsteps):
        tiles = options.get("tiles", {})
        arr = (_lib.EnsMember * len(states))()
        dts = torch.zeros((len(states), nsteps + 1), dtype=torch.float64, device=states[0].device)
        for i, U in enumerate(states):
            arr[i] = self._member(U, self.grids[i], self.bcs[i], self._tile_of(U, tiles),
                                  bool(options.get("split", 1)))
            arr[i].dts = dts[i].data_ptr()
        need = _lib.lib.mfvm_ens_table_bytes(arr, len(states), self.integrator)
        if need < 0:
            _lib.check(_lib.E_UNSUP if self.integrator == _lib.INT_EULER else _lib.E_ARG)
        if self._table is None or self._table.numel() < need:
            self._table = torch.empty(int(need), dtype=torch.uint8, device=states[0].device)
        o = _lib.EnsOpts()
        o.order = ORDERS[options.get("order", "cost")]
        o.persist = int(options.get("persist", 0))
        o.graph = int(options.get("graph", 0))
        o.remap = int(options.get("remap", 1))
        cur = torch.cuda.current_stream()
        if o.graph > 0 and cur.cuda_stream == 0:
            # graph capture needs a stream other than the null stream
            if self._stream is None:
                self._stream = torch.cuda.Stream()
            side = self._stream
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                rc = _lib.lib.mfvm_ens_advance(arr, len(states), ctypes.byref(self.p), ctypes.byref(o),
                                               self.integrator, int(self.compensated), float(cfl),
                                               int(nsteps), c.ptr(self._table), int(need),
                                               ctypes.c_void_p(side.cuda_stream))
            cur.wait_stream(side)
        else:
            rc = _lib.lib.mfvm_ens_advance(arr, len(states), ctypes.byref(self.p), ctypes.byref(o),
                                           self.integrator, int(self.compensated), float(cfl),
                                           int(nsteps), c.ptr(self._table), int(need), c.stream())
        _lib.check(rc)
        return dts

    # ---- public API ------------------------------------------------------------
    def advance(self, cfl, nsteps, options=None):
        """``nsteps`` steps of every member; returns the (members, nsteps+1)
        device tensor of each member's time steps."""
        opts = dict(self.options)
        if options:
            opts.update(options)
        return self._run(self.states, opts, cfl, nsteps)

    def cells(self):
        return sum(U.n[0] * U.n[1] * U.n[2] for U in self.states)

    def _shadow(self):
        out = []
        for U in self.states:
            cfg = {"nx1": U.n[0], "nx2": U.n[1], "nx3": U.n[2]}
            V = State(ns.VARS, cfg, dtype=U.dtype, device=U.device)
            V.buf.copy_(U.buf)
            if self.compensated:
                V.scratch("lo", tuple(U.buf.shape)).copy_(U.scratch("lo", tuple(U.buf.shape)))
            out.append(V)
        return out

    def autotune(self, cfl, steps=3, verbose=False, repeats=2, min_gain=0.01):
        """Pick split or fused stages, tiles per size class, work order,
        persistence and XCD order by coordinate descent over timed runs of
        ``steps`` steps on copies of the members.  Each candidate's time is
        the least of ``repeats`` runs, and a candidate replaces the current
        choice only when it is faster by more than ``min_gain`` (a fraction),
        so that run-to-run noise does not decide.  Stores and returns the
        chosen options."""
        shadow = self._shadow()
        log = []

        def timed(opts):
            o = dict(opts, graph=0)
            self._run(shadow, o, cfl, 1)  # warm-up (and table sizing)
            ms = float("inf")
            for _ in range(repeats):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self._run(shadow, o, cfl, steps)
                torch.cuda.synchronize()
                ms = min(ms, (time.perf_counter() - t0) * 1e3 / steps)
            log.append((dict(o, tiles=dict(o["tiles"])), ms))
            if verbose:
                print(f"autotune split={o['split']} {o['order']:>10} persist={o['persist']} "
                      f"remap={o['remap']} tiles={o['tiles']} -> {ms:.3f} ms/step", flush=True)
            return ms

        best = dict(DEFAULT_OPTIONS, tiles={})
        best_ms = timed(best)
        trial = dict(best, split=1 - best["split"])
        ms = timed(trial)
        if ms < best_ms * (1.0 - min_gain):
            best, best_ms = trial, ms
        classes = {}
        for U in self.states:
            classes[tuple(U.n)] = classes.get(tuple(U.n), 0) + U.n[0] * U.n[1] * U.n[2]
        for shape in sorted(classes, key=lambda k: -classes[k]):
            cands = TILES_2D if shape[2] == 1 else TILES_3D
            for t in cands:
                t = (t[0], t[1], min(t[2], shape[2]))
                if t[0] > max(64, shape[0]) or t == self._tile_of_shape(best, shape):
                    continue
                tiles = dict(best["tiles"])
                tiles[shape] = t
                trial = dict(best, tiles=tiles)
                ms = timed(trial)
                if ms < best_ms * (1.0 - min_gain):
                    best, best_ms = trial, ms
        for key, values in (("order", ("member", "interleave", "cost")), ("persist", (1, 2, 0)),
                            ("remap", (0, 1))):
            for v in values:
                if best[key] == v:
                    continue
                trial = dict(best, **{key: v})
                ms = timed(trial)
                if ms < best_ms * (1.0 - min_gain):
                    best, best_ms = trial, ms
        best["graph"] = DEFAULT_OPTIONS["graph"]
        self.options = best
        self.tuning_log = log
        return best

    @staticmethod
    def _tile_of_shape(opts, shape):
        return tuple(opts["tiles"].get(shape, default_tile(shape)))
