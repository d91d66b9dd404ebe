"""Train the learned flux-correction closure on the GPU and score it.

    python -m mfvm.train_closure [--train 256] [--heldout 32] [--pairs 32]
                                 [--epochs 20] [--out PATH] [--report PATH]

Kelvin-Helmholtz, 128^2 coarse grid: training data are 4x block averages of
512^2 MUSCL/minmod/HLL/SSPRK3 runs (mu = 1e-4, CFL 0.4) to t = 2 from seeded
perturbations (seeds 1000.. for training, 5000.. held out); the network is
trained with the library's BF16 MFMA forward and backward kernels and Adam;
the held-out runs score plain 128^2, corrected 128^2 and plain 256^2 against
the coarse-grained 512^2 solution at t = 2 (per-field L1) and time them.
"""

import argparse
import json
import os
import sys
import time

import torch

from . import closure
from .grid import boundary_conditions as bc
from .numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
from .physics import navier_stokes as ns

PARAMS = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4, "gravitational_acceleration": [0.0, 0.0, 0.0]}
SCHEMES = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
           "slope_limiter": slope_limiters.minmod, "diffusive_solver": diffusive_solvers.centered}
BCS = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "closure_kh128.npz")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--heldout", type=int, default=32)
    ap.add_argument("--pairs", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--cosine", action="store_true", help="cosine decay of the Adam step size")
    ap.add_argument("--t-end", type=float, default=2.0)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default="")
    ap.add_argument("--eval-only", default="", metavar="WEIGHTS",
                    help="score saved weights on the held-out seeds without training")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t0 = time.perf_counter()
    if args.eval_only:
        net = closure.FluxCorrectionCNN.load(args.eval_only)
        rep = closure.evaluate(net, list(range(5000, 5000 + args.heldout)), PARAMS, SCHEMES, BCS,
                               t_end=args.t_end, log=log)
        rep["weights"] = os.path.basename(args.eval_only)
        rep["total_s"] = time.perf_counter() - t0
        line = json.dumps(rep)
        print(line, flush=True)
        if args.report:
            with open(args.report, "w") as f:
                f.write(line + "\n")
        return 0
    net = closure.FluxCorrectionCNN(seed=0)
    tr = closure.Trainer(net, PARAMS, SCHEMES, BCS, n_coarse=128, factor=4)
    X, T = tr.generate(list(range(1000, 1000 + args.train)), t_end=args.t_end, pairs=args.pairs, log=log)
    std = tr.set_scales(T)
    log(f"data: {X.shape[0]} samples, target std {std.tolist()}, {time.perf_counter() - t0:.0f} s")
    hist = tr.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, cosine=args.cosine, log=log)
    del X, T
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    log(f"saved {args.out}, {time.perf_counter() - t0:.0f} s")
    rep = closure.evaluate(net, list(range(5000, 5000 + args.heldout)), PARAMS, SCHEMES, BCS,
                           t_end=args.t_end, log=log)
    rep["loss_history"] = hist
    rep["train_runs"] = args.train
    rep["pairs_per_run"] = args.pairs
    rep["total_s"] = time.perf_counter() - t0
    line = json.dumps(rep)
    print(line, flush=True)
    if args.report:
        with open(args.report, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
