"""Command-line entry point (drop-in for main.py):

    python -m mfvm --setup kelvin_helmholtz_128

runs ``mfvm.setups.<name>.get_setup()`` through :func:`mfvm.solver.run`
(or, for ``ensemble_kh64``, all members through ``run_ensemble``)."""

import argparse
import importlib
import time

from . import solver


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run a finite-volume simulation setup.")
    ap.add_argument("--setup", required=True, help="Setup name (e.g. kelvin_helmholtz_128).")
    args = ap.parse_args(argv)
    mod = importlib.import_module(f"mfvm.setups.{args.setup}")
    t0 = time.perf_counter()
    if hasattr(mod, "get_setups"):
        states, _ = solver.run_ensemble(mod.get_setups())
        print(f"{len(states)} members advanced in {time.perf_counter() - t0:.2f} s")
    else:
        solver.run(mod.get_setup())
        print(f"done in {time.perf_counter() - t0:.2f} s")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
