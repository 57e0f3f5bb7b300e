"""Sweeps the oracle's fc3d_nsgs needs in natural order on the 1M-contact pile.

The oracle (the C restatement of the reference fc3d_nsgs with its default
options: hybrid local solver 565, LIGHT_WITH_FULL_FINAL) is run from a zero
reaction in the reference's own natural contact order on the granular pile of
1,000,000 contacts / 400,000 bodies (mu ~ U[0.1, 0.9], seed 5), to an FP64
fc3d_compute_error <= 1e-6.  The sweep count, final residual and wall time are
written to cpu_natural_order_pile1m.json beside this script.  bench.py scales
its timed sample of oracle sweeps by this count (a full solve takes minutes on
one core, longer than a bench run should).

Run from the repository root: python tests/golden/cpu_natural_order_pile1m.py
"""
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)


def main():
    from oracle import pyoracle as O
    from siconos_amd import numerics as N
    t0 = time.perf_counter()
    g = N.generate_granular_pile(400_000, 1_000_000, 0.1, 0.9, 5)
    gen_s = time.perf_counter() - t0
    print(f"generated in {gen_s:.1f} s, nnzb {g.W.nnzb}", flush=True)
    t0 = time.perf_counter()
    info, r, u, it, res = O.nsgs(g.W, g.q, g.mu, tol=1e-6, max_iter=20000)
    dt = time.perf_counter() - t0
    out = {"workload": "granular pile, 1,000,000 contacts / 400,000 bodies, mu ~ U[0.1,0.9], "
                       "seed 5, tol 1e-6, zero start",
           "order": "natural (the reference's)",
           "solver": "oracle fc3d_nsgs, reference default options (565, LIGHT_WITH_FULL_FINAL)",
           "info": int(info), "sweeps": int(it), "residual": float(res),
           "seconds": dt, "ms_per_sweep": dt / max(it, 1) * 1e3, "cores": 1,
           "host": platform.processor() or platform.machine()}
    print(json.dumps(out), flush=True)
    with open(os.path.join(HERE, "cpu_natural_order_pile1m.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
