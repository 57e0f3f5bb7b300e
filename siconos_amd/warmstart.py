"""Learned warm start: training pairs from the GPU solver, and training.

The network (fc3d_hip.h, ``fc3d_hip_mlp_*``) maps 22 features of a contact's
local problem to its converged reaction, scaled by W_cc(0,0) / |q_c|.

The pairs come from the GPU solver itself: each training problem is solved
from a zero start to the requested tolerance. Then
``fc3d_hip_problem_features`` reads the features together with the scaled
converged reaction. Training runs on the GPU, with the library's own MFMA
forward and backward kernels and Adam.
"""
import numpy as np

from . import numerics as N


def collect(problems, tol=1e-6, max_iter=2000, device=0):
    """Features and targets of every contact of the converged problems.

    ``problems`` yields (SparseBlock3, q, mu). Problems that do not converge
    are skipped, so their contacts never become targets. Returns (x, t,
    sweeps), where sweeps lists the zero-start sweep count of each problem.
    """
    xs, ts, sweeps = [], [], []
    params = N.hip_params(tol=tol, max_iter=max_iter, warm_start=1)
    for W, q, mu in problems:
        H = N.HipProblem(W, q, mu, device=device)
        s = H.solve(params)
        if s.info != 0 or not s.residual <= tol:
            continue
        f, t = H.features(target=True)
        xs.append(f)
        ts.append(t)
        sweeps.append(s.iterations)
        del H
    if not xs:
        raise RuntimeError("no training problem converged")
    return np.concatenate(xs), np.concatenate(ts), sweeps


def train(x, t, seed=0, epochs=30, batch=1024, lr=1e-3, device=0):
    """A new model trained on (x, t); returns (model, last minibatch loss)."""
    m = N.Mlp(seed=seed, device=device)
    loss = m.train(x, t, epochs=epochs, batch=batch, lr=lr, seed=seed + 1,
                   set_normalisation=True)
    return m, loss
