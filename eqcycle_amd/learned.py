"""Learned initial guess workflow for the quasi-static solve (config 4 / 5).

1. ``record``: run the plain solver with snapshot recording on; each pass-1
   quasi-static solve stores (fault increment dF, interior increment dd).
2. ``fit_basis``: POD by the method of snapshots -- the Gram matrix S^T S is
   formed on the GPU, eigen-decomposed here (numpy), and Phi = S V / sigma is
   formed on the GPU; the training targets are sigma_k V_jk (Phi^T dd_j).
3. ``train``: Adam on the conv1d + MLP network (hand-written GPU kernels).
4. ``enable``: set eq_solver_opts.guess = learned; the previous-step guess stays
   the fallback (energy-optimal step along the learned correction).
5. ``fit_linear``: the linear least-squares map from the same input to the
   same coefficients (guess = linear), the bar the network has to beat: the
   quasi-static response is linear in dF, so a linear predictor in POD space
   is the natural competitor.
"""
from __future__ import annotations

import numpy as np


def record(ctx, total_time=None, max_steps=0, capacity=2000, nbasis=64, seed=4):
    ctx.learn_init(capacity=capacity, nbasis=nbasis, seed=seed)
    ctx.learn_record(True)
    out = ctx.run(total_time=total_time, max_steps=max_steps, snapshots=False)
    ctx.learn_record(False)
    return out


def fit_basis(ctx, nbasis=None, tol=1e-8, noise=1e-3):
    """POD of the usable snapshots.  A pass-1 solve to relative residual tol
    leaves an error of order tol |F| in the displacement, so an increment with
    |dF| < tol |F| / noise is mostly solver error; such snapshots are left out."""
    nb = nbasis or ctx.nbasis
    G = ctx.learn_gram()
    norms = ctx.learn_norms()
    use = (norms[:, 0] > (tol / noise) * norms[:, 1]) & (norms[:, 0] > 0)
    idx = np.nonzero(use)[0]
    w, Vs = np.linalg.eigh(G[np.ix_(idx, idx)])
    order = np.argsort(w)[::-1][:nb]
    w = np.maximum(w[order], 0.0)
    V = np.zeros((G.shape[0], nb))
    V[idx] = Vs[:, order]
    sigma = np.sqrt(w)
    ctx.learn_set_basis(np.ascontiguousarray(V), sigma, use.astype(np.int32))
    energy = w.sum() / max(np.trace(G[np.ix_(idx, idx)]), 1e-300)
    return dict(sigma=sigma, energy_captured=float(energy), used=int(use.sum()), total=int(len(use)), use=use)


def fit_linear(ctx, use):
    """Least squares [X 1] [A^T; b^T] = Y over the snapshots in use (X: the
    network's normalised inputs, Y: its targets), in FP64 on the host, handed
    to the device as the linear predictor.  Returns the fit's relative
    residual on the training set."""
    data = ctx.learn_get()
    X = data["X"][use].astype(np.float64)
    Y = data["Y"][use].astype(np.float64)
    Xa = np.hstack([X, np.ones((X.shape[0], 1))])
    beta, *_ = np.linalg.lstsq(Xa, Y, rcond=None)
    A, b = beta[:-1].T, beta[-1]
    ctx.learn_set_linear(A, b)
    res = np.linalg.norm(Xa @ beta - Y) / max(np.linalg.norm(Y), 1e-300)
    return dict(train_rel_residual=float(res), rows=int(X.shape[0]))


def train(ctx, epochs=300, batch=64, lr=1e-3, seed=11):
    return ctx.learn_train(epochs=epochs, batch=batch, lr=lr, seed=seed)


def enable(ctx, tol, pc="amg", prec="fp64"):
    ctx.set_solver(tol=tol, pc=pc, prec=prec, guess="learned")
