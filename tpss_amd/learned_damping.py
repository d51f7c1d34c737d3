"""Learned per-patch damping: training-data generation, training and use in GMG.

Training data: on seeded distorted meshes (interior vertices moved by
amplitude*h*U[-1,1]^dim) every vertex patch contributes its features (the
singular values of its cell Jacobians) and the optimal local damping that the
smoother computes for it with its own local solvers and operator
(SchwarzPreconditioner.optimal_damping).  The network is trained on the
device by the library's own forward/backward/Adam kernels.

    python -m tpss_amd.learned_damping --out tpss_amd/models/damping_dgq3_3d.npy
    python -m tpss_amd.learned_damping --out tpss_amd/models/damping_dgq3_3d.npy --calibrate
"""
import argparse
import os
import time

import numpy as np

from . import (F32, F64, MULTIPLICATIVE, COLOR_REDBLACK, DampingMLP, DeviceVector, GMGSolver,
               Level, SchwarzSmoother, fill_random)
from .meshes import distorted_vertices  # noqa: F401 -- the training meshes

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "damping_dgq3_3d.npy")


def patch_samples(dim, degree, n, amplitude, seed, ip_factor=4.0, n_samples=4,
                  omega_min=0.2, omega_max=1.2):
    """(features, optimal damping) of every patch of one distorted level."""
    V = distorted_vertices(dim, n, amplitude, seed)
    lv = Level(dim, degree, n, ip_factor=ip_factor, vertices=V, precision=F64)
    sm = SchwarzSmoother(lv, variant=MULTIPLICATIVE, coloring=COLOR_REDBLACK, relaxation=1.0)
    X = sm.patch_features()
    y = sm.optimal_damping(n_samples=n_samples, seed=1000 + seed, omega_min=omega_min,
                           omega_max=omega_max)
    return X, y


# Seven 32^3 meshes at 0.2 h (seeds 101..107), four random errors per patch in
# the target, 40 epochs: the settings the shipped model was trained with.  A
# range of amplitudes (--amplitudes 0.05,0.08,...,0.2) is accepted, but a model
# trained on it damped the strongly distorted 64^3 test mesh worse (49 CG
# iterations against 14).  Amplitudes stop at 0.2 h: from 0.25 h on, perturbed
# trilinear cells of a 32^3 mesh have corners with negative Jacobian determinant.
TRAIN_AMPLITUDES = (0.2,) * 7


def generate(dim=3, degree=3, n=32, amplitudes=TRAIN_AMPLITUDES, seed0=101, n_samples=4):
    """Training set: one distorted mesh per entry of amplitudes (seeds seed0,
    seed0 + 1, ...)."""
    Xs, ys = [], []
    for i, a in enumerate(amplitudes):
        X, y = patch_samples(dim, degree, n, a, seed0 + i, n_samples=n_samples)
        Xs.append(X)
        ys.append(y)
    return np.concatenate(Xs), np.concatenate(ys)


def train(X, y, epochs=40, batch=2048, lr=2e-3, seed=4, hidden=64):
    mlp = DampingMLP(X.shape[1], hidden, 0.2, 1.2, seed=seed)
    loss0, _ = mlp.loss_grad(X[:8192], y[:8192])
    loss = mlp.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, loss0, loss


# The network predicts the locally optimal damping of each patch (mean 0.90 on
# 0.2 h meshes).  As a multigrid smoother the V-cycle contracts faster at a
# lower level of damping, so the output range is scaled by one factor
# calibrated on a training mesh (calibrate_scale: 64^3 Q3, 0.2 h, seed 201,
# mixed-precision GMG-CG; the factor with the smallest CG reduction rate over
# 0.60, 0.64, ..., 0.88).  scale=1 gives the network's own prediction.
OUTPUT_SCALE = 0.72


def load(path=DEFAULT_MODEL, n_in=24, hidden=64, scale=OUTPUT_SCALE):
    """The damping network with its output range [0.2, 1.2] scaled by scale."""
    return DampingMLP(n_in, hidden, 0.2 * scale, 1.2 * scale).load(path)


def calibrate_scale(mlp, n=64, degree=3, amplitude=0.2, seed=201,
                    scales=tuple(0.60 + 0.04 * i for i in range(8)), rel_tol=1e-8):
    """Output scale of an (unscaled) damping network with the smallest GMG-CG
    reduction rate on one distorted training mesh (mixed precision, MVP with
    red-black colors): returns (best scale, {scale: (iterations, rate)})."""
    V = distorted_vertices(3, n, amplitude, seed)
    levels = int(np.log2(n // 2)) + 1
    res = {}
    for s in scales:
        g = GMGSolver(3, degree, 2, levels, ip_factor=4.0, vertices=V, mg_precision=F32,
                      variant=MULTIPLICATIVE, coloring=COLOR_REDBLACK, relaxation=1.0)
        for level in range(1, levels):
            for post in (False, True):
                sm = g.smoother(level, post)
                sm.set_patch_damping(mlp.forward(sm.patch_features()).astype(np.float64) * s)
        b = DeviceVector(g.n_dofs)
        fill_random(b, 2)
        x = DeviceVector.zeros(g.n_dofs)
        info, _ = g.solve(x, b, rel_tol=rel_tol, max_iter=100, raise_on_fail=False)
        res[round(float(s), 4)] = (info.iterations, info.reduction_rate)
        del g
    best = min(res, key=lambda k: res[k][1])
    return best, res


def apply_to_gmg(gmg, mlp):
    """Per-patch damping from the network on every smoothed level (pre and post)."""
    for level in range(1, gmg.n_levels):
        for post in (False, True):
            gmg.smoother(level, post).set_damping_mlp(mlp)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--amplitudes", default=",".join(str(a) for a in TRAIN_AMPLITUDES))
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--samples", type=int, default=4,
                    help="random errors per patch in the optimal-damping target")
    ap.add_argument("--calibrate", action="store_true",
                    help="only calibrate the output scale of the model at --out")
    args = ap.parse_args()
    if args.calibrate:
        best, res = calibrate_scale(load(args.out, scale=1.0))
        for sc, (it, rate) in res.items():
            print(f"scale {sc:.2f}: {it} CG iterations, reduction rate {rate:.4f}", flush=True)
        print(f"calibrated output scale: {best:.2f}", flush=True)
        return
    t0 = time.time()
    amps = [float(a) for a in args.amplitudes.split(",")]
    X, y = generate(3, args.degree, args.n, amps, n_samples=args.samples)
    t1 = time.time()
    print(f"training data: {X.shape[0]} patches from {len(amps)} meshes in {t1 - t0:.1f} s; "
          f"target mean {y.mean():.4f} std {y.std():.4f} min {y.min():.3f} max {y.max():.3f}",
          flush=True)
    mlp, l0, l1 = train(X, y, epochs=args.epochs)
    print(f"training: loss {l0:.3e} -> {l1:.3e} in {time.time() - t1:.1f} s "
          f"(variance of the target {0.5 * y.var():.3e})", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    mlp.save(args.out)
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()
