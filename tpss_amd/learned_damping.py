"""Learned per-patch damping: training-data generation, training and use in GMG.

Training data: on seeded distorted meshes every vertex patch contributes its
features (the singular values of its cell Jacobians) and the optimal local
damping that the smoother computes for it with its own local solvers and
operator (SchwarzPreconditioner.optimal_damping: the A-norm line search along
the patch correction, exactly 1 where the local solver is the exact patch
inverse).  The network is trained on the device by the library's own
forward/backward/Adam kernels.

Use as a smoother damping: the line search is the best damping of one patch
correction on its own; inside the V-cycle the multiplicative smoother does
better with the deviation from 1 amplified, omega_j = 1 - g (1 - w_j), w_j the
network's prediction, g one gain calibrated on training meshes.  Patches whose
local solver is exact keep omega = 1 whatever g is.

    python -m tpss_amd.learned_damping --out tpss_amd/models/damping_dgq3_3d.npy
    python -m tpss_amd.learned_damping --out tpss_amd/models/damping_dgq3_3d.npy --calibrate
    python -m tpss_amd.learned_damping --elasticity
    python -m tpss_amd.learned_damping --elasticity --calibrate
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

#: output range of the network: the line-search damping, clamped
DAMPING_RANGE = (0.6, 1.1)


def patch_samples(dim, degree, n, amplitude, seed, pattern="uniform", ip_factor=4.0,
                  n_samples=4):
    """(features, optimal damping) of every patch of one distorted level."""
    V = distorted_vertices(dim, n, amplitude, seed, pattern=pattern)
    lv = Level(dim, degree, n, ip_factor=ip_factor, vertices=V, precision=F64)
    sm = SchwarzSmoother(lv, variant=MULTIPLICATIVE, coloring=COLOR_REDBLACK, relaxation=1.0)
    X = sm.patch_features()
    y = sm.optimal_damping(n_samples=n_samples, seed=1000 + seed, omega_min=DAMPING_RANGE[0],
                           omega_max=DAMPING_RANGE[1])
    return X, y


# Seven 32^3 Q3 meshes at 0.2 h with the three distortion patterns (the whole
# mesh, growing with x, only x > 1/2), four random errors per patch in the
# target.  Amplitudes stop at 0.2 h: from 0.25 h on, perturbed trilinear cells of
# a 32^3 mesh can have corners with negative Jacobian determinant.
TRAIN_MESHES = ((0.2, "uniform", 101), (0.2, "uniform", 102), (0.2, "graded", 103),
                (0.2, "graded", 104), (0.2, "half", 105), (0.2, "half", 106),
                (0.2, "uniform", 107))


def generate(dim=3, degree=3, n=32, meshes=TRAIN_MESHES, n_samples=4):
    """Training set: the patches of every (amplitude, pattern, seed) mesh."""
    Xs, ys = [], []
    for amp, pattern, seed in meshes:
        X, y = patch_samples(dim, degree, n, amp, seed, pattern=pattern, n_samples=n_samples)
        Xs.append(X)
        ys.append(y)
    return np.concatenate(Xs), np.concatenate(ys)


def train(X, y, epochs=40, batch=2048, lr=2e-3, seed=4, hidden=64):
    mlp = DampingMLP(X.shape[1], hidden, DAMPING_RANGE[0], DAMPING_RANGE[1], seed=seed)
    loss0, _ = mlp.loss_grad(X[:8192], y[:8192])
    loss = mlp.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, loss0, loss


# Gain of the deviation from 1 (omega = 1 - g (1 - w)), calibrated on two 64^3
# Q3 meshes that are neither training nor evaluation meshes (0.2 h, seeds 201
# "half" and 202 "uniform", mixed-precision GMG-CG to 1e-8): the gain with the
# smallest mean CG reduction rate over 1.5, 1.75, 2.0, 2.25, 2.5, 3.0, 3.5
# (calibrate_gain).
DEVIATION_GAIN = 3.0


def load(path=DEFAULT_MODEL, n_in=24, hidden=64, gain=DEVIATION_GAIN):
    """The damping network with omega = 1 - gain (1 - w) applied to its
    prediction w: the output range DAMPING_RANGE mapped affinely."""
    lo, hi = DAMPING_RANGE
    return DampingMLP(n_in, hidden, 1.0 - gain * (1.0 - lo), 1.0 - gain * (1.0 - hi)).load(path)


def apply_to_gmg(gmg, mlp):
    """Per-patch damping from the network on every smoothed level (pre and post)."""
    for level in range(1, gmg.n_levels):
        for post in (False, True):
            gmg.smoother(level, post).set_damping_mlp(mlp)


def gmg_solve(V, n, degree=3, mlp=None, relaxation=1.0, precision=F32, rel_tol=1e-8):
    """GMG-CG (multiplicative red-black smoother, ip 4, 2^3 coarse cells) on the
    distorted n^3 mesh V with the network's damping or a constant one:
    (iterations, reduction rate, seconds)"""
    levels = int(np.log2(n // 2)) + 1
    g = GMGSolver(3, degree, 2, levels, ip_factor=4.0, vertices=V, mg_precision=precision,
                  variant=MULTIPLICATIVE, coloring=COLOR_REDBLACK, relaxation=relaxation)
    if mlp is not None:
        apply_to_gmg(g, mlp)
    b = DeviceVector(g.n_dofs)
    fill_random(b, 2)
    x = DeviceVector.zeros(g.n_dofs)
    info, _ = g.solve(x, b, rel_tol=rel_tol, max_iter=100, raise_on_fail=False)
    return info.iterations, info.reduction_rate, info.seconds


def calibrate_gain(path=DEFAULT_MODEL, n=64, meshes=((0.2, "half", 201), (0.2, "uniform", 202)),
                   gains=(1.5, 1.75, 2.0, 2.25, 2.5, 3.0, 3.5)):
    """Gain with the smallest mean GMG-CG reduction rate on the calibration meshes:
    returns (best gain, {gain: [(iterations, rate) per mesh]})."""
    Vs = [distorted_vertices(3, n, a, s, pattern=pat) for a, pat, s in meshes]
    res = {}
    for g in gains:
        mlp = load(path, gain=g)
        res[g] = [gmg_solve(V, n, mlp=mlp)[:2] for V in Vs]
    best = min(res, key=lambda k: np.mean([r for _, r in res[k]]))
    return best, res


# Linear elasticity (3D SIPG DG Q2, Lame mu = 1, lambda = 10, ip 10, Cartesian
# meshes): the block-diagonal local solvers drop the lambda coupling of the
# components, so the line-search damping of a patch is below 1 (about 0.97).  The
# features of a Cartesian patch are all 1 (singular values scaled by n): the
# network learns one damping for every patch, from the patches of two training
# levels, and its deviation from 1 is amplified by a gain calibrated on a GMG-CG
# solve, as for the Poisson network.
ELASTICITY_MODEL = os.path.join(HERE, "models", "damping_elasticity_dgq2_3d.npy")
ELASTICITY = (1.0, 10.0)
ELASTICITY_IP = 10.0
ELASTICITY_TRAIN_SIZES = (32, 48)
# gain with the smallest GMG-CG reduction rate on the 32^3 level (mixed precision)
# over 1, 1.5, 2, 2.5, 3, 4 (calibrate_gain_elasticity)
ELASTICITY_GAIN = 2.0


def elasticity_samples(n, n_samples=4, seed=7):
    """(features, optimal damping) of every patch of the n^3 elasticity level."""
    lv = Level(3, 2, n, ip_factor=ELASTICITY_IP, precision=F64, elasticity=ELASTICITY)
    sm = SchwarzSmoother(lv, variant=MULTIPLICATIVE, coloring=COLOR_REDBLACK, relaxation=1.0)
    X = sm.patch_features()
    y = sm.optimal_damping(n_samples=n_samples, seed=seed, omega_min=DAMPING_RANGE[0],
                           omega_max=DAMPING_RANGE[1])
    return X, y


def generate_elasticity(sizes=ELASTICITY_TRAIN_SIZES, n_samples=4):
    Xs, ys = [], []
    for i, n in enumerate(sizes):
        X, y = elasticity_samples(n, n_samples=n_samples, seed=7 + 100 * i)
        Xs.append(X)
        ys.append(y)
    return np.concatenate(Xs), np.concatenate(ys)


def load_elasticity(path=ELASTICITY_MODEL, gain=ELASTICITY_GAIN):
    return load(path, gain=gain)


def gmg_solve_elasticity(n, mlp=None, relaxation=1.0, precision=F32, rel_tol=1e-8):
    """GMG-CG (multiplicative red-black block-diagonal smoother, 2^3 coarse cells)
    for the n^3 elasticity problem with the network's damping or a constant one:
    (iterations, reduction rate, seconds)"""
    levels = int(np.log2(n // 2)) + 1
    g = GMGSolver(3, 2, 2, levels, ip_factor=ELASTICITY_IP, mg_precision=precision,
                  variant=MULTIPLICATIVE, coloring=COLOR_REDBLACK, relaxation=relaxation,
                  elasticity=ELASTICITY)
    if mlp is not None:
        apply_to_gmg(g, mlp)
    b = DeviceVector(g.n_dofs)
    fill_random(b, 2)
    x = DeviceVector.zeros(g.n_dofs)
    info, _ = g.solve(x, b, rel_tol=rel_tol, max_iter=100, raise_on_fail=False)
    return info.iterations, info.reduction_rate, info.seconds


def calibrate_gain_elasticity(path=ELASTICITY_MODEL, n=32, gains=(1.0, 1.5, 2.0, 2.5, 3.0, 4.0)):
    res = {g: [gmg_solve_elasticity(n, mlp=load(path, gain=g))[:2]] for g in gains}
    best = min(res, key=lambda k: res[k][0][1])
    return best, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--samples", type=int, default=4,
                    help="random errors per patch in the optimal-damping target")
    ap.add_argument("--calibrate", action="store_true",
                    help="only calibrate the deviation gain of the model at --out")
    ap.add_argument("--elasticity", action="store_true",
                    help="the elasticity network (3D Q2, mu 1, lambda 10) instead of the "
                         "distorted-mesh Poisson one")
    args = ap.parse_args()
    if args.elasticity:
        out = args.out if args.out != DEFAULT_MODEL else ELASTICITY_MODEL
        if args.calibrate:
            best, res = calibrate_gain_elasticity(out)
            for g, runs in res.items():
                print(f"gain {g:.2f}: " + ", ".join(f"{it} CG iterations (rate {r:.4f})"
                                                   for it, r in runs), flush=True)
            print(f"calibrated deviation gain: {best:.2f}", flush=True)
            return
        t0 = time.time()
        X, y = generate_elasticity(n_samples=args.samples)
        t1 = time.time()
        print(f"training data: {X.shape[0]} patches from levels {ELASTICITY_TRAIN_SIZES} in "
              f"{t1 - t0:.1f} s; target mean {y.mean():.4f} std {y.std():.4f} min {y.min():.3f} "
              f"max {y.max():.3f}", flush=True)
        mlp, l0, l1 = train(X, y, epochs=args.epochs)
        print(f"training: loss {l0:.3e} -> {l1:.3e} in {time.time() - t1:.1f} s "
              f"(variance of the target {0.5 * y.var():.3e}); prediction "
              f"{float(mlp.forward(X[:1])[0]):.4f}", flush=True)
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        mlp.save(out)
        print("saved", out, flush=True)
        return
    if args.calibrate:
        best, res = calibrate_gain(args.out)
        for g, runs in res.items():
            print(f"gain {g:.2f}: " + ", ".join(f"{it} CG iterations (rate {r:.4f})"
                                               for it, r in runs), flush=True)
        print(f"calibrated deviation gain: {best:.2f}", flush=True)
        return
    t0 = time.time()
    X, y = generate(3, args.degree, args.n, n_samples=args.samples)
    t1 = time.time()
    print(f"training data: {X.shape[0]} patches from {len(TRAIN_MESHES)} meshes in "
          f"{t1 - t0:.1f} s; target mean {y.mean():.4f} std {y.std():.4f} min {y.min():.3f} "
          f"max {y.max():.3f}", flush=True)
    mlp, l0, l1 = train(X, y, epochs=args.epochs)
    print(f"training: loss {l0:.3e} -> {l1:.3e} in {time.time() - t1:.1f} s "
          f"(variance of the target {0.5 * y.var():.3e})", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    mlp.save(args.out)
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()
