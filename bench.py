"""Benchmark of the vertex-patch Schwarz smoother hot path on one MI355X.

Workload: 3D Poisson, SIPG DG Q3 (64 DoFs/cell) on 64^3 Cartesian cells
(16.8M DoFs), multiplicative vertex-patch smoother (2x2x2-cell patches, 512
DoFs/patch, 16 red-black colors), seeded U[-1,1] right-hand side (synthetic).

A "step" is one SchwarzSmoother::step (Richardson: residual + one full
multiplicative Schwarz sweep over all patches), inputs resident in HBM.
value = patch-DoFs/s of the FP64 smoother (the reference's arithmetic; the
FP32 and BF16-MFMA smoothers are reported beside it), plus the
GMG-preconditioned CG time to a 1e-8 relative residual for the FP64 and the
mixed V-cycles (outer CG, residual and coarse solve in FP64).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published
# measured HBM bytes per launch of the fused color pass (rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, each scaled by what the counter reads for a copy kernel of known size),
# written by the profiling recipe in that file's header
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "mvp_fused_pass_64cubed_q3_hbm_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps first: the clocks settle within about 10 ms of work")
    ap.add_argument("--n", type=int, default=64, help="cells per direction")
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gmg", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the distorted-mesh GMG-CG with learned vs fixed damping")
    ap.add_argument("--no-fit", action="store_true",
                    help="skip the 48^3 Q4 distorted GMG-CG with surrogate vs fitted local solvers")
    ap.add_argument("--no-elasticity", action="store_true",
                    help="skip the 3D DG Q2 linear elasticity smoother and GMG-CG")
    ap.add_argument("--variants", default="f64,f32,bf16",
                    help="smoother variants to time besides f32 (comma list of f64, bf16)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="N",
                    help="only run 5 FP64 -> FP32 conversions of N values (8N bytes read, 4N "
                         "written): the reference kernel for scaling HBM counter readings")
    ap.add_argument("--write-traffic", nargs=4, metavar=("FETCH", "WRITE", "CAL_FETCH", "CAL_WRITE"),
                    help="only turn four rocprofv3 --pmc counter CSVs (FETCH_SIZE, WRITE_SIZE of "
                         "'bench.py --steps 2 --warmup 1 --no-cpu --no-gmg --no-elasticity "
                         "--variants f64,f32' and of '--calibrate-copy N') into the traffic file")
    ap.add_argument("--operators", action="store_true",
                    help="only time the level operators the smoothers' residuals use: SIPG "
                         "on the 64^3 Q3 mesh with interior vertices moved by 0.2 h and the "
                         "64^3 Q2 elasticity operator, FP32 and FP64, 10 applies each")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-cores CPU run (0: the CPUs this process may use, "
                         "capped by OMP_NUM_THREADS)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the all-cores CPU run")
    ap.add_argument("--cpu-steps-1core", type=int, default=1, help="steps of the 1-core CPU run")
    ap.add_argument("--no-cpu-gmg", action="store_true",
                    help="skip the CPU GMG-CG solve of the baseline (about 25 s on 16 cores)")
    return ap.parse_args()


def color_pass_bytes(sm, n, ndc, s_v, first_reads_outer=True):
    """Compulsory HBM bytes of one fused color pass, averaged over the color
    sequence of a sweep: read b and write x on the cells of the color's
    patches, read every outer face-neighbour cell of the color once (none in
    the first pass of a sweep that starts from x = 0)."""
    V = sm.patch_vertices()
    cp = sm.color_ptr()
    dim = V.shape[1]
    per_color = []
    for c in range(sm.n_colors):
        vs = V[cp[c]:cp[c + 1]]
        cells = set()
        for off in np.ndindex(*([2] * dim)):
            cc = vs - 1 + np.array(off)
            cells.update((cc[:, 0] + n * cc[:, 1] + (n * n * cc[:, 2] if dim == 3 else 0)).tolist())
        arr = np.array(sorted(cells))
        coords = np.stack([(arr // n ** d) % n for d in range(dim)], axis=1)
        outer = set()
        for d in range(dim):
            for sgn in (-1, 1):
                nb = coords.copy()
                nb[:, d] += sgn
                ok = (nb[:, d] >= 0) & (nb[:, d] < n)
                outer.update(sum(nb[ok, e] * n ** e for e in range(dim)).tolist())
        outer -= cells
        per_color.append((2 * len(cells) * ndc * s_v, len(outer) * ndc * s_v))
    seq = list(range(sm.n_colors))
    total = sum(own + (outer if (k > 0 or first_reads_outer) else 0.0)
                for k, (own, outer) in enumerate(per_color[c] for c in seq))
    return total / len(seq)


def write_traffic(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, n=64, p=3):
    """HBM bytes per launch of the fused color pass: FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch, mean over the dispatches of the kernel), each scaled by
    what the counter reads for the copy kernel k_convert<float, double> (8N bytes
    read, 4N written; N from its grid size)."""
    import csv

    def means(path):
        acc = {}
        for r in csv.DictReader(open(path)):
            acc.setdefault(r["Kernel_Name"], []).append((float(r["Counter_Value"]),
                                                         int(r["Grid_Size"])))
        return acc

    cf, cw = means(cal_fetch_csv), means(cal_write_csv)
    copy = next(k for k in cf if "k_convert<float, double>" in k)
    n_copy = max(g for _, g in cf[copy])
    fetch_factor = 8.0 * n_copy / 1024.0 / np.mean([v for v, _ in cf[copy]])
    write_factor = 4.0 * n_copy / 1024.0 / np.mean([v for v, _ in cw[copy]])
    f, w = means(fetch_csv), means(write_csv)
    out = {"fetch_factor": fetch_factor, "write_factor": write_factor,
           "calibration": f"k_convert<float, double>, N = {n_copy}"}
    for name, tag in (("f64", "double, double, 0>"), ("f32", "float, float, 0>")):
        k = next((k for k in f if "k_patch_mvp_fused<3, 8, " + tag in k), None)
        if k is None:
            continue
        fk, wk = float(np.mean([v for v, _ in f[k]])), float(np.mean([v for v, _ in w[k]]))
        out[f"{name}_{n}_{p}"] = {"kernel": k.split("(")[0], "dispatches": len(f[k]),
                                  "fetch_kib": fk, "write_kib": wk,
                                  "bytes_per_launch": 1024.0 * (fetch_factor * fk + write_factor * wk)}
    with open(TRAFFIC_FILE, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def operators_section(t, n=64, reps=10):
    """y = A x for the moved-vertex SIPG operator (3D Q3, 0.2 h, ip 4) and the
    elasticity operator (3D Q2, mu 1, lambda 10, ip 10) on n^3 cells"""
    from tpss_amd import learned_damping as ld
    V = ld.distorted_vertices(3, n, 0.2, seed=3)
    out = {}
    for name, prec in (("f32", t.F32), ("f64", t.F64)):
        for op, kw in (("moved_vertex_q3", dict(degree=3, ip_factor=4.0, vertices=V)),
                       ("elasticity_q2", dict(degree=2, ip_factor=10.0, elasticity=(1.0, 10.0)))):
            lv = t.Level(3, kw.pop("degree"), n, precision=prec, **kw)
            x = t.DeviceVector(lv.n_dofs, prec)
            y = t.DeviceVector(lv.n_dofs, prec)
            t.fill_random(x, 1)
            for _ in range(reps):  # warm-up: the clocks settle within about 10 ms of work
                lv.vmult(y, x)
            t.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                lv.vmult(y, x)
            t.synchronize()
            dt = (time.perf_counter() - t0) / reps
            s_v = 8 if prec == t.F64 else 4
            out[f"{op}_{name}"] = {"ms_per_apply": 1e3 * dt, "n_dofs": lv.n_dofs,
                                   "read_write_gbs": 2 * lv.n_dofs * s_v / dt / 1e9}
            del lv, x, y
    return out


def host_cpus():
    """(CPUs this process may use, CPUs of the host, CPU model name)"""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return usable, os.cpu_count() or usable, model


def time_smoother(t, lv, sm, steps, warmup, precision):
    b = t.DeviceVector(lv.n_dofs, precision)
    t.fill_random(b, 2)
    x = t.DeviceVector.zeros(lv.n_dofs, precision)
    for _ in range(warmup):
        sm.step(x, b)
    t.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sm.step(x, b)
    t.synchronize()
    t1 = time.perf_counter()
    # per-kernel averages from the section timers (HIP events around every
    # launch on the launch stream) in a second series of steps, so that the
    # events do not add to the step time above
    sm.set_timing(True)
    for _ in range(steps):
        sm.step(x, b)
    t.synchronize()
    sections = {name: (secs, calls) for name, secs, calls in sm.get_time_data()}
    sm.set_timing(False)
    return (t1 - t0) / steps, sections


def elasticity_section(t, args, n=64, p=2, mu=1.0, lam=10.0, ip=10.0):
    """3D linear elasticity, SIPG DG Q2 (3 x 27 DoFs per cell) on n^3 cells, Lame mu = 1,
    lambda = 10, ip 10: block-diagonal vertex-patch smoother steps (648 patch-DoFs) and
    GMG-CG to 1e-8 (FP64 and FP32 V-cycle), seeded U[-1,1] body-force vector."""
    out = {"workload": f"3D SIPG DG Q{p} linear elasticity {n}^3 cells, mu {mu}, lambda {lam}, "
                       f"ip {ip}, block-diagonal vertex patches"}
    steps = max(2, args.steps // 4)
    for name, variant, relax, prec, lprec in (
            ("mvp_f32", t.MULTIPLICATIVE, 1.0, t.F32, t.F32),
            ("avp_f32", t.ADDITIVE, 0.15, t.F32, t.F32),
            ("mvp_f64", t.MULTIPLICATIVE, 1.0, t.F64, t.F64),
            ("mvp_bf16_mfma", t.MULTIPLICATIVE, 1.0, t.F32, t.BF16)):
        lv = t.Level(3, p, n, ip_factor=ip, precision=prec, elasticity=(mu, lam))
        sm = t.SchwarzSmoother(lv, variant=variant, relaxation=relax, local_precision=lprec)
        sec, sections = time_smoother(t, lv, sm, steps, 1, prec)
        apply_t, apply_n = sections["[SchwarzPrecond] Apply inverses:"]
        res_t, res_n = sections["[SchwarzPrecond] Update residuals:"]
        pd = sm.n_patches * sm.n_dofs_per_patch
        out[name] = {"ms_per_step": 1e3 * sec, "patch_dofs_per_s": pd / sec,
                     "apply_inverses_ms_per_step": 1e3 * apply_t / steps,
                     "update_residuals_ms_per_step": 1e3 * res_t / steps,
                     "n_dofs": lv.n_dofs, "n_patches": sm.n_patches,
                     "patch_dofs": sm.n_dofs_per_patch}
        del sm, lv
    if not args.no_gmg:
        # the reference damping (omega = 1) and the learned one (the shipped
        # elasticity network: one damping for the Cartesian patches, its deviation
        # from 1 doubled)
        from tpss_amd import learned_damping as ld
        mlp = ld.load_elasticity()
        out["learned_damping"] = float(mlp.forward(np.ones((1, 24), dtype=np.float32))[0])
        for name, mp, net, lp in (("gmg_cg_f64", t.F64, None, None),
                                  ("gmg_cg_mixed_f32", t.F32, None, None),
                                  ("gmg_cg_mixed_bf16_mfma", t.F32, None, t.BF16),
                                  ("gmg_cg_f64_learned", t.F64, mlp, None),
                                  ("gmg_cg_mixed_f32_learned", t.F32, mlp, None)):
            g = t.GMGSolver(3, p, 2, int(np.log2(n // 2)) + 1, ip_factor=ip, mg_precision=mp,
                            variant=t.MULTIPLICATIVE, relaxation=1.0, elasticity=(mu, lam),
                            local_precision=lp)
            if net is not None:
                ld.apply_to_gmg(g, net)
            b = t.DeviceVector(g.n_dofs)
            t.fill_random(b, 2)
            x = t.DeviceVector.zeros(g.n_dofs)
            g.solve(x, b, rel_tol=1e-8, max_iter=100, raise_on_fail=False)  # warm-up solve
            info, hist = g.solve(x, b, rel_tol=1e-8, max_iter=100, raise_on_fail=False)
            out[name] = {"seconds": info.seconds, "iterations": info.iterations,
                         "n_frac": info.n_frac, "reduction_rate": info.reduction_rate,
                         "final_rel_residual": hist[-1] / hist[0]}
            del g
    return out


def fit_section(t, n=48, p=4, n_root=3, amplitude=0.15):
    """3D SIPG DG Q4 on 48^3 cells whose 3^3 root mesh is distorted by 0.15 h_root
    and refined (the reference's construction, so the GMG levels are nested),
    ip x4: mixed-precision GMG-CG to 1e-8 with surrogate-length and with fitted
    (separable Kronecker fit of the exact local matrices) patch solvers, for the
    multiplicative smoother and the additive one with the reference's 3D damping."""
    from tpss_amd import meshes
    V = meshes.nested_distorted_vertices(3, n_root, n, amplitude, seed=3)
    levels = int(round(np.log2(n // n_root))) + 1
    out = {"workload": f"3D SIPG DG Q{p} {n}^3 cells, {n_root}^3 root mesh distorted by "
                       f"{amplitude} h_root and refined, mixed-precision GMG-CG to 1e-8"}
    for ls_name, ls in (("surrogate", t.LOCAL_SURROGATE), ("fit", t.LOCAL_FIT)):
        for sm_name, var, relax in (("mvp", t.MULTIPLICATIVE, 1.0), ("avp_0.15", t.ADDITIVE, 0.15)):
            t.synchronize()
            t0 = time.perf_counter()
            g = t.GMGSolver(3, p, n_root, levels, ip_factor=4.0, vertices=V, mg_precision=t.F32,
                            variant=var, relaxation=relax, local_solver=ls)
            t.synchronize()
            setup = time.perf_counter() - t0
            b = t.DeviceVector(g.n_dofs)
            t.fill_random(b, 2)
            x = t.DeviceVector.zeros(g.n_dofs)
            info, hist = g.solve(x, b, rel_tol=1e-8, max_iter=200, raise_on_fail=False)
            out[f"{ls_name}_{sm_name}"] = {"setup_seconds": setup, "seconds": info.seconds,
                                           "iterations": info.iterations,
                                           "reduction_rate": info.reduction_rate}
            del g
    return out


def cpu_baseline(n, p, args):
    """The C oracle (FP64 restatement of the reference smoother, full residual per
    color, OpenMP over the patches of a color) on the same level, protocol of
    apps/poisson_mpiscaling.cc:409-471 (setup timed apart, then timed steps on a
    seeded vector) with a bounded sample: cpu_steps steps on all cores, then
    cpu_steps_1core steps on one core, instead of 10 samples x 20 steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as o
    usable, total, model = host_cpus()
    threads = args.cpu_threads
    if threads <= 0:
        threads = usable
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    out = {}
    for label, th, steps in (("all_cores", threads, args.cpu_steps),
                             ("one_core", 1, args.cpu_steps_1core)):
        t0 = time.perf_counter()
        sm = o.Smoother(3, p, n, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK, threads=th)
        setup = time.perf_counter() - t0
        b = o.splitmix_uniform(2, sm.N)
        x = np.zeros(sm.N)
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            sm.step(x, b)
            times.append(time.perf_counter() - t0)
        med = float(np.median(times))
        out[label] = {"threads": th, "steps": steps, "setup_seconds": setup,
                      "median_step_seconds": med, "min_step_seconds": float(min(times)),
                      "patch_dofs_per_s": sm.P * sm.ndp / med}
        del sm
    if not args.no_cpu_gmg:
        # the same GMG-CG to 1e-8 as the GPU's FP64 path (6 levels, 2^3 coarse cells,
        # MVP pre- and post-smoothing with the full residual per color), all cores
        t0 = time.perf_counter()
        g = o.GMG(3, p, 2, int(np.log2(n // 2)) + 1, threads=threads)
        setup = time.perf_counter() - t0
        b = o.splitmix_uniform(2, g.N)
        t0 = time.perf_counter()
        _, it, hist = g.cg(b, rel_tol=1e-8)
        out["gmg_cg"] = {"threads": threads, "setup_seconds": setup,
                         "seconds": time.perf_counter() - t0, "iterations": int(it),
                         "final_rel_residual": float(hist[-1] / hist[0])}
        del g
    out["host"] = {"usable_cpus": usable, "host_cpus": total, "model": model}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the smoother path is single-GPU; see the sharding design note).", file=sys.stderr)
        sys.exit(2)
    if args.write_traffic:
        write_traffic(*args.write_traffic, n=args.n, p=args.degree)
        return
    import tpss_amd as t

    if args.operators:
        print(json.dumps({"operators": operators_section(t, n=args.n)}))
        return

    if args.calibrate_copy:
        import ctypes
        src = t.DeviceVector(args.calibrate_copy, t.F64)
        dst = t.DeviceVector(args.calibrate_copy, t.F32)
        t.fill_random(src, 1)
        for _ in range(5):
            t._check(t.lib.tpss_vector_convert(dst._p, t.F32, src._p, t.F64,
                                               ctypes.c_int64(args.calibrate_copy), None))
        t.synchronize()
        print(json.dumps({"calibration": "k_convert<float, double>", "n": args.calibrate_copy,
                          "read_bytes": 8 * args.calibrate_copy,
                          "write_bytes": 4 * args.calibrate_copy, "launches": 5}))
        return

    n, p, dim = args.n, args.degree, 3
    ndc = (p + 1) ** dim
    results = {}
    wanted = set(args.variants.split(","))
    for name, prec, lprec in (("f64", t.F64, t.F64), ("f32", t.F32, t.F32),
                              ("bf16", t.F32, t.BF16)):
        if name != "f64" and name not in wanted:  # the FP64 variant is the headline value
            continue
        lv = t.Level(dim, p, n, precision=prec)
        sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE, coloring=t.COLOR_REDBLACK,
                               relaxation=1.0, local_precision=lprec)
        sec, sections = time_smoother(t, lv, sm, args.steps, args.warmup, prec)
        s_v = 8 if prec == t.F64 else 4
        pd = sm.n_patches * sm.n_dofs_per_patch
        apply_t, apply_n = sections["[SchwarzPrecond] Apply inverses:"]
        res_t, res_n = sections["[SchwarzPrecond] Update residuals:"]
        # FP64/FP32: residual-free color passes on x (every pass reads the outer
        # neighbour cells); BF16: Richardson residual b - A x (read x, b, write r:
        # 3 N s_v), then color passes on the correction, which starts from zero
        fused = name != "bf16"
        patch_bytes = color_pass_bytes(sm, n, ndc, s_v, first_reads_outer=fused)
        res_bytes = 3.0 * lv.n_dofs * s_v
        kernels = {
            "color_pass": {"avg_ms": 1e3 * apply_t / max(apply_n, 1), "launches": apply_n,
                           "bytes": patch_bytes},
            "residual": {"avg_ms": 1e3 * res_t / max(res_n, 1), "launches": res_n,
                         "bytes": res_bytes},
        }
        for k in kernels.values():
            k["gbs"] = k["bytes"] / (k["avg_ms"] * 1e-3) / 1e9 if k["avg_ms"] > 0 else 0.0
        dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
        results[name] = {"ms_per_step": 1e3 * sec, "patch_dofs_per_s": pd / sec,
                         "kernels": kernels, "dominant": dom, "n_patches": sm.n_patches,
                         "n_dofs": lv.n_dofs}
        del sm, lv

    gmg = {}
    if not args.no_gmg:
        for name, mp, lp in (("f64", t.F64, t.F64), ("mixed_f32", t.F32, t.F32),
                             ("mixed_bf16", t.F32, t.BF16)):
            g = t.GMGSolver(dim, p, 2, int(np.log2(n // 2)) + 1, mg_precision=mp,
                            variant=t.MULTIPLICATIVE, coloring=t.COLOR_REDBLACK, relaxation=1.0,
                            local_precision=lp)
            b = t.DeviceVector(g.n_dofs)
            t.fill_random(b, 2)
            x = t.DeviceVector.zeros(g.n_dofs)
            g.solve(x, b, rel_tol=1e-8)  # warm-up solve
            info, hist = g.solve(x, b, rel_tol=1e-8)
            gmg[name] = {"seconds": info.seconds, "iterations": info.iterations,
                         "n_frac": info.n_frac, "reduction_rate": info.reduction_rate,
                         "final_rel_residual": hist[-1] / hist[0]}
            del g

    learned = {}
    model = os.path.join(ROOT, "tpss_amd", "models", "damping_dgq3_3d.npy")
    if not args.no_gmg and not args.no_learned and os.path.exists(model):
        # learned damping: same level with interior vertices moved by 0.2 h (seed 3),
        # ip_factor x4 as the reference does on distorted meshes; the network's
        # deviation gain was calibrated on other meshes (seeds 201, 202); the best
        # constant damping swept on this mesh (0.7) for comparison
        from tpss_amd import learned_damping as ld
        V = ld.distorted_vertices(dim, n, 0.2, seed=3)
        mlp = ld.load(model)
        learned["damping_deviation_gain"] = ld.DEVIATION_GAIN
        for name, mp, use, relax in (("f64_fixed", t.F64, False, 1.0),
                                     ("f64_learned", t.F64, True, 1.0),
                                     ("mixed_f32_fixed", t.F32, False, 1.0),
                                     ("mixed_f32_fixed_0.7", t.F32, False, 0.7),
                                     ("mixed_f32_learned", t.F32, True, 1.0)):
            g = t.GMGSolver(dim, p, 2, int(np.log2(n // 2)) + 1, ip_factor=4.0, vertices=V,
                            mg_precision=mp, variant=t.MULTIPLICATIVE,
                            coloring=t.COLOR_REDBLACK, relaxation=relax)
            if use:
                ld.apply_to_gmg(g, mlp)
            b = t.DeviceVector(g.n_dofs)
            t.fill_random(b, 2)
            x = t.DeviceVector.zeros(g.n_dofs)
            info, hist = g.solve(x, b, rel_tol=1e-8, max_iter=100, raise_on_fail=False)
            learned[name] = {"seconds": info.seconds, "iterations": info.iterations,
                             "n_frac": info.n_frac, "reduction_rate": info.reduction_rate,
                             "final_rel_residual": hist[-1] / hist[0]}
            del g

    fitted = {}
    if not args.no_gmg and not args.no_fit:
        fitted = fit_section(t)

    elast = {}
    if not args.no_elasticity:
        elast = elasticity_section(t, args)

    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline(n, p, args)
        a1, c1, hst = cb["all_cores"], cb["one_core"], cb["host"]
        cpu = {"value": a1["patch_dofs_per_s"], "unit": "patch-DoFs/s", "cores": a1["threads"],
               "kind": "port",
               "sample": f"FP64 MVP smoothing steps (full residual per color, as the reference) of "
                         f"the C oracle on the same {n}^3 Q{p} level: median of {a1['steps']} "
                         f"steps on {a1['threads']} threads ({a1['median_step_seconds']:.2f} "
                         f"s/step) and of {c1['steps']} on 1 core "
                         f"({c1['median_step_seconds']:.2f} s/step); capped from 10 samples x "
                         f"20 steps to keep the run within minutes",
               "one_core": c1["patch_dofs_per_s"], "runs": cb,
               "host": f"{hst['model']}, {hst['usable_cpus']} usable of {hst['host_cpus']} CPUs"}
        if "gmg_cg" in cb:
            gc = cb["gmg_cg"]
            cpu["gmg_cg_seconds"] = gc["seconds"]
            cpu["sample"] += (f"; GMG-CG to 1e-8 (FP64, same hierarchy and smoother): "
                              f"{gc['iterations']} iterations in {gc['seconds']:.2f} s on "
                              f"{gc['threads']} threads")

    main_v = results["f64"]
    dom = main_v["dominant"]
    k = main_v["kernels"][dom]
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and dom == "color_pass":
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get(f"f64_{n}_{p}", {}).get("bytes_per_launch")
    out = {
        "metric": "smoother patch-DoFs/s + GMG-CG time-to-1e-8: mixed/learned vs FP64, 1 MI355X",
        "value": main_v["patch_dofs_per_s"],
        "unit": "patch-DoFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_v["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1] right-hand side, Cartesian unit cube)",
        "config": {"workload": f"3D SIPG DG Q{p} Poisson {n}^3 cells, multiplicative vertex-patch "
                               f"Schwarz smoother step (16 red-black colors, 512 DoFs/patch), "
                               f"FP64 vectors and patch solves",
                   "n_dofs": main_v["n_dofs"], "n_patches": main_v["n_patches"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": k["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": k["gbs"] / HBM_PEAK_GBS,
                     "algorithmic_bytes_per_launch": k["bytes"], "traffic": traffic},
        "cpu_baseline": cpu,
        "variants": {name: {"patch_dofs_per_s": r["patch_dofs_per_s"],
                            "ms_per_step": r["ms_per_step"],
                            "kernels": r["kernels"]} for name, r in results.items()},
        "gmg_cg": gmg,
        "gmg_cg_distorted_learned_damping": learned,
        "gmg_cg_distorted_fitted_solvers": fitted,
        "elasticity": elast,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
