#!/usr/bin/env python3
"""Headline benchmark: matrix-free J.v of the Neo-Hookean p=3 form on one MI355X.

Workload (default): Functional<H1<3,3>(H1<3,3>)> with MaterialStressFunctor<NeoHookean>
(K=10, G=0.25), 64^3 hexes of [0,1]^3 with interior vertices perturbed by
U[-0.2h, 0.2h] (seed 42), 4^3 Gauss points, FP64.  Displacement
u = 0.01 sin(pi x)(1, 0.5, 0.25) + U[-a, a] with a = 1e-3 * 16 / n (smooth: no
inverted elements); direction U[-1, 1].
A step is one application of the gradient operator (J.v) to the resident
direction vector.  The residual and the linearization are timed alongside.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "GDOF/s matrix-free J·v (p=3 hex, FP64 vs FP32); Newton solve s to 1e-8"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (n, order, perturb, params, qfunction); d = minimal stored tangent per point
    "neohookean_p3_64": (64, 3, 0.2, (1.0, 10.0, 0.25), "neohookean"),
    "neohookean_p3_24": (24, 3, 0.2, (1.0, 10.0, 0.25), "neohookean"),  # quick runs of the same path
    "neohookean_p4_48": (48, 4, 0.2, (1.0, 10.0, 0.25), "neohookean"),
    "neohookean_p2_128": (128, 2, 0.0, (1.0, 10.0, 0.25), "neohookean"),
    "neohookean_p2_32": (32, 2, 0.0, (1.0, 10.0, 0.25), "neohookean"),  # quick runs of the p = 2 path
    # J2 with linear hardening: E, nu, sigma_y, Hi, Hk, density, geometric nonlinearity
    "j2_p2_96": (96, 2, 0.0, (200.0, 0.3, 0.5, 2.0, 0.0, 1.0, 1.0), "j2"),
    "j2_p2_32": (32, 2, 0.0, (200.0, 0.3, 0.5, 2.0, 0.0, 1.0, 1.0), "j2"),  # quick runs of the J2 path
}
TANGENT_ENTRIES = {"neohookean": 45, "j2": 81}


def unet_flops(nv):
    """Multiply-adds x 2 of one U-Net forward pass on an nv^3 vertex grid (3x3x3 convolutions only)."""
    import serac_amd as sx

    levels = [nv]
    for _ in range(3):
        levels.append((levels[-1] - 1) // 2 + 1)
    out_level = [0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0, 0]
    return sum(2 * 27 * (ca + cb) * co * levels[lv] ** 3 for (ca, cb, co), lv in zip(sx.UNet.LAYERS, out_level))


def unet_bench(sx, n, steps, warmup):
    """U-Net inference and training step (forward, backward, Adam) timed with device syncs on the host clock."""
    net = sx.UNet(n + 1, n + 1, n + 1, seed=1)
    rng = np.random.default_rng(0)
    x = sx.DeviceVector.from_numpy(rng.standard_normal(net.nvox * 3), dtype=np.float32)
    t = sx.DeviceVector.from_numpy(rng.standard_normal(net.nvox * 3), dtype=np.float32)
    y = sx.DeviceVector(net.nvox * 3, dtype=np.float32)
    reps = max(1, min(steps, 20))
    loss = ctypes_double()
    for _ in range(max(1, warmup)):
        sx.capi.call("sx_unet_forward", net.h, x.ptr, y.ptr)
    sx.capi.call("sx_device_synchronize")
    tf = time.perf_counter()
    for _ in range(reps):
        sx.capi.call("sx_unet_forward", net.h, x.ptr, y.ptr)
    sx.capi.call("sx_device_synchronize")
    t_fwd = (time.perf_counter() - tf) / reps
    for _ in range(max(1, warmup)):
        sx.capi.call("sx_unet_backward", net.h, x.ptr, t.ptr, loss)
        sx.capi.call("sx_unet_adam", net.h, 1e-4)
    sx.capi.call("sx_device_synchronize")
    tb = time.perf_counter()
    for _ in range(reps):
        sx.capi.call("sx_unet_backward", net.h, x.ptr, t.ptr, loss)
        sx.capi.call("sx_unet_adam", net.h, 1e-4)
    sx.capi.call("sx_device_synchronize")
    t_train = (time.perf_counter() - tb) / reps
    fl = unet_flops(n + 1)
    return {
        "grid": [n + 1] * 3,
        "params": net.nparam,
        "forward_ms": round(t_fwd * 1e3, 4),
        "train_step_ms": round(t_train * 1e3, 4),
        "forward_gflop": round(fl / 1e9, 3),
        "train_step_gflop": round(3 * fl / 1e9, 3),
        "forward_tflops": round(fl / t_fwd / 1e12, 2),
        "train_tflops": round(3 * fl / t_train / 1e12, 2),
        "flops_note": "2 x multiply-adds of the 3x3x3 convolutions; a training step counts 3x the forward",
        "kernels": "direct convolution on FP32 FMA units (BF16 weights and activations, FP32 accumulation)",
        "peak_tflops_fp32_vector": 157.3,
        "peak_tflops_bf16_mfma_dense": 2500.0,
    }


def ctypes_double():
    import ctypes

    return ctypes.byref(ctypes.c_double())


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# HBM traffic of the J.v kernel measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by plain kernels of known size); see the summary file for the commands.
PROFILED_TRAFFIC = {
    # (workload, scatter, XCD remap, scatter order)
    ("neohookean_p3_64", "atomic", 0, 2): "profiles/jvp_neohookean_p3_64.md",
}


def profiled_traffic(workload, scatter, remap, order):
    path = PROFILED_TRAFFIC.get((workload, scatter, remap, order))
    if not path or not os.path.exists(os.path.join(ROOT, path)):
        return None
    for line in open(os.path.join(ROOT, path)):
        if "traffic_bytes_per_launch:" in line:
            return int(line.split("traffic_bytes_per_launch:")[1].split()[0])
    return None


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, p, perturb, params, seconds, qf="neohookean"):
    """The oracle's J.v (the reference algorithm restated in C++) on an n^3
    mesh of the same workload (the workload's own mesh by default), timed on
    one host core and on all the host threads this job may use (OpenMP element
    ranges, each thread scattering into its own node band, allocated before
    the timed region): median of 5 batches each.  The oracle is loaded in its
    x86-64-v4 (AVX-512) build when the host has AVX-512."""
    import oracle as O
    from serac_amd import workloads as W

    isa = O.use_isa("widest")
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    m = O.Mesh.cuboid(n, n, n)
    if perturb > 0:
        m.perturb(42, perturb)
    s = O.Space(m, p, 3)
    f = O.Form(s)
    if qf == "j2":
        f.add_domain_integral_state(O.QF_J2_LINEAR, params, np.zeros(10))
        U = np.zeros(s.ndof)
        U[0::3] = 0.03 * s.node_coords()[:, 2]
    else:
        f.add_domain_integral(O.QF_NEOHOOKEAN, params)
        U = W.neohookean_displacement(s.node_coords(), noise=W.noise_for_mesh(n))
    dU = W.direction(s.nnodes, 3)
    f.linearize_threads(U, threads)

    def median_per_apply(nt, budget):
        t1 = f.time_jvp(dU, 1) if nt == 1 else f.time_jvp_threads(dU, 1, nt)
        reps = max(1, int(budget / 5 / max(t1, 1e-6)))
        times = [(f.time_jvp(dU, reps) if nt == 1 else f.time_jvp_threads(dU, reps, nt)) / reps for _ in range(5)]
        return float(np.median(times)), reps

    t_one, r_one = median_per_apply(1, seconds / 2)
    t_all, r_all = median_per_apply(threads, seconds / 2)
    return {
        "value": s.ndof / t_all / 1e9,
        "unit": "GDOF/s",
        "cores": threads,
        "kind": "port",
        "single_core_value": s.ndof / t_one / 1e9,
        "all_vs_one_core": round(t_one / t_all, 2),
        "cpu_model": host_cpu_model(),
        "nproc": os.cpu_count(),
        "isa": "x86-64-v4 (AVX-512)" if isa == "v4" else "x86-64-v3 (AVX2)",
        "sample": f"oracle J.v (stored 81-entry tangent, sum factorization), {n}^3 p={p} {qf}, {s.ndof} DOFs; "
                  f"median of 5 batches of {r_all} applications on {threads} threads (OpenMP element ranges, "
                  f"per-thread node bands allocated outside the timed region) and of {r_one} on one core",
    }


def load_path(space, args, qf):
    """Essential DOFs and their values after each load step.  Neo-Hookean: uniaxial
    compression (x = 0 clamped, the x = 1 face moved by -strain in x; y, z free there)
    in `load_steps` equal steps.  J2: cyclic shear of the z = 1 face against the
    clamped z = 0 face (u_x = gamma, u_y = u_z = 0 on top), gamma going 0 -> +g ->
    -g -> 0 in `load_steps` equal increments (g = 3 %: the steps are 4g / load_steps)."""
    L = args.load_steps
    if qf == "j2":
        bot = space.boundary_dofs(0b010000, 0x7)
        top = space.boundary_dofs(0b100000, 0x7)
        ess = np.union1d(bot, top).astype(np.int32)
        top_x = np.isin(ess, space.boundary_dofs(0b100000, 0x1))
        g, q = args.shear, L // 4
        gam = [g * k / q for k in range(1, q + 1)]
        gam += [g - 2 * g * k / (2 * q) for k in range(1, 2 * q + 1)]
        gam += [-g + g * k / (L - 3 * q) for k in range(1, L - 3 * q + 1)]
        return ess, [np.where(top_x, gk, 0.0) for gk in gam], {"path": "cyclic shear of z = 1 against z = 0",
                                                               "gamma": [round(v, 6) for v in gam]}
    ess0 = space.boundary_dofs(0b000001, 0x7)
    ess1 = space.boundary_dofs(0b000010, 0x1)
    ess = np.union1d(ess0, ess1).astype(np.int32)
    final = np.where(np.isin(ess, ess1), -args.strain, 0.0)
    return ess, [final * (k / L) for k in range(1, L + 1)], {"path": "uniaxial compression", "strain": args.strain}


def mode_precisions(sx, mode):
    """(tangent precision, J.v arithmetic) of a solve mode name"""
    if mode.startswith("mixed_bf16"):
        return sx.capi.SX_PREC_BF16, sx.capi.SX_PREC_F32
    if mode.startswith("mixed_f32"):
        return sx.capi.SX_PREC_F32, sx.capi.SX_PREC_F32
    if mode.startswith("f64"):
        return sx.capi.SX_PREC_F64, sx.capi.SX_PREC_F64
    raise SystemExit(f"bench: unknown solve mode {mode}")


def run_solve(sx, args, space, make_form, precisions, mesh_label, qf="neohookean"):
    """The load-stepped path of the workload (load_path; warm start, Newton + Krylov
    per step) in each solve mode; wall time per mode."""
    ess, targets, path_info = load_path(space, args, qf)
    solve, finals = {}, {}
    for mode in precisions.split(","):
        tprec, arith = mode_precisions(sx, mode)
        # FP64 and FP32-arithmetic forms each take their own scatter option (both atomic by default)
        scatter = args.scatter if mode.startswith("f64") else args.split_scatter
        form = make_form(scatter)
        form.set_option(sx.capi.SX_OPT_PRINT_LEVEL, args.print_level)
        form.set_option(sx.capi.SX_OPT_LINE_SEARCH, args.line_search)
        form.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, tprec)
        form.set_option(sx.capi.SX_OPT_JVP_ARITHMETIC, arith)
        if "_gmres" in mode:
            # (flexible) GMRES: the J2 tangent is not symmetric under geometric nonlinearity
            form.set_option(sx.capi.SX_OPT_LINEAR_SOLVER, sx.capi.SX_LINSOLVER_GMRES)
            form.set_option(sx.capi.SX_OPT_GMRES_RESTART, args.gmres_restart)
        mg = None
        if "_mg" in mode:
            # the multigrid V-cycle (set up once per linearization, inside the timed solve); with
            # _bf16smooth its level-0 smoothing reads BF16 copies of the FP32 tangent (and, with
            # _bf16smooth32, runs on FP32 vectors)
            mg = sx.Multigrid(space, fine_degree=args.mg_fine_degree, coarse_degree=args.mg_coarse_degree,
                              smoothing_range=args.mg_range)
            mg.set_reuse(args.mg_reuse)
            if "_bf16smooth" in mode:
                mg.set_smoother_precision(sx.capi.SX_PREC_BF16)
            if "_bf16smooth32" in mode:
                mg.set_smoother_vectors(sx.capi.SX_PREC_F32)
            if "_learned" not in mode:
                form.set_multigrid(mg)
        learned, t_train, train_loss = None, 0.0, None
        if "_learned" in mode:
            # a linear U-Net trained (timed) on pairs from the linearization at
            # the undeformed state, then flexible CG in every load step; with
            # _mg the network corrects the V-cycle (trained on the error the
            # V-cycle leaves); _nonet: the same flexible CG with the network's
            # correction off (gain 0, no training), the ablation
            tt = time.perf_counter()
            learned = sx.Learned(space, seed=11)
            learned.set_activation(sx.capi.SX_UNET_IDENTITY, seed=11)
            learned.configure(max_pairs=args.learn_pairs, stall_window=20)
            if mg is not None:
                learned.set_multigrid(mg)
            if mode.endswith("_nonet"):
                learned.set_gain(0.0)
                log(f"{mode}: network correction off (gain 0)")
            else:
                _, g0 = form(0.0, sx.DeviceVector(space.ndof), wrt=0)
                learned.generate_pairs(g0, ess, args.learn_pairs, seed=5)
                del g0
                train_loss = learned.train(args.learn_steps, lr=args.learn_lr, seed=3, batch=args.learn_batch)
                sx.capi.call("sx_device_synchronize")
                t_train = time.perf_counter() - tt
                log(f"{mode}: trained {args.learn_steps} Adam steps x {args.learn_batch} pairs of "
                    f"{learned.info()['pairs']} in {t_train:.2f} s (pair generation included), loss of the last "
                    f"tenth {train_loss:.4f}, gain {learned.gain():.3f}")
            form.set_preconditioner(learned, sx.capi.SX_PRECOND_LEARNED)
        x = sx.DeviceVector(space.ndof)
        sx.capi.call("sx_device_synchronize")
        # one load step per call (the same arithmetic as one call over all
        # steps: each call starts from the essential values x holds) so that
        # every step reports as it finishes
        reps, t_solve = [], 0.0
        for k, target in enumerate(targets, start=1):
            ts = time.perf_counter()
            reps += sx.quasistatic_solve(form, x, ess, target, 1, warm_start=True,
                                         rel_tol=args.newton_tol, lin_rel_tol=args.cg_tol,
                                         lin_max_iter=args.cg_max)
            sx.capi.call("sx_device_synchronize")
            t_solve += time.perf_counter() - ts
            log(f"{mode} load step {k}/{args.load_steps}: Newton {reps[-1]['newton_iterations']}, "
                f"CG {reps[-1]['linear_iterations']}, {t_solve:.2f} s so far")
        solve[mode] = {
            "seconds": round(t_solve, 3),
            "load_steps": args.load_steps,
            **path_info,
            "newton_rel_tol": args.newton_tol,
            "max_line_search_iterations": args.line_search,
            "cg_rel_tol": args.cg_tol,
            "precond": "Jacobi",
            "linear_solver": "GMRES" if "_gmres" in mode else "CG",
            "tangent": {0: "f64", 1: "f32", 2: "bf16"}[tprec],
            "jvp_arithmetic": {0: "f64", 1: "f32"}[arith],
            "newton_iterations": [r["newton_iterations"] for r in reps],
            "line_search_steps": [r["line_search_steps"] for r in reps],
            "cg_iterations": [r["linear_iterations"] for r in reps],
            "final_rel_residual": [r["final_norm"] / r["initial_norm"] if r["initial_norm"] else 0.0
                                   for r in reps],
        }
        solve[mode]["scatter"] = scatter
        solve[mode]["mesh"] = mesh_label
        if mg is not None:
            info = mg.info()
            solve[mode]["precond"] = "geometric multigrid V-cycle (Chebyshev-Jacobi, Galerkin vertex-grid levels)"
            solve[mode]["multigrid"] = {
                "fine_degree": args.mg_fine_degree, "coarse_degree": args.mg_coarse_degree,
                "smoothing_range": args.mg_range, "setups": info["setups"], "reuse": args.mg_reuse,
                "smoother_tangent": "bf16 copy" if "_bf16smooth" in mode else "the form's",
                "smoother_vectors": "fp32" if "_bf16smooth32" in mode else "fp64",
                "levels": [lv["dims"] for lv in info["levels"]],
            }
            if learned is None:
                form.set_multigrid(None)
            del mg
        if learned is not None:
            # the wall time above includes the training
            solve[mode]["seconds"] = round(t_solve + t_train, 3)
            base = "the multigrid V-cycle" if "_mg" in mode else "Jacobi"
            solve[mode]["precond"] = (f"{base} in flexible CG, network correction off (gain 0)"
                                      if mode.endswith("_nonet") else
                                      f"learned linear U-Net correction + {base} in flexible CG (Jacobi-CG fallback)")
            solve[mode]["learned"] = {
                "train_seconds": round(t_train, 3),
                "adam_steps": args.learn_steps,
                "batch": args.learn_batch,
                "lr": args.learn_lr,
                "pairs": learned.info()["pairs"],
                "pairs_from": "the linearization at the undeformed state: r = J_bc e for random smooth e",
                "activation": "identity",
                "gain": learned.gain(),
                "train_loss_last_tenth": train_loss,
                "fallbacks_to_jacobi_cg": learned.info()["fallbacks"],
            }
            form.set_preconditioner(None, sx.capi.SX_PRECOND_JACOBI)
            del learned
        log(f"{mode} solve {t_solve:.2f} s: Newton {solve[mode]['newton_iterations']}, "
            f"CG {solve[mode]['cg_iterations']}")
        finals[mode] = x.numpy()
        del x, form
    ref_mode = next((m for m in ("f64", "f64_mg") if m in finals), None)
    if ref_mode is not None:
        ref = finals[ref_mode]
        for mode, xm in finals.items():
            # final displacement against the FP64 solve (Jacobi-CG, else multigrid-CG)
            solve[mode]["x_err_reference"] = ref_mode
            solve[mode]["x_max_err_rel"] = float(np.abs(xm - ref).max() / np.abs(ref).max())
            solve[mode]["x_rms_err_rel"] = float(np.sqrt(np.mean((xm - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    del finals
    if "f64" in solve:
        for mode in solve:
            solve[mode]["speedup_vs_f64"] = round(solve["f64"]["seconds"] / solve[mode]["seconds"], 3)
    return solve


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="neohookean_p3_64", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-sample-n", type=int, default=0,
                    help="mesh intervals per side of the CPU baseline sample (0: the workload's own mesh, at most 64)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="skip the FP32 / BF16 J.v measurements")
    ap.add_argument("--solve", action="store_true",
                    help="time the load-stepped Newton + Krylov solve on the workload's mesh (without it, a 24^3 "
                         "solve of the same load path runs in FP64 and mixed precision, with Jacobi and multigrid)")
    ap.add_argument("--no-default-solve", action="store_true", help="skip the 24^3 solve")
    ap.add_argument("--load-steps", type=int, default=10)
    ap.add_argument("--strain", type=float, default=0.05)
    ap.add_argument("--shear", type=float, default=0.03, help="J2 workloads: the cyclic shear amplitude")
    ap.add_argument("--gmres-restart", type=int, default=50)
    ap.add_argument("--newton-tol", type=float, default=1e-8)
    ap.add_argument("--line-search", type=int, default=0,
                    help="max_line_search_iterations of the Newton loops of the solves (0: plain Newton)")
    ap.add_argument("--print-level", type=int, default=0,
                    help="> 0: the Newton loops of the solves report every iteration and linear solve on stderr")
    ap.add_argument("--cg-tol", type=float, default=1e-6)
    ap.add_argument("--cg-max", type=int, default=20000)
    ap.add_argument("--solve-precision", default="f64",
                    help="comma-separated solve modes: f64 (plain FP64 Jacobi-CG), mixed_f32 / mixed_bf16 "
                         "(FP32 element arithmetic with an FP32 / BF16 tangent inside FP64 Newton and Krylov), "
                         "mixed_f32_learned (mixed_f32 with flexible CG and the learned U-Net preconditioner, "
                         "trained first on pairs from the undeformed linearization; training time included), "
                         "f64_mg / mixed_f32_mg / mixed_bf16_mg (f64 / mixed_f32 / mixed_bf16 with multigrid-CG), "
                         "mixed_f32_mg_bf16smooth (mixed_f32_mg with BF16 tangent copies in the level-0 smoothing), "
                         "mixed_f32_mg_bf16smooth32 (and FP32 vectors in the level-0 smoothing)")
    ap.add_argument("--mg-fine-degree", type=int, default=2, help="Chebyshev degree on the order-p level")
    ap.add_argument("--mg-coarse-degree", type=int, default=2, help="Chebyshev degree on the vertex levels")
    ap.add_argument("--mg-range", type=float, default=12.0, help="Chebyshev smoothing range")
    ap.add_argument("--mg-reuse", type=int, default=1,
                    help="linear solves one multigrid hierarchy serves inside a load step (1: set up per linearization)")
    ap.add_argument("--learn-steps", type=int, default=2000, help="Adam steps of the learned preconditioner")
    ap.add_argument("--learn-lr", type=float, default=1e-3)
    ap.add_argument("--learn-batch", type=int, default=8, help="pairs per Adam step")
    ap.add_argument("--learn-pairs", type=int, default=300,
                    help="pairs generated from the first linearization (random smooth corrections e, r = J e)")
    ap.add_argument("--unet", action="store_true",
                    help="also time the learned preconditioner's U-Net (inference and one training step) on the "
                         "workload's vertex grid")
    ap.add_argument("--scatter", default="atomic", choices=["colored", "atomic"])
    ap.add_argument("--split-scatter", default="atomic", choices=["colored", "atomic"],
                    help="scatter mode of the FP32 / BF16 J.v measurements and of the FP32-arithmetic solves")
    ap.add_argument("--xcd-remap", type=int, default=0,
                    help="1: each XCD takes a contiguous run of element blocks; 0: the hardware's round-robin order")
    ap.add_argument("--scatter-order", type=int, default=2,
                    help="2: element values added in address order through LDS after merging the nodes shared with "
                         "the previous element of the workgroup; 1: without the merge; 0: one component per "
                         "instruction")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a known-size read (dot of two L-vectors) and write (fill) for counter scaling")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X "
              "(multi-GPU sharding is design-only); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import serac_amd as sx
    from serac_amd import workloads as W

    n, p, perturb, params, qf = WORKLOADS[args.workload]
    sx.capi.call("sx_init", 0)
    dev = sx.device_name()
    log(f"device {dev}; workload {args.workload}")

    t0 = time.time()
    mesh = sx.Mesh.cuboid(n, n, n)
    if perturb > 0:
        mesh.perturb(42, perturb)
    space = sx.Space(mesh, p, 3)
    nq1 = p + 1
    qdatas = []

    def make_form(scatter, sp=None):
        sp = space if sp is None else sp
        f = sx.Functional(sp, [sp])
        f.set_option(sx.capi.SX_OPT_SCATTER,
                     sx.capi.SX_SCATTER_ATOMIC if scatter == "atomic" else sx.capi.SX_SCATTER_COLORED)
        f.set_option(sx.capi.SX_OPT_XCD_REMAP, args.xcd_remap)
        f.set_option(sx.capi.SX_OPT_SCATTER_ORDER, args.scatter_order)
        if qf == "j2":
            qd = sx.QuadratureData(sp.ne, nq1 ** 3, np.zeros(10))
            qdatas.append(qd)
            f.add_domain_integral(sx.SX_QF_J2_LINEAR, params, qdata=qd)
        else:
            f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, params)
        return f

    form = make_form(args.scatter)
    xyz = space.node_coords()
    if qf == "j2":
        # simple shear to 3 % (the end of the cyclic path): every point is plastic
        Uh = np.zeros(space.ndof)
        Uh[0::3] = 0.03 * xyz[:, 2]
        U = sx.DeviceVector.from_numpy(Uh)
        del Uh
    else:
        U = sx.DeviceVector.from_numpy(W.neohookean_displacement(xyz, noise=W.noise_for_mesh(n)))
    dU = sx.DeviceVector.from_numpy(W.direction(space.nnodes, 3))
    dR = sx.DeviceVector(space.ndof)
    del xyz
    log(f"setup {time.time() - t0:.1f} s: {space.ndof} DOFs, {space.ne} elements")

    nqp = space.ne * nq1 ** 3

    # residual and linearization (timed with the host clock around a sync)
    form(0.0, U)
    tr = time.perf_counter()
    for _ in range(3):
        form(0.0, U)
    t_res = (time.perf_counter() - tr) / 3
    form(0.0, U, wrt=0)
    tl = time.perf_counter()
    _, grad = form(0.0, U, wrt=0)
    t_lin = time.perf_counter() - tl
    tangent_bytes = grad.tangent_bytes()
    # Jacobi diagonal from the stored tangent (one warm call, then the mean of three)
    dg = grad.diagonal()
    td = time.perf_counter()
    for _ in range(3):
        grad.diagonal(out=dg)
    t_diag = (time.perf_counter() - td) / 3
    dg.free()
    log(f"residual {t_res * 1e3:.2f} ms, linearize {t_lin * 1e3:.2f} ms, diagonal {t_diag * 1e3:.2f} ms, "
        f"tangent {tangent_bytes / 1e9:.2f} GB")

    # J.v: warmup, then exactly `steps` applications bracketed by device syncs
    if args.warmup > 0:
        grad.apply_timed(dU, dR, args.warmup)
    sx.capi.call("sx_device_synchronize")
    tw = time.perf_counter()
    ms_total, ms_kern, launches = grad.apply_timed(dU, dR, args.steps)
    sx.capi.call("sx_device_synchronize")
    wall = time.perf_counter() - tw
    ms_step = ms_total / args.steps
    value = W.gdof_per_s(space.ndof, ms_step / 1e3)
    jv = dR.numpy()
    r64 = form(0.0, U).numpy().copy()
    if not np.isfinite(jv).all() or not np.isfinite(r64).all():
        raise SystemExit("bench: non-finite residual or J.v; the measurement is invalid")
    jv_norm = float(np.linalg.norm(jv))
    # dominant kernel: the colour launches of the J.v element kernel
    d_sym = TANGENT_ENTRIES[qf]
    b_alg = W.jvp_algorithmic_bytes(space.ndof, nqp, space.ne, space.npe, d_sym)
    kern_ms_per_apply = ms_kern / args.steps
    achieved = b_alg / (kern_ms_per_apply / 1e3) / 1e9
    log(f"J.v {ms_step:.3f} ms/step ({value:.2f} GDOF/s), element kernels {kern_ms_per_apply:.3f} ms "
        f"over {launches} launches, {achieved:.0f} GB/s algorithmic, wall {wall:.2f} s")

    # precision-split operator: FP32 vectors and arithmetic, tangent in FP32 or BF16
    split = {}
    if not args.no_split:
        # FP32 scatter-adds go faster as coloured read-modify-writes than as
        # atomics, so the split operator runs on a coloured form
        grad = None
        del form
        sform = make_form(args.split_scatter)
        dU32 = sx.DeviceVector.from_numpy(W.direction(space.nnodes, 3), dtype=np.float32)
        dR32 = sx.DeviceVector(space.ndof, dtype=np.float32)
        for name, tprec, s_d in (("fp32", sx.capi.SX_PREC_F32, 4), ("fp32_bf16", sx.capi.SX_PREC_BF16, 2)):
            sform.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, tprec)
            r32, g32 = sform(0.0, U, wrt=0)
            # the precision-split Newton's residual: measured against the FP64 form's
            r32h = r32.numpy()
            del r32
            if args.warmup > 0:
                g32.apply_timed(dU32, dR32, args.warmup)
            t32, k32, _ = g32.apply_timed(dU32, dR32, args.steps)
            jv32 = dR32.numpy().astype(np.float64)
            if not np.isfinite(jv32).all():
                raise SystemExit(f"bench: non-finite {name} J.v; the measurement is invalid")
            b32 = W.jvp_algorithmic_bytes(space.ndof, nqp, space.ne, space.npe, d_sym, s_v=4, s_d=s_d)
            ach = b32 / (k32 / args.steps / 1e3) / 1e9
            split[name] = {
                "value": round(W.gdof_per_s(space.ndof, t32 / args.steps / 1e3), 4),
                "ms_per_step": round(t32 / args.steps, 5),
                "roofline_achieved_GBs": round(ach, 1),
                "roofline_frac": round(ach / HBM_PEAK_GBS, 4),
                "algorithmic_bytes_per_apply": b32,
                "rel_err_vs_fp64": float(np.linalg.norm(jv32 - jv) / np.linalg.norm(jv)),
                # max and RMS error of J.v against the FP64 operator, relative to its max and RMS
                "jv_max_err_rel": float(np.abs(jv32 - jv).max() / np.abs(jv).max()),
                "jv_rms_err_rel": float(np.sqrt(np.mean((jv32 - jv) ** 2)) / np.sqrt(np.mean(jv ** 2))),
                # the residual the precision-split Newton evaluates (FP64 by design) against the FP64 form's
                "residual_max_err_rel": float(np.abs(r32h - r64).max() / np.abs(r64).max()),
                "residual_rms_err_rel": float(np.sqrt(np.mean((r32h - r64) ** 2)) / np.sqrt(np.mean(r64 ** 2))),
            }
            del r32h
            log(f"{name} J.v {t32 / args.steps:.3f} ms/step ({split[name]['value']:.2f} GDOF/s), "
                f"{ach:.0f} GB/s algorithmic, rel. error vs FP64 {split[name]['rel_err_vs_fp64']:.2e}")
            split[name]["scatter"] = args.split_scatter
            del g32
        del dU32, dR32, sform
        form = make_form(args.scatter)
    del jv, r64

    unet = None
    if args.unet:
        unet = unet_bench(sx, n, args.steps, args.warmup)
        log(f"U-Net on {n + 1}^3 vertices: forward {unet['forward_ms']:.3f} ms ({unet['forward_tflops']:.1f} TFLOP/s), "
            f"training step {unet['train_step_ms']:.3f} ms ({unet['train_tflops']:.1f} TFLOP/s)")

    solve = None
    if args.solve:
        del form
        solve = run_solve(sx, args, space, make_form, args.solve_precision,
                          f"{n}^3 hexes, p={p}, perturbed {perturb}h (the workload's mesh)", qf)
        form = make_form(args.scatter)
    solve_full = None
    if not args.solve and qf == "neohookean" and not args.no_default_solve:
        # the full-size load path in mixed precision with multigrid-CG, its
        # level-0 smoothing on FP32 vectors and BF16 tangent copies (about 8 s
        # at 64^3); its FP64 Jacobi-CG baseline takes minutes and runs with
        # --solve --solve-precision f64
        del form
        solve_full = run_solve(sx, args, space, make_form, "mixed_f32_mg_bf16smooth32",
                               f"{n}^3 hexes, p={p}, perturbed {perturb}h (the workload's mesh)")
        solve_full["note"] = ("FP64 Jacobi-CG on this load path takes minutes and is not rerun by default: "
                              "bench.py --solve --solve-precision f64")
        form = make_form(args.scatter)
        # the same load path on a 24^3 mesh, FP64 Jacobi-CG against mixed
        # precision and multigrid-CG
        ns = 24
        smesh = sx.Mesh.cuboid(ns, ns, ns)
        if perturb > 0:
            smesh.perturb(42, perturb)
        sspace = sx.Space(smesh, p, 3)
        solve = run_solve(sx, args, sspace, lambda sc: make_form(sc, sspace), "f64,mixed_f32,f64_mg,mixed_f32_mg,mixed_f32_mg_bf16smooth32",
                          f"{ns}^3 hexes, p={p}, perturbed {perturb}h ({sspace.ndof} DOFs)")
        del sspace, smesh

    stream_read = None
    if args.calibrate:
        # reference kernels for HBM counter scaling: k_dot_partial reads 2 * 8 * ndof
        # bytes, k_fill writes 8 * ndof bytes
        import ctypes

        res = ctypes.c_double()
        sx.capi.call("sx_vec_dot", U.ptr, dU.ptr, space.ndof, ctypes.byref(res))
        dR.zero()
        log(f"calibration: dot reads {2 * 8 * space.ndof} B, fill writes {8 * space.ndof} B")
        # streaming read rate of the library's own dot product over two vectors
        # as large as the FP64 tangent (each call synchronises, so the host's
        # launch and copy-back time is included)
        nbig = int(tangent_bytes // 16)
        va, vb = sx.DeviceVector(nbig), sx.DeviceVector(nbig)
        va.zero()
        vb.zero()
        sx.capi.call("sx_vec_dot", va.ptr, vb.ptr, nbig, ctypes.byref(res))
        ts = time.perf_counter()
        for _ in range(10):
            sx.capi.call("sx_vec_dot", va.ptr, vb.ptr, nbig, ctypes.byref(res))
        stream_read = 16 * nbig * 10 / (time.perf_counter() - ts) / 1e9
        va.free()
        vb.free()
        log(f"calibration: dot over 2 x {8 * nbig / 1e9:.2f} GB reads {stream_read:.0f} GB/s")

    cpu = None
    if not args.no_cpu_baseline:
        tc = time.time()
        cpu = cpu_baseline(args.cpu_sample_n or min(n, 64), p, perturb, params, args.cpu_seconds, qf)
        log(f"cpu baseline {cpu['value']:.5f} GDOF/s ({time.time() - tc:.1f} s)")

    rec = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {'J2 plasticity' if qf == 'j2' else 'Neo-Hookean'} J.v, {n}^3 hexes, "
                        f"p={p}, {nq1}^3 Gauss points, perturbed {perturb}h",
            "ndof": space.ndof,
            "elements": space.ne,
            "qpoints": nqp,
            "parallelism": "1 GPU",
            "scatter": args.scatter,
            "scatter_order": args.scatter_order,
            "xcd_remap": args.xcd_remap,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(args.workload, args.scatter, args.xcd_remap, args.scatter_order),
            "traffic_source": PROFILED_TRAFFIC.get((args.workload, args.scatter, args.xcd_remap, args.scatter_order)),
            "kernel": f"k_jvp<{nq1},3,{'GRAD_FULL' if qf == 'j2' else 'GRAD_SYM'},double,double,double>"
                      + (" (one launch per element colour)" if args.scatter == "colored"
                         else " (one launch, atomic scatter)"),
            "algorithmic_bytes_per_apply": b_alg,
            "kernel_ms_per_apply": round(kern_ms_per_apply, 5),
            "launches_per_apply": launches,
            "stream_read_GBs": round(stream_read, 1) if stream_read else None,
        },
        "cpu_baseline": cpu,
        "residual_ms": round(t_res * 1e3, 3),
        "linearize_ms": round(t_lin * 1e3, 3),
        "diagonal_ms": round(t_diag * 1e3, 3),
        "tangent_bytes": tangent_bytes,
        "jv_norm": jv_norm,
        "precision_split": split,
        "newton_solve": solve,
        "newton_solve_full_size": solve_full,
        "learned_unet": unet,
        "device": dev,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
