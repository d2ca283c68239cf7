"""Benchmark: timesteps/s of the fractional-step hot path on one MI355X.

Workload: lid-driven cavity on the unit cube, 64^3 hex27-velocity /
hex8-pressure elements (129^3 velocity nodes, 65^3 pressure nodes), 27-point
Gauss quadrature, Re = 1000 (viscosity 1/1000 with lid speed 1), dt = 0.0025,
3 inner iterations per step at tolerance 1e-3, pressure solved to a relative
residual of 1e-8 from the previous Pdot in every mode (the reference's GPU
path warm-starts its CG the same way; --x0 zero gives its CPU path's cold
start). Inputs are synthetic (generated mesh, zero initial fields plus the
lid BC) and resident in HBM before timing starts.

One "step" is one time step of the reference's timeLoop (BCH:2859-3118):
convection residual, step1, step2 (assembly of R2 + the pressure solve),
step3, convergence checks, K*Acc, for up to 3 inner iterations. The same
trajectory is run with the plain FP64 Jacobi-PCG, with the mixed-precision
form (FP32 assembly of K, G, G^T and the convection, FP32 inner PCG, FP64
refinement residual) and with learned+mixed
(flexible PCG in FP32 preconditioned by Jacobi plus the trained multilevel
network, FP64 reliable updates and stop test, Jacobi-PCG fallback), and with
the direct solve (band Cholesky of Z factored once in the first warm-up
step, the MATLAB prototype's pressure step); `value` is the headline
(learned+mixed) mode, the other rates, the pressure-solve speed-up over FP64
and the field errors against the FP64 run sit beside it.
The network is set up, fed with the solver's own pressure solves and
trained before the timed region (train_s in the output).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cfd-with-cuda_amd"))

METRIC = "timesteps/s, 64^3 hex27 cavity: learned+mixed vs plain FP64 PCG at 1e-8 residual"


def step_kernels(vt):
    """rocprofv3 names of the step's row kernels for matrix values of type vt
    ("double", or "float" in the mixed forms, whose step1 also gathers K's
    operands from their FP32 shadows): step1 with K*Acc_prev fused and G*Pn
    kept (inner iterations >= 2), step2's right-hand side, step3."""
    if vt == "float":
        return ("k_step1<8, 2, 2, true, true, 2, float, true>", "k_step2_rhs<32, 1, float, true>",
                "k_step3<4, 1, float>")
    return ("k_step1<8, 2, 2, true, true, 2, double, false>", "k_step2_rhs<16, 8, double, false>",
            "k_step3<8, 2, double>")


def zspmv_name(prec, npat, ncls):
    """rocprofv3 name of the Z SpMV kernel the solvers run in this form."""
    if prec and ncls > 0:
        return "k_zcls_spmv<8, false, 2>"
    if npat > 0:   # the row-pattern form runs 8 lanes x 16 loads per row in both precisions
        return "k_pcg_spmv<float, 8, 16, true>" if prec else "k_pcg_spmv<double, 8, 16, true>"
    return "k_pcg_spmv<float, 32, 4, false>" if prec else "k_pcg_spmv<double, 16, 8, false>"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


X0_POLICY = {"zero": 0, "previous": 1}
GRAPHS = 0   # --graphs


def run_mode(solver, mode, steps, warmup, rtol, maxit, matrix_bits=64, x0="zero"):
    """Warm-up then timed steps in one solver mode; matrix_bits 32 gives the
    mixed forms' FP32 assembly (K, G, G^T copies and the convection); x0 the
    pressure solves' initial guess (zero, or the previous Pdot)."""
    import numpy as np
    solver.reset(mode=mode, rtol=rtol, maxit=maxit)
    solver.set_graphs(GRAPHS)
    solver.set_solver_x0(X0_POLICY[x0])
    solver.set_matrix_precision(matrix_bits)
    solver.set_convection_precision(matrix_bits)
    solver.timing_reset(True)
    if warmup:
        solver.run(warmup)
    solver.timing_reset(True)
    solver.sync()
    t0 = time.perf_counter()
    st = solver.run(steps)
    solver.sync()
    t1 = time.perf_counter()
    tm = solver.timing()
    assert len(st) == steps, "steady state reached early"
    return dict(
        wall_s=t1 - t0, steps_per_s=steps / (t1 - t0), pcg_iters=sum(s["pcg"] for s in st),
        inner_iters=sum(s["inner"] for s in st), max_relres=max(s["relres"] for s in st),
        solve_ms=tm["pcg_ms"], Un=solver.download("Un"), Pn=solver.download("Pn"),
        mon=st[-1]["mon"], per_step=[dict(pcg=s["pcg"], inner=s["inner"], wall_s=s["wall_s"]) for s in st],
    )


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    runs (FETCH_SIZE and WRITE_SIZE of bench.py --probe, each scaled by
    the same counter on a copy of known size), or None when not measured."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_cavity64.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get("fsx::" + kernel)
    return None if k is None else round(k["hbm_bytes_per_launch"])


def train_learned(solver, args):
    """Untimed preparation of the learned mode: capture the pressure solves of
    a few mixed-precision steps, build training pairs from truncated
    Jacobi-PCG iterates, train with Adam on the GPU."""
    import fsx
    t0 = time.perf_counter()
    solver.reset(mode=fsx.SOLVER_MIXED, rtol=args.rtol, maxit=args.maxit)
    solver.set_matrix_precision(32)
    solver.set_convection_precision(32)
    n1 = args.n + 1
    solver.unet_form(args.learned_form)
    solver.unet_setup((n1, n1, n1), seed=args.seed)
    solver.unet_capture(64)
    solver.run(args.train_capture_steps)
    ncap = solver.unet_captured()
    solver.unet_capture(0)
    nsamp = solver.unet_prepare([0, 5, 20, 50, 100])
    loss0 = solver.unet_train(1, batch=8, lr=1e-12, seed=args.seed)
    loss = solver.unet_train(args.train_iters, batch=4, lr=args.lr, seed=args.seed + 1)
    solver.sync()
    return dict(train_s=time.perf_counter() - t0, captured=ncap, samples=nsamp, loss_initial=loss0, loss_final=loss,
                train_iters=args.train_iters, params=solver.unet_nparams())


def cpu_baseline(mesh, args, gpu_pcg_per_step, gpu_inner_per_step):
    """The CPU oracle on the host cores, on a bounded sample of the same 64^3
    workload: setup (untimed), then one inner iteration (step1, step2 with the
    PCG capped at `cpu_pcg_iters` iterations, step3) timed, extrapolated to a
    full step with the GPU run's measured inner and PCG iteration counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsx_oracle
    if not os.path.exists(fsx_oracle._LIB_PATH):
        fsx_oracle.build()
    cores = len(os.sched_getaffinity(0))
    fsx_oracle.set_threads(cores)
    t0 = time.perf_counter()
    o = fsx_oracle.Oracle()
    o.set_mesh(mesh.arrays())
    o.set_params(NGP=args.ngp, viscosity=1.0 / args.re, dt=args.dt, maxIter=3, tolerance=1e-3, converge=1e-6,
                 pcg_rtol=1e-30, pcg_maxit=args.cpu_pcg_iters)
    o.setup()
    t_setup = time.perf_counter() - t0
    o.set_d("UnpHalf_prev", o.d("Un"))
    t0 = time.perf_counter()
    o.step1(1)
    t1 = time.perf_counter()
    o.step2(1)
    t2 = time.perf_counter()
    o.step3(1)
    t3 = time.perf_counter()
    tm = o.timings()
    t_pcg_iter = tm["pcg_s"] / max(1, tm["pcg_iters"])
    t_other = (t1 - t0) + (t2 - t1 - tm["pcg_s"]) + (t3 - t2)
    t_step = gpu_inner_per_step * t_other + gpu_pcg_per_step * t_pcg_iter
    o.close()
    return dict(value=1.0 / t_step, unit="timesteps/s", cores=cores, kind="port",
                sample=(f"64^3 cavity on the CPU oracle (C++/OpenMP, {cores} threads): one inner iteration timed "
                        f"(step1 {t1 - t0:.2f} s, step2 non-solve {t2 - t1 - tm['pcg_s']:.2f} s, step3 {t3 - t2:.2f} s) "
                        f"and {tm['pcg_iters']} FP64 Jacobi-PCG iterations ({t_pcg_iter * 1e3:.1f} ms each), "
                        f"extrapolated to {gpu_inner_per_step:.2f} inner iterations and {gpu_pcg_per_step:.1f} PCG "
                        f"iterations per step as measured on the GPU; setup {t_setup:.1f} s untimed"),
                t_step_s=t_step)


def duct_main(args):
    """The bent-duct workload: 16 x 16 cross-section elements, 40 + 40 + 80
    elements along the axis (5-unit inlet, bend of centreline radius 2,
    10-unit outlet), node numbering permuted from a seed, parabolic inlet of
    peak 1, Re 200, dt 0.005, consistent-mass velocity updates (the three
    components as one 3-RHS Jacobi-PCG per update) and the FP64 / mixed
    pressure solvers. Prints one JSON line of its own (not the headline)."""
    import fsx
    mesh = fsx.Mesh.duct(n=16, n_in=40, n_arc=40, n_out=80, seed=args.seed)
    mesh.set_params(NGP=args.ngp, dt=0.005, maxIter=3, tolerance=1e-3, converge=1e-6, viscosity=1.0 / 200)
    log("setup bent duct ...")
    solver = fsx.Solver(mesh, device=0, NGP=args.ngp, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=args.rtol,
                        pcg_maxit=args.maxit, solver_mode=0)
    solver.set_mass(fsx.MASS_CONSISTENT, rtol=1e-10, maxit=2000)
    res = {}
    for name, mode in (("fp64", fsx.SOLVER_FP64), ("mixed", fsx.SOLVER_MIXED)):
        r = run_mode(solver, mode, args.steps, args.warmup, args.rtol, args.maxit)
        m_it, m_rel, _ = solver.mass_stats()
        res[name] = {"steps_per_s": round(r["steps_per_s"], 4), "pcg_iters": r["pcg_iters"],
                     "inner_iters": r["inner_iters"], "solve_ms": round(r["solve_ms"], 2),
                     "mass_iters_last_solve": m_it, "mass_relres_last_solve": m_rel}
        log(f"  {name}: {res[name]}")
    out = {"metric": "timesteps/s, bent duct, consistent mass (3-RHS block PCG)", "value": res["mixed"]["steps_per_s"],
           "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "90-degree bent square duct 16x16x160 hex27/hex8, permuted numbering, Re 200, "
                                  f"dt 0.005, consistent mass, pressure rtol {args.rtol:g}",
                      "NN": solver.sizes["NN"], "NNp": solver.sizes["NNp"], "nnz_M": solver.sizes["nnzM"]},
           "modes": res}
    solver.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--re", type=float, default=1000.0)
    ap.add_argument("--dt", type=float, default=0.0025)
    ap.add_argument("--ngp", type=int, default=27)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--maxit", type=int, default=20000)
    ap.add_argument("--modes", default="fp64,mixed,learned,direct",
                    help="solver modes to run, of fp64, mixed, learned and direct (band Cholesky of Z, the MATLAB "
                         "prototype's pressure solve)")
    ap.add_argument("--train-iters", type=int, default=400)
    ap.add_argument("--train-capture-steps", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--learned-form", type=int, default=1, choices=(0, 1),
                    help="preconditioner of the learned mode: 1 two-level (Jacobi smoothing around the network, "
                         "tiled Z products), 0 additive (dinv r + network)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--probe", action="store_true",
                    help="only set up the case, run one FP64 time step and one with FP32 matrices and convection, time "
                         "the step kernels and the Z SpMV kernels and run a 256 MiB "
                         "calibration copy (the workload of the rocprofv3 --pmc traffic runs)")
    ap.add_argument("--x0", default="previous", choices=("zero", "previous"),
                    help="initial guess of every pressure solve, in all modes: the previous Pdot (the reference's "
                         "GPU path, CUDAcodes.cu:425-442; default) or zero (its CPU path)")
    ap.add_argument("--workload", default="cavity", choices=("cavity", "duct"),
                    help="duct: the bent-duct case with consistent-mass updates, printed as a line of its own")
    ap.add_argument("--graphs", type=int, default=0, choices=(0, 1),
                    help="enqueue the solver iterations as replayed HIP graphs (1) or as individual launches (0)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-pcg-iters", type=int, default=30)
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = args.graphs
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the solver core is built "
              f"and measured for one GPU; sharding by mesh partitions is designed, not built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import fsx

    if args.workload == "duct":
        duct_main(args)
        return
    mesh = fsx.Mesh.cavity(args.n).set_params(NGP=args.ngp, dt=args.dt, maxIter=3, tolerance=1e-3, converge=1e-6,
                                              viscosity=1.0 / args.re)
    log(f"setup {args.n}^3 cavity ...")
    t0 = time.perf_counter()
    solver = fsx.Solver(mesh, device=0, NGP=args.ngp, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=args.rtol,
                        pcg_maxit=args.maxit, solver_mode=0)
    log(f"setup done in {time.perf_counter() - t0:.1f} s: {solver.setup_times()} sizes {solver.sizes}")
    if args.probe:
        solver.run(1)
        solver.sync()
        for kid in (fsx.KERNEL_STEP1, fsx.KERNEL_KACC, fsx.KERNEL_STEP2_RHS, fsx.KERNEL_STEP3):
            solver.bench_kernel(kid, reps=3)
        # the same with the mixed forms' FP32 matrices and convection
        solver.set_matrix_precision(32)
        solver.set_convection_precision(32)
        solver.run(1)
        solver.sync()
        for kid in (fsx.KERNEL_STEP1, fsx.KERNEL_KACC, fsx.KERNEL_STEP2_RHS, fsx.KERNEL_STEP3):
            solver.bench_kernel(kid, reps=3)
        solver.set_matrix_precision(64)
        solver.set_convection_precision(64)
        ms32, b32 = solver.bench_spmv(precision=1, reps=20)
        ms64, b64 = solver.bench_spmv(precision=0, reps=20)
        copy_bytes = 256 << 20
        msc = solver.bench_copy(copy_bytes, reps=5)
        print(json.dumps({"spmv_f32": {"ms": ms32, "bytes": b32}, "spmv_f64": {"ms": ms64, "bytes": b64},
                          "copy_bytes_each_way": copy_bytes, "copy_ms": msc,
                          "copy_GBs": 2 * copy_bytes / (msc * 1e-3) / 1e9}), flush=True)
        solver.close()
        return
    modes = {"fp64": fsx.SOLVER_FP64, "mixed": fsx.SOLVER_MIXED, "learned": fsx.SOLVER_LEARNED,
             "direct": fsx.SOLVER_DIRECT}
    res = {}
    train = None
    for name in args.modes.split(","):
        if name == "learned":
            log("learned mode: capture, prepare and train the network (untimed) ...")
            train = train_learned(solver, args)
            log(f"  {train}")
        log(f"mode {name}: {args.warmup} warmup + {args.steps} timed steps")
        res[name] = run_mode(solver, modes[name], args.steps, args.warmup, args.rtol, args.maxit,
                             matrix_bits=64 if name in ("fp64", "direct") else 32, x0=args.x0)
        r = res[name]
        if name == "direct":   # the factor is formed in the first warm-up step
            w, fms, fbytes = solver.direct_stats()
            r["extra"] = {"half_bandwidth": w, "factor_ms": round(fms, 1), "band_GB": round(fbytes / 1e9, 2)}
            log(f"  direct: {r['extra']}")
        log(f"  {r['steps_per_s']:.3f} steps/s, PCG iters {r['pcg_iters']}, solve {r['solve_ms']:.1f} ms, "
            f"max relres {r['max_relres']:.2e}")
    head = next((m for m in ("learned", "mixed") if m in res), list(res)[0])
    h = res[head]
    prec = 1 if head in ("mixed", "learned") else 0
    # live per-launch device times (HIP events on the context's stream) of the
    # step's row kernels and the headline solver's Z SpMV; the roofline entry
    # is the one with the most device time per step in the headline mode
    steps_inner = h["inner_iters"] / args.steps
    solver.set_matrix_precision(32 if prec else 64)   # the headline mode's matrices
    names = step_kernels("float" if prec else "double")
    kern = {}
    for kid, kname, per_step in ((fsx.KERNEL_STEP1, names[0], steps_inner),
                                 (fsx.KERNEL_STEP2_RHS, names[1], steps_inner),
                                 (fsx.KERNEL_STEP3, names[2], steps_inner)):
        ms_k, by_k = solver.bench_kernel(kid, reps=20)
        kern[kname] = dict(launch_ms=ms_k, bytes=by_k, ms_per_step=ms_k * per_step)
    ms, bytes_per = solver.bench_spmv(precision=prec, reps=100)
    ms64, bytes64 = solver.bench_spmv(precision=0, reps=100)
    npat, _ = solver.z_patterns()
    _, ncls, _ = solver.z_forms()
    zname = zspmv_name(prec, npat, ncls)
    kern[zname] = dict(launch_ms=ms, bytes=bytes_per, ms_per_step=ms * h["pcg_iters"] / args.steps)
    kname = max(kern, key=lambda k: kern[k]["ms_per_step"])
    achieved = kern[kname]["bytes"] / (kern[kname]["launch_ms"] * 1e-3) / 1e9
    h = res[head]
    out = {
        "metric": METRIC,
        "value": round(h["steps_per_s"], 4),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * h["wall_s"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"learned": "bf16+f32+f64", "mixed": "f32+f64"}.get(head, "f64"),
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity {args.n}^3 hex27/hex8, NGP {args.ngp}, Re {args.re:g}, "
                               f"dt {args.dt:g}, 3 inner iterations, pressure rtol {args.rtol:g}, "
                               f"pressure x0 {args.x0}",
                   "mode": head, "NN": solver.sizes["NN"], "NNp": solver.sizes["NNp"],
                   "nnz_Z": solver.sizes["nnzZ"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname),
                     "algorithmic_bytes_per_launch": kern[kname]["bytes"],
                     "launch_ms": round(kern[kname]["launch_ms"], 5)},
        "kernels": {k: {"launch_ms": round(v["launch_ms"], 5), "algorithmic_bytes": v["bytes"],
                        "achieved_GBs": round(v["bytes"] / (v["launch_ms"] * 1e-3) / 1e9, 1),
                        "ms_per_step": round(v["ms_per_step"], 4), "traffic": measured_traffic(k)}
                    for k, v in kern.items()},
        "modes": {k: {"steps_per_s": round(v["steps_per_s"], 4), "pcg_iters": v["pcg_iters"],
                      "inner_iters": v["inner_iters"], "solve_ms": round(v["solve_ms"], 2),
                      "max_relres": v["max_relres"], **v.get("extra", {})} for k, v in res.items()},
        "spmv_fp64": {"launch_ms": round(ms64, 5), "achieved_GBs": round(bytes64 / (ms64 * 1e-3) / 1e9, 1)},
        "z_row_patterns": npat,
    }
    if train is not None:
        out["learned_training"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in train.items()}
    if "fp64" in res and head != "fp64":
        f = res["fp64"]
        out["pressure_solve_speedup"] = round(f["solve_ms"] / h["solve_ms"], 3)
        out["vs_fp64_steps"] = round(h["steps_per_s"] / f["steps_per_s"], 3)
        out["velocity_rel_err_vs_fp64"] = float(np.linalg.norm(h["Un"] - f["Un"]) / np.linalg.norm(f["Un"]))
        out["pressure_rel_err_vs_fp64"] = float(np.linalg.norm(h["Pn"] - f["Pn"]) / np.linalg.norm(f["Pn"]))
    if not args.no_cpu:
        base = res.get("fp64", h)
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(mesh, args, base["pcg_iters"] / args.steps,
                                           base["inner_iters"] / args.steps)
    else:
        out["cpu_baseline"] = None
    solver.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
