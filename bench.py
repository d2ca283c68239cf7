"""Benchmark: timesteps/s of the fractional-step hot path on one MI355X.

Headline workload (the 64^3 cavity): lid-driven cavity on the unit cube,
64^3 hex27-velocity / hex8-pressure elements (129^3 velocity nodes, 65^3
pressure nodes), 27-point Gauss quadrature, Re = 1000 (viscosity 1/1000 with
lid speed 1), dt = 0.0025, 3 inner iterations per step at tolerance 1e-3,
pressure solved to a relative residual of 1e-8, the FP64 mode from the
previous Pdot (the reference's GPU path warm-starts its CG the same way; --x0
zero gives its CPU path's cold start), the mixed and learned modes from the
Z-norm projection of the solution onto the last 16 solutions (--x0-accel).
Inputs are synthetic (generated mesh, zero initial fields plus the lid BC)
and resident in HBM before timing starts.

One "step" is one time step of the reference's timeLoop (BCH:2859-3118):
convection residual, step1, step2 (assembly of R2 + the pressure solve),
step3, convergence checks, K*Acc, for up to 3 inner iterations. Each solver
mode runs the same trajectory from the initial fields: W untimed warm-up
steps, the K timed steps of the contract, then the rest of the 100 steps of
the configuration, so that `trajectory` reports all 100 steps as well:
  fp64     the plain FP64 Jacobi-PCG (the reference's solver, the baseline);
  mixed    FP32 assembly of K, G, G^T and the convection, FP32 inner PCG,
           FP64 refinement residual and stop test;
  learned  learned+mixed, the headline: flexible PCG in FP32 with the
           two-level preconditioner (Jacobi smoothing around the trained
           network), FP64 reliable updates and stop test, Jacobi fallback;
  learned_untrained  the same network at its initial (multilevel) weights;
  direct   sparse Cholesky of Z factored once (the MATLAB prototype's step):
           nested-dissection order, multifrontal factor;
  direct_band  (not run by default) the same with Z factored as a band.
The network is trained before any timing on the pressure solves of seeded
cavity runs (lid speed U[0.5, 2], Re U[100, 1000]) on the same mesh, never on
the timed trajectory; its weights are saved to a file and loaded back for the
timed runs (learned_training names the draws).

`workloads` adds the other configurations on the GPU, each with the CPU
oracle beside it: the 8^3 cavity (Re 100, 50 steps), the 32^3 cavity (Re 400,
200 steps) and the bent duct (16x16 x 160 elements, permuted numbering,
consistent-mass velocity updates, Re 200). `cpu_baseline` is the CPU oracle
(C++/OpenMP restatement of the reference's CPU path) timed over whole time
steps of the headline workload.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cfd-with-cuda_amd"))

METRIC = "timesteps/s, 64^3 hex27 cavity: learned+mixed vs plain FP64 PCG at 1e-8 residual"
TRAJECTORY_STEPS = 100   # the 64^3 configuration's steps
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK = 157.3   # TFLOP/s, dense


def step_kernels(vt, classes=False):
    """rocprofv3 names of the step's row kernels for matrix values of type vt
    ("double", or "float" in the mixed forms, whose step1 also gathers K's
    operands from their FP32 shadows): step1 with K*Acc_prev fused and G*Pn
    kept (inner iterations >= 2), step2's right-hand side, step3; `classes`:
    the mixed forms read G and G^T as row classes."""
    if vt == "float" and classes:
        return ("k_step1<8, 2, 2, true, true, 2, float, true>", "k_step2_rhs<8, 4, float, true, 2>",
                "k_step3<4, 2, float, true>")
    if vt == "float":
        return ("k_step1<8, 2, 2, true, true, 2, float, true>", "k_step2_rhs<32, 1, float, true>",
                "k_step3<4, 1, float>")
    return ("k_step1<8, 2, 2, true, true, 2, double, false>", "k_step2_rhs<16, 8, double, false>",
            "k_step3<8, 2, double>")


ELEM_KERNELS = ("k_elem_step1<true>", "k_step1_nodes<true, 2>")


def elem_flops(ne, ngp):
    """Useful FP32 flops of one k_elem_step1<true> launch: 2 x the multiply-adds
    of its contractions at the real Gauss points and nodes, per element and
    point: field set 0 (Jacobian 8x9, reference gradients 27x9, u0 27x3, K back
    27x9, convection back 27x3) and set 1 (Jacobian, gradients, K back)."""
    macs = (72 + 243 + 81 + 243 + 81) + (72 + 243 + 243)
    return 2.0 * macs * ngp * ne


def zspmv_name(prec, npat, ncls):
    """rocprofv3 name of the Z SpMV kernel the FP64 / mixed solvers run
    (npat: the solvers' row-pattern count, ncls: the FP32 row classes)."""
    if prec and ncls > 0:
        return "k_zcls_spmv<8, false, 2>"
    if npat > 0:   # the row-pattern form runs 8 lanes x 16 loads per row in both precisions
        return "k_pcg_spmv<float, 8, 16, true>" if prec else "k_pcg_spmv<double, 8, 16, true>"
    return "k_pcg_spmv<float, 32, 4, false>" if prec else "k_pcg_spmv<double, 16, 8, false>"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


X0_POLICY = {"zero": 0, "previous": 1, "projection": 2}
GRAPHS = 0   # --graphs


def timed(solver, n):
    """n time steps bracketed by device synchronisations: (stats, wall s)."""
    solver.sync()
    t0 = time.perf_counter()
    st = solver.run(n) if n > 0 else []
    solver.sync()
    return st, time.perf_counter() - t0


def run_mode(solver, mode, steps, warmup, rtol, maxit, matrix_bits=64, x0="zero", total=TRAJECTORY_STEPS):
    """One trajectory in one solver mode from the initial fields: `warmup`
    untimed steps, the `steps` timed ones, then the rest up to `total` steps;
    matrix_bits 32 gives the mixed forms' FP32 assembly (K, G, G^T copies and
    the convection); x0 the pressure solves' initial guess."""
    import numpy as np
    solver.reset(mode=mode, rtol=rtol, maxit=maxit)
    solver.set_graphs(GRAPHS)
    solver.set_solver_x0(X0_POLICY[x0])
    solver.set_matrix_precision(matrix_bits)
    solver.set_convection_precision(matrix_bits)
    solver.timing_reset(True)
    sw, tw = timed(solver, warmup)
    solve_w = solver.timing()["pcg_ms"]
    st, tk = timed(solver, steps)
    solve_k = solver.timing()["pcg_ms"] - solve_w
    Un, Pn = solver.download("Un"), solver.download("Pn")
    rest = max(0, total - warmup - steps)
    sr, tr = timed(solver, rest)
    tm = solver.timing()
    assert len(st) == steps and len(sw) == warmup and len(sr) == rest, "steady state reached early"
    allst = sw + st + sr
    ntraj = len(allst)
    return dict(
        wall_s=tk, steps_per_s=steps / tk, pcg_iters=sum(s["pcg"] for s in st),
        inner_iters=sum(s["inner"] for s in st), max_relres=max(s["relres_max"] for s in st),
        noconv=sum(s["noconv"] for s in st), solve_ms=solve_k, Un=Un, Pn=Pn,
        traj=dict(steps=ntraj, steps_per_s=ntraj / (tw + tk + tr), solve_ms=tm["pcg_ms"],
                  pcg_iters=sum(s["pcg"] for s in allst), inner_iters=sum(s["inner"] for s in allst),
                  max_relres=max(s["relres_max"] for s in allst), noconv=sum(s["noconv"] for s in allst)),
        Un_traj=solver.download("Un"), Pn_traj=solver.download("Pn"))


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
    """Untimed preparation of the learned mode: the network trained on
    seeded cavity runs (fsx_train), its weights saved to a file and loaded
    back; returns the training record and the initial weights."""
    import fsx_train
    n1 = args.n + 1
    rec, w_init = fsx_train.train_seeded(solver, (n1, n1, n1), seed=args.train_seed, runs=args.train_runs,
                                         steps=args.train_steps, train_iters=args.train_iters, lr=args.lr,
                                         target=(1.0, args.re), form=args.learned_form, unet_seed=args.seed,
                                         onpolicy_rounds=args.train_rounds)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "unet.npy")
        meta = fsx_train.save_weights(path, solver.unet_weights(), rec)
        w, meta2 = fsx_train.load_weights(path)
    solver.unet_set_weights(w)
    rec["weights_sha256"] = meta2["sha256"]
    return rec, w, w_init


def cpu_oracle(mesh, steps, threads, ngp, viscosity, dt, rtol, maxit=20000):
    """The CPU oracle (C++/OpenMP restatement of the reference's CPU path:
    cold-started Jacobi-PCG) on `mesh`: setup untimed, then `steps` whole
    time steps timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsx_oracle
    if not os.path.exists(fsx_oracle._LIB_PATH):
        fsx_oracle.build()
    fsx_oracle.set_threads(threads)
    t0 = time.perf_counter()
    o = fsx_oracle.Oracle()
    o.set_mesh(mesh.arrays())
    o.set_params(NGP=ngp, viscosity=viscosity, dt=dt, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=rtol,
                 pcg_maxit=maxit)
    o.setup()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    done = len(o.run(steps))
    t = time.perf_counter() - t0
    tm = o.timings()
    o.close()
    return dict(steps=done, steps_per_s=done / t, wall_s=t, setup_s=t_setup, pcg_iters=tm["pcg_iters"])


def cpu_cores():
    return len(os.sched_getaffinity(0))


def cavity_workload(fsx, N, re, dt, steps, cpu_steps, args, label):
    """The plain FP64 GPU path against the CPU oracle on the N^3 cavity:
    `steps` time steps on each (cpu_steps on the CPU)."""
    mesh = fsx.Mesh.cavity(N).set_params(NGP=args.ngp, dt=dt, maxIter=3, tolerance=1e-3, converge=1e-6,
                                         viscosity=1.0 / re)
    s = fsx.Solver(mesh, device=0, NGP=args.ngp, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=args.rtol,
                   pcg_maxit=args.maxit, solver_mode=0)
    s.reset(mode=fsx.SOLVER_FP64, rtol=args.rtol, maxit=args.maxit)
    s.set_solver_x0(X0_POLICY[args.x0])
    s.timing_reset(True)
    st, t = timed(s, steps)
    out = dict(workload=label, steps=len(st), steps_per_s=len(st) / t, solve_ms=s.timing()["pcg_ms"],
               pcg_iters=sum(x["pcg"] for x in st), max_relres=max(x["relres_max"] for x in st), mode="fp64")
    s.close()
    if cpu_steps > 0 and not args.no_cpu:
        c = cpu_oracle(mesh, cpu_steps, cpu_cores(), args.ngp, 1.0 / re, dt, args.rtol)
        out["cpu"] = dict(c, cores=cpu_cores(), kind="port")
    mesh.close()
    return out


def duct_workload(fsx, args, steps, warmup, cpu_steps):
    """The bent duct: 16 x 16 cross-section elements, 40 + 40 + 80 elements
    along the axis (5-unit inlet, bend of centreline radius 2, 10-unit
    outlet), node numbering permuted from a seed, parabolic inlet of peak 1,
    Re 200, dt 0.005, consistent-mass velocity updates (the three components
    as one 3-RHS Jacobi-PCG per update) with the FP64 and the mixed pressure
    solvers."""
    mesh = fsx.Mesh.duct(n=16, n_in=40, n_arc=40, n_out=80, seed=args.seed)
    mesh.set_params(NGP=args.ngp, dt=0.005, maxIter=3, tolerance=1e-3, converge=1e-6, viscosity=1.0 / 200)
    log("setup bent duct ...")
    s = fsx.Solver(mesh, device=0, NGP=args.ngp, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=args.rtol,
                   pcg_maxit=args.maxit, solver_mode=0)
    s.set_mass(fsx.MASS_CONSISTENT, rtol=1e-10, maxit=2000)
    res = {}
    for name, mode, bits in (("fp64", fsx.SOLVER_FP64, 64), ("mixed", fsx.SOLVER_MIXED, 32)):
        if name not in args.duct_modes.split(","):
            continue
        s.set_mass_precision(bits)   # the mixed forms' mass solves: FP32 segments, FP64 residuals
        r = run_mode(s, mode, steps, warmup, args.rtol, args.maxit, matrix_bits=bits,
                     x0=args.x0 if name == "fp64" else args.x0_accel, total=warmup + steps)
        m_it, m_rel, m_tot = s.mass_stats()
        res[name] = {"steps_per_s": round(r["steps_per_s"], 4), "pcg_iters": r["pcg_iters"],
                     "inner_iters": r["inner_iters"], "solve_ms": round(r["solve_ms"], 2),
                     "max_relres": r["max_relres"], "mass_iters_last_solve": m_it, "mass_relres_last_solve": m_rel}
        log(f"  duct {name}: {res[name]}")
    head = "mixed" if "mixed" in res else "fp64"
    out = {"workload": "90-degree bent square duct 16x16x160 hex27/hex8, permuted numbering, Re 200, dt 0.005, "
                       f"consistent mass, pressure rtol {args.rtol:g}",
           "steps": steps, "warmup": warmup, "value": res[head]["steps_per_s"], "mode": head,
           "NN": s.sizes["NN"], "NNp": s.sizes["NNp"], "nnz_M": s.sizes["nnzM"], "modes": res}
    s.close()
    if cpu_steps > 0 and not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fsx_oracle
        fsx_oracle.set_threads(cpu_cores())
        t0 = time.perf_counter()
        o = fsx_oracle.case_from_mesh(mesh.arrays(), NGP=args.ngp, viscosity=1.0 / 200, dt=0.005, maxIter=3,
                                      tolerance=1e-3, converge=1e-6, pcg_rtol=args.rtol, pcg_maxit=args.maxit)
        # the duct's inlet profile: per-node BC values from the topology, as the GPU run has them
        topo = fsx.Topology(mesh)
        un = o.d("Un")
        nn = o.sizes()["NN"]
        nodes, vals = topo.get("BCnodes"), topo.get("BCvalues").reshape(-1, 3)
        for c in range(3):
            un[c * nn + nodes] = vals[:, c]
        o.set_d("Un", un)
        o.set_mass(1, rtol=1e-10, maxit=2000)
        ts = time.perf_counter() - t0
        t0 = time.perf_counter()
        done = len(o.run(cpu_steps))
        t = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=done / t, unit="timesteps/s", cores=cpu_cores(), kind="port",
                                   sample=f"the same duct on the CPU oracle (C++/OpenMP, {cpu_cores()} threads, "
                                          f"FP64 Jacobi-PCG from zero, consistent mass): {done} whole time steps "
                                          f"timed ({t:.1f} s), setup {ts:.1f} s untimed")
        o.close()
        topo.close()
    mesh.close()
    return out


def headline(args):
    """The 64^3 cavity in every solver mode of --modes: the JSON line's headline
    fields (None after --probe)."""
    import numpy as np
    import fsx

    mesh = fsx.Mesh.cavity(args.n, cluster=args.cluster).set_params(NGP=args.ngp, dt=args.dt, maxIter=3,
                                                                    tolerance=1e-3, converge=1e-6,
                                                                    viscosity=1.0 / args.re)
    log(f"setup {args.n}^3 cavity (cluster {args.cluster:g}) ...")
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
        for kid in (fsx.KERNEL_STEP1, fsx.KERNEL_KACC, fsx.KERNEL_STEP2_RHS, fsx.KERNEL_STEP3,
                    fsx.KERNEL_STEP1_ELEMENTS, fsx.KERNEL_STEP1_NODES):
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
             "learned_untrained": fsx.SOLVER_LEARNED, "direct": fsx.SOLVER_DIRECT, "direct_band": fsx.SOLVER_DIRECT}
    names = args.modes.split(",")
    res = {}
    train = None
    w_trained = w_init = None
    for name in names:
        if name in ("learned", "learned_untrained") and train is None:
            log("learned mode: seeded training runs, prepare and train the network (untimed) ...")
            train, w_trained, w_init = train_learned(solver, args)
            log(f"  {({k: v for k, v in train.items() if k != 'cases'})}")
        if name == "learned":
            solver.unet_set_weights(w_trained)
        elif name == "learned_untrained":
            solver.unet_set_weights(w_init)
        if name in ("direct", "direct_band"):
            solver.set_direct_order(fsx.DIRECT_ND if name == "direct" else fsx.DIRECT_BAND)
        log(f"mode {name}: {args.warmup} warm-up + {args.steps} timed steps, trajectory of {args.trajectory}")
        res[name] = run_mode(solver, modes[name], args.steps, args.warmup, args.rtol, args.maxit,
                             matrix_bits=64 if name in ("fp64", "direct", "direct_band") else 32,
                             x0=args.x0 if name in ("fp64", "direct", "direct_band") else args.x0_accel,
                             total=max(args.trajectory, args.warmup + args.steps))
        r = res[name]
        if name in ("direct", "direct_band"):   # the factor is formed in the first warm-up step
            w, fms, fbytes = solver.direct_stats()
            if name == "direct":
                _, ffl, nfr, depth = solver.direct_info()
                r["extra"] = {"order": "nested dissection", "factor_ms": round(fms, 1),
                              "factor_GB": round(fbytes / 1e9, 3), "factor_flops": ffl, "fronts": nfr, "depth": depth}
            else:
                r["extra"] = {"order": "band", "half_bandwidth": w, "factor_ms": round(fms, 1),
                              "band_GB": round(fbytes / 1e9, 2)}
            log(f"  {name}: {r['extra']}")
        log(f"  timed {r['steps_per_s']:.3f} steps/s, PCG iters {r['pcg_iters']}, solve {r['solve_ms']:.1f} ms; "
            f"trajectory {r['traj']['steps_per_s']:.3f} steps/s, solve {r['traj']['solve_ms']:.1f} ms, "
            f"iters {r['traj']['pcg_iters']}, max relres {r['traj']['max_relres']:.2e}")
    if "learned" in res:
        solver.unet_set_weights(w_trained)
    head = next((m for m in ("learned", "mixed") if m in res), names[0])
    h = res[head]
    prec = 1 if head in ("mixed", "learned") else 0
    # live per-launch device times (HIP events on the context's stream) of the
    # step's row kernels and the Z SpMV of the FP64 / mixed solvers; the
    # roofline entry is the one with the most device time per step in the
    # headline mode
    steps_inner = h["inner_iters"] / args.steps
    solver.set_matrix_precision(32 if prec else 64)   # the headline mode's matrices and convection
    solver.set_convection_precision(32 if prec else 64)
    gcls = solver.g_class_info()
    knames = step_kernels("float" if prec else "double", classes=bool(prec and gcls[0] > 0))
    kern = {}
    # step1 of the mixed forms runs element by element (its K Acc_prev form in
    # inner iterations >= 2): its two passes instead of the CSR k_step1
    step1 = ((fsx.KERNEL_STEP1_ELEMENTS, ELEM_KERNELS[0], steps_inner - 1),
             (fsx.KERNEL_STEP1_NODES, ELEM_KERNELS[1], steps_inner - 1)) if prec else \
        ((fsx.KERNEL_STEP1, knames[0], steps_inner),)
    for kid, kname, per_step in step1 + ((fsx.KERNEL_STEP2_RHS, knames[1], steps_inner),
                                         (fsx.KERNEL_STEP3, knames[2], steps_inner)):
        ms_k, by_k = solver.bench_kernel(kid, reps=20)
        kern[kname] = dict(launch_ms=ms_k, bytes=by_k, ms_per_step=ms_k * per_step)
    if prec:
        kern[ELEM_KERNELS[0]]["flops"] = elem_flops(solver.sizes["NE"], args.ngp)
    ms64, bytes64 = solver.bench_spmv(precision=0, reps=100)
    npat, _ = solver.z_patterns()
    _, ncls, _ = solver.z_forms()
    if "fp64" in res:
        kern[zspmv_name(0, npat, ncls)] = dict(launch_ms=ms64, bytes=bytes64,
                                               ms_per_step=0.0)   # not a kernel of the headline mode
    kname = max(kern, key=lambda k: kern[k]["ms_per_step"])
    kd = kern[kname]
    if "flops" in kd:   # the element pass is bound by the matrix cores (FP32 MFMA)
        roof = {"bound": "mfma", "kernel": kname,
                "achieved": round(kd["flops"] / (kd["launch_ms"] * 1e-3) / 1e12, 2), "peak": FP32_MFMA_PEAK,
                "unit": "TFLOP/s", "flops_per_launch": kd["flops"]}
    else:
        roof = {"bound": "hbm", "kernel": kname,
                "achieved": round(kd["bytes"] / (kd["launch_ms"] * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s"}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["traffic"] = measured_traffic(kname)
    roof["algorithmic_bytes_per_launch"] = kd["bytes"]
    roof["launch_ms"] = round(kd["launch_ms"], 5)
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
        "config": {"workload": f"lid-driven cavity {args.n}^3 hex27/hex8" +
                               (f" (sinh clustering {args.cluster:g})" if args.cluster else "") +
                               f", NGP {args.ngp}, Re {args.re:g}, dt {args.dt:g}, 3 inner iterations, "
                               f"pressure rtol {args.rtol:g}, pressure x0 {args.x0} (fp64), {args.x0_accel} (mixed, learned), timed steps "
                               f"{args.warmup + 1}-{args.warmup + args.steps} of {args.trajectory}",
                   "mode": head, "NN": solver.sizes["NN"], "NNp": solver.sizes["NNp"],
                   "nnz_Z": solver.sizes["nnzZ"], "parallelism": "single GPU",
                   "g_classes": dict(zip(("G", "G_entries", "GT", "GT_entries"), gcls)) if prec else None},
        "roofline": roof,
        "kernels": {k: {"launch_ms": round(v["launch_ms"], 5), "algorithmic_bytes": v["bytes"],
                        "achieved_GBs": round(v["bytes"] / (v["launch_ms"] * 1e-3) / 1e9, 1),
                        "ms_per_step": round(v["ms_per_step"], 4), "traffic": measured_traffic(k),
                        **({"flops": v["flops"],
                            "achieved_TFLOPs": round(v["flops"] / (v["launch_ms"] * 1e-3) / 1e12, 2),
                            "frac_fp32_mfma": round(v["flops"] / (v["launch_ms"] * 1e-3) / 1e12 / FP32_MFMA_PEAK, 4)}
                           if "flops" in v else {})}
                    for k, v in kern.items()},
        "modes": {k: {"steps_per_s": round(v["steps_per_s"], 4), "pcg_iters": v["pcg_iters"],
                      "inner_iters": v["inner_iters"], "solve_ms": round(v["solve_ms"], 2),
                      "max_relres": v["max_relres"], "noconv": v["noconv"], **v.get("extra", {})}
                  for k, v in res.items()},
        "trajectory": {k: {kk: (round(vv, 4) if isinstance(vv, float) and kk != "max_relres" else vv)
                           for kk, vv in v["traj"].items()} for k, v in res.items()},
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
        ft, ht = f["traj"], h["traj"]
        out["trajectory_pressure_solve_speedup"] = round(ft["solve_ms"] / ht["solve_ms"], 3)
        out["trajectory_velocity_rel_err_vs_fp64"] = float(np.linalg.norm(h["Un_traj"] - f["Un_traj"]) /
                                                           np.linalg.norm(f["Un_traj"]))
        out["trajectory_pressure_rel_err_vs_fp64"] = float(np.linalg.norm(h["Pn_traj"] - f["Pn_traj"]) /
                                                           np.linalg.norm(f["Pn_traj"]))
    solver.close()
    if not args.no_cpu:
        log(f"cpu baseline (oracle, {args.cpu_steps} whole time steps) ...")
        c = cpu_oracle(mesh, args.cpu_steps, cpu_cores(), args.ngp, 1.0 / args.re, args.dt, args.rtol)
        out["cpu_baseline"] = dict(value=c["steps_per_s"], unit="timesteps/s", cores=cpu_cores(), kind="port",
                                   sample=(f"the same {args.n}^3 cavity on the CPU oracle (C++/OpenMP restatement of "
                                           f"the reference's CPU path, {cpu_cores()} threads, FP64 Jacobi-PCG from "
                                           f"zero): its first {c['steps']} whole time steps timed "
                                           f"({c['wall_s']:.1f} s, {c['pcg_iters']} PCG iterations), setup "
                                           f"{c['setup_s']:.1f} s untimed"))
        log(f"  {out['cpu_baseline']}")
    else:
        out["cpu_baseline"] = None
    mesh.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--re", type=float, default=1000.0)
    ap.add_argument("--dt", type=float, default=0.0025)
    ap.add_argument("--cluster", type=float, default=0.0,
                    help="sinh clustering of the cavity mesh towards the walls (the reference generator's, 0 uniform)")
    ap.add_argument("--ngp", type=int, default=27)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--maxit", type=int, default=20000)
    ap.add_argument("--modes", default="fp64,mixed,learned,learned_untrained,direct",
                    help="solver modes to run, of fp64, mixed, learned, learned_untrained, direct and direct_band (none: only the "
                         "--workloads)")
    ap.add_argument("--train-iters", type=int, default=800, help="Adam steps of each training round")
    ap.add_argument("--train-rounds", type=int, default=2,
                    help="on-policy training rounds after the first (each re-prepares pairs from the learned "
                         "solver's own truncated iterates)")
    ap.add_argument("--train-runs", type=int, default=6, help="seeded cavity runs the network is trained on")
    ap.add_argument("--train-steps", type=int, default=3, help="time steps of each training run")
    ap.add_argument("--train-seed", type=int, default=2024)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--learned-form", type=int, default=1, choices=(0, 1),
                    help="preconditioner of the learned mode: 1 two-level (Jacobi smoothing around the network, "
                         "tiled Z products), 0 additive (dinv r + network)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--probe", action="store_true",
                    help="only set up the case, run one FP64 time step and one with FP32 matrices and convection, time "
                         "the step kernels and the Z SpMV kernels and run a 256 MiB "
                         "calibration copy (the workload of the rocprofv3 --pmc traffic runs)")
    ap.add_argument("--x0", default="previous", choices=("zero", "previous", "projection"),
                    help="initial guess of the FP64 mode's pressure solves (and of the 8^3 and 32^3 workloads): the "
                         "previous Pdot (the reference's GPU path, CUDAcodes.cu:425-442; default) or zero (its CPU path)")
    ap.add_argument("--x0-accel", default="projection", choices=("zero", "previous", "projection"),
                    help="initial guess of the mixed and learned modes' pressure solves: the Z-norm projection onto "
                         "the last 16 solutions (default), the previous Pdot or zero")
    ap.add_argument("--workloads", default="cavity8,cavity32,duct",
                    help="other configurations to run beside the headline (empty for none)")
    ap.add_argument("--graphs", type=int, default=0, choices=(0, 1),
                    help="enqueue the FP64 and mixed solver iterations as replayed HIP graphs (1) or as individual "
                         "launches (0)")
    ap.add_argument("--duct-modes", default="fp64,mixed", help="pressure solver modes of the bent-duct workload")
    ap.add_argument("--trajectory", type=int, default=TRAJECTORY_STEPS,
                    help="steps of each mode's trajectory (the 64^3 configuration's 100; fewer for profiled runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3, help="whole 64^3 time steps the CPU baseline times")
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = args.graphs
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the solver core is built "
              f"and measured for one GPU; sharding by mesh partitions is designed, not built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import fsx

    out = headline(args) if args.modes != "none" else {"metric": METRIC, "value": None}
    if out is None:   # --probe
        return
    work = {}
    for wname in [w for w in args.workloads.split(",") if w]:
        if wname == "cavity8":
            log("workload: 8^3 cavity, Re 100, dt 0.01, 50 steps")
            work[wname] = cavity_workload(fsx, 8, 100.0, 0.01, 50, 50, args,
                                          "lid-driven cavity 8^3 hex27/hex8, Re 100, dt 0.01, 50 steps")
        elif wname == "cavity32":
            log("workload: 32^3 cavity, Re 400, dt 0.005, 200 steps (GPU and CPU)")
            work[wname] = cavity_workload(fsx, 32, 400.0, 0.005, 200, 200, args,
                                          "lid-driven cavity 32^3 hex27/hex8, Re 400, dt 0.005, 200 steps")
        elif wname == "duct":
            work[wname] = duct_workload(fsx, args, steps=min(args.steps, 20), warmup=min(args.warmup, 5),
                                        cpu_steps=2)
        log(f"  {wname}: {({k: v for k, v in work[wname].items() if k != 'modes'})}")
    if work:
        out["workloads"] = work
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
