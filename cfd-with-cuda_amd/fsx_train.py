"""Training of the learned pressure preconditioner on seeded cavity runs.

The learned preconditioner (fsx_unet_* in include/fsx.h) is trained on pairs
(right-hand side, solution) of the pressure solves of lid-driven cavity runs
whose lid speed and Reynolds number are drawn from a seeded generator: lid
speed uniform in [0.5, 2], Re uniform in [100, 1000], viscosity lid / Re on
the unit cube. The runs use the mesh of the case the network is then used
on, but not that case's own trajectory: a timed run at (lid 1, Re 1000) sees
right-hand sides the network was not trained on.

    python fsx_train.py --n 64 --runs 6 --steps 3 --out unet_cavity64.npy

writes the FP32 parameter vector (numpy .npy, in the layout of
fsx_unet_get_weights) and a JSON sidecar that names the draws, the training
settings and the losses; load_weights() reads both back.
"""
import argparse
import hashlib
import json
import os
import time

import numpy as np

import fsx

LID_RANGE = (0.5, 2.0)
RE_RANGE = (100.0, 1000.0)


def draw_cases(seed, n, lid_range=LID_RANGE, re_range=RE_RANGE):
    """n seeded (lid speed, Re) draws, uniform in the given ranges."""
    rng = np.random.default_rng(seed)
    lids = rng.uniform(*lid_range, size=n)
    res = rng.uniform(*re_range, size=n)
    return [(float(a), float(b)) for a, b in zip(lids, res)]


def train_seeded(solver, grid, seed=2024, runs=6, steps=3, ks=(0, 5, 20, 50, 100), train_iters=800, batch=4,
                 lr=1e-4, target=(1.0, 1000.0), form=1, unet_seed=7, random_pairs=0, random_sweeps=0,
                 onpolicy_rounds=2, onpolicy_ks=(2, 4, 7, 10), offpolicy=False):
    """Sets up the network on `solver`'s pressure grid and trains it on the
    pressure solves of `runs` seeded cavity runs of `steps` time steps each,
    in the mixed form (FP32 matrices and convection, as the learned mode
    runs), then puts the solver back on the target case (lid, Re) at its
    initial fields. Returns (training record, initial weights); the trained
    weights stay in the solver (solver.unet_weights()).

    The training pairs are the residuals of truncated solves of the captured
    right-hand sides from zero: by default those of the learned solver itself
    (onpolicy_ks iterations deep, with the network as it is at the time), first
    with the initial network and then again after each of `onpolicy_rounds`
    training rounds, appended, so that each round sees the modes the network
    it trained last handles worst. offpolicy=True takes the first pairs from
    FP64 Jacobi-PCG truncated at `ks` instead; random_pairs adds pairs of
    drawn errors (fsx_unet_prepare_random) with random_sweeps smoothing
    sweeps."""
    t0 = time.perf_counter()
    solver.unet_form(form)
    solver.unet_setup(grid, seed=unet_seed)
    w_init = solver.unet_weights()
    cases = draw_cases(seed, runs)
    per_run = steps * max(1, int(solver.opt.maxIter))
    solver.unet_capture(per_run * runs)
    for lid, re in cases:
        solver.set_case(viscosity=lid / re, lid=lid)
        solver.reset(mode=fsx.SOLVER_MIXED)
        solver.set_matrix_precision(32)
        solver.set_convection_precision(32)
        solver.run(steps)
    ncap = solver.unet_captured()
    solver.unet_capture(0)
    lid, re = target
    solver.set_case(viscosity=lid / re, lid=lid)
    if offpolicy:
        nsamp = solver.unet_prepare(list(ks))
    else:
        nsamp = solver.unet_prepare(list(onpolicy_ks), learned=True)
    if random_pairs > 0:
        nsamp = solver.unet_prepare_random(random_pairs, sweeps=random_sweeps, seed=seed + 2)
    loss0 = solver.unet_train(1, batch=8, lr=1e-12, seed=seed)
    loss = solver.unet_train(train_iters, batch=batch, lr=lr, seed=seed + 1)
    for rnd in range(onpolicy_rounds):
        # residuals the current network leaves (the learned solver's own
        # truncated iterates) hold the modes it handles worst
        nsamp = solver.unet_prepare(list(onpolicy_ks), learned=True, append=True)
        loss = solver.unet_train(train_iters, batch=batch, lr=lr, seed=seed + 3 + rnd)
    solver.sync()
    return dict(train_s=time.perf_counter() - t0, seed=seed, cases=[dict(lid=a, Re=b) for a, b in cases],
                steps_per_run=steps, captured=ncap, samples=nsamp, truncations=list(ks), loss_initial=loss0,
                loss_final=loss, train_iters=train_iters, batch=batch, lr=lr, params=solver.unet_nparams(),
                form=form, gain=solver.unet_gain(), target=dict(lid=target[0], Re=target[1]),
                random_pairs=random_pairs, random_sweeps=random_sweeps, onpolicy_rounds=onpolicy_rounds,
                onpolicy_ks=list(onpolicy_ks), offpolicy=offpolicy), w_init


def save_weights(path, w, meta):
    """weights as .npy (float32) and meta as <path>.json with their SHA-256"""
    w = np.ascontiguousarray(w, dtype=np.float32)
    np.save(path, w)
    meta = dict(meta, sha256=hashlib.sha256(w.tobytes()).hexdigest(), nparams=int(w.size))
    with open(path + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    return meta


def load_weights(path):
    """(weights, meta) as save_weights wrote them; the digest is checked"""
    w = np.load(path, allow_pickle=False).astype(np.float32)
    with open(path + ".json") as f:
        meta = json.load(f)
    if hashlib.sha256(w.tobytes()).hexdigest() != meta.get("sha256"):
        raise ValueError(f"{path}: weights do not match their recorded digest")
    return w, meta


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=64, help="cavity elements per axis (the pressure grid has n+1 nodes)")
    ap.add_argument("--runs", type=int, default=6)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--train-iters", type=int, default=800)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--form", type=int, default=1, choices=(0, 1))
    ap.add_argument("--dt", type=float, default=0.0025)
    ap.add_argument("--rounds", type=int, default=2, help="on-policy training rounds after the first")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    mesh = fsx.Mesh.cavity(a.n).set_params(NGP=27, dt=a.dt, maxIter=3, tolerance=1e-3, converge=1e-6,
                                           viscosity=1e-3)
    s = fsx.Solver(mesh, device=0, NGP=27, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=1e-8,
                   pcg_maxit=20000, solver_mode=0)
    n1 = a.n + 1
    rec, _ = train_seeded(s, (n1, n1, n1), seed=a.seed, runs=a.runs, steps=a.steps, train_iters=a.train_iters,
                          lr=a.lr, form=a.form, onpolicy_rounds=a.rounds)
    meta = save_weights(a.out, s.unet_weights(), dict(rec, grid=[n1] * 3))
    print(json.dumps(meta), flush=True)
    s.close()


if __name__ == "__main__":
    main()
