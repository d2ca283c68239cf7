"""GPU checks of the learned pressure preconditioner (fsx_unet_* C-ABI).

The U-net kernels are compared with the numpy oracle (oracle/unet_ref.py) in
its device arithmetic model (BF16 weights and MFMA operands, FP32
elsewhere). The oracle sums in float64 while the device sums in FP32 in
another order, and a BF16 rounding of an intermediate value can land on the
other side of a tie, so the bars are normwise relative errors:
  * forward output: 2e-3;
  * loss: 1e-3; parameter gradient, per layer: 1e-2.
Scaling the input by 2 commutes with every rounding step, so that check is
bit-exact. The learned solves must meet the same FP64 stop test as the
Jacobi solver (||b - Z x|| <= 1e-8 ||b||, checked here with the FP64 Z).
"""
import numpy as np
import pytest

import unet_ref as U

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def cavity_solver(fsx, N, ngp=8, mode=1, rtol=1e-8, steps=0, dt=0.01, viscosity=0.01):
    mesh = fsx.Mesh.cavity(N)
    mesh.set_params(NGP=ngp, dt=dt, maxIter=3, tolerance=1e-3, converge=1e-6, viscosity=viscosity)
    s = fsx.Solver(mesh, device=0, NGP=ngp, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=rtol,
                   pcg_maxit=20000, solver_mode=mode)
    if steps:
        s.run(steps)
    return s


@pytest.mark.parametrize("N", [8, 16])
def test_unet_forward_matches_oracle(fsx_lib, N):
    say(f"unet forward: grid {N + 1}^3")
    s = cavity_solver(fsx_lib, N)
    g = (N + 1,) * 3
    s.unet_setup(g, seed=11)
    w = s.unet_weights().astype(np.float64)
    x = np.random.default_rng(N).normal(size=(N + 1) ** 3).astype(np.float32)
    y = s.unet_apply(x)
    gain = s.unet_gain()
    y_ref, _ = U.forward(x, w, g, gain=gain, exact=False)
    e = rel(y, y_ref)
    say(f"  rel err {e:.2e}")
    assert e < 2e-3
    # scaling by 2 commutes with every BF16/FP32 rounding: bit-exact
    assert np.array_equal(s.unet_apply(2 * x), 2 * y)
    # weights round-trip, and other weights change the output as the oracle says
    w2 = U.init_weights(5).astype(np.float32)
    s.unet_set_weights(w2)
    assert np.array_equal(s.unet_weights(), w2)
    y2 = s.unet_apply(x)
    assert rel(y2, U.forward(x, w2.astype(np.float64), g, gain=gain, exact=False)[0]) < 2e-3
    s.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_unet_gradient_matches_oracle(fsx_lib, kind):
    say(f"unet loss gradient: grid 9^3, loss kind {kind}")
    s = cavity_solver(fsx_lib, 8)
    g = (9, 9, 9)
    s.unet_setup(g, seed=3)
    gain = float(np.float32(0.6))
    s.unet_gain(gain)
    w = s.unet_weights().astype(np.float64)
    rng = np.random.default_rng(1)
    x = rng.normal(size=729).astype(np.float32)
    t = rng.normal(size=729).astype(np.float32)
    loss, grad = s.unet_loss_grad(x, t, kind=kind)
    if kind == 0:
        loss_ref, grad_ref, _ = U.loss_grad(x, t, w, g, gain=gain, exact=False)
    elif kind == 1:
        # energy norm on the context's FP64 Z (the device trains on its FP32 copy)
        y, cache = U.forward(x, w, g, gain=gain, exact=False)
        t64 = t.astype(np.float64)
        tt = float(t64 @ s.spmv("Z", t64))
        zd = s.spmv("Z", y - t64)
        loss_ref = float((y - t64) @ zd) / tt
        grad_ref = U.backward(2.0 * zd / tt, cache)
    else:
        # energy after post-smoothing, f = (I - S Z)(y - t), S = omega D^-1:
        # the loss of f' Z f is quadratic in y with gradient 2 (I - Z S) Z f
        y, cache = U.forward(x, w, g, gain=gain, exact=False)
        t64 = t.astype(np.float64)
        Z = host_z(s)
        dinv = (1.0 / Z.diagonal()).astype(np.float32).astype(np.float64)
        sw = float(np.float32(s.unet_form_get()[1])) * dinv
        tt = float(t64 @ (Z @ t64))
        f = (y - t64) - sw * (Z @ (y - t64))
        zf = Z @ f
        loss_ref = float(f @ zf) / tt
        grad_ref = U.backward(2.0 * (zf - Z @ (sw * zf)) / tt, cache)
        # the same gradient by central differences of the loss along y
        dy = np.random.default_rng(2).normal(size=y.size)

        def lossy(yy):
            ff = (yy - t64) - sw * (Z @ (yy - t64))
            return float(ff @ (Z @ ff)) / tt
        h = 1e-3 * np.linalg.norm(y) / np.linalg.norm(dy)
        fd = (lossy(y + h * dy) - lossy(y - h * dy)) / (2 * h)
        assert abs(fd - 2.0 * (zf - Z @ (sw * zf)) @ dy / tt) <= 1e-6 * abs(fd) + 1e-12
    say(f"  loss {loss:.6f} oracle {loss_ref:.6f}")
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    offs, total = U.layer_offsets()
    assert grad.size == total
    for li, ((cin, cout, _), off) in enumerate(zip(U.CHANNELS, offs)):
        n = cout * U.kpad(cin)
        e = rel(grad[off:off + n], grad_ref[off:off + n])
        say(f"  layer {U.LAYER_NAMES[li]}: rel err {e:.2e}")
        assert e < 1e-2
    s.close()


def true_relres(s, b, x):
    return float(np.linalg.norm(b - s.spmv("Z", x)) / np.linalg.norm(b))


def test_learned_solve_trains_and_beats_jacobi(fsx_lib):
    say("learned solve: cavity 16^3, NGP 8, capture 3 steps, train, solve a later right-hand side")
    s = cavity_solver(fsx_lib, 16, steps=0)
    s.unet_setup((17, 17, 17), seed=7)
    s.unet_capture(64)
    s.run(3)
    ncap = s.unet_captured()
    s.unet_capture(0)
    s.run(1)
    b = s.download("R2")
    assert ncap >= 3
    nsamp = s.unet_prepare([0, 2, 5, 10, 20, 40])
    assert ncap <= nsamp <= 6 * ncap   # iterates converged below 1e-5 |b| are skipped
    l0 = s.unet_train(1, batch=16, lr=1e-9, seed=1)
    l1 = s.unet_train(400, batch=4, lr=1e-3, seed=2)
    say(f"  captured {ncap}, samples {nsamp}, loss {l0:.4f} -> {l1:.4f}, gain {s.unet_gain():.3e}")
    assert l1 < l0
    s.set_solver(1, 1e-8, 20000)
    xj, itj, relj, rc = s.solve_pressure(b)
    assert rc == 0
    s.set_solver(2, 1e-8, 20000)
    xl, itl, rell, rc = s.solve_pressure(b)
    lit, fb = s.unet_stats()
    rj, rl = true_relres(s, b, xj), true_relres(s, b, xl)
    say(f"  jacobi(mixed) {itj} it relres {rj:.2e}; learned {itl} it (flexible {lit}, fallback {fb}) relres {rl:.2e}")
    assert rc == 0
    assert rl <= 1.0001e-8
    assert itl < itj
    # both solutions solve the same system to 1e-8
    assert rel(xl, xj) < 1e-5
    s.close()


def test_learned_fallback_and_zero_rhs(fsx_lib):
    say("learned fallback: cavity 8^3, untrained network, flexible CG capped at 3 iterations")
    s = cavity_solver(fsx_lib, 8, steps=1)
    b = s.download("R2")
    s.unet_setup((9, 9, 9), seed=2)
    s.unet_limit(3)
    s.set_solver(2, 1e-8, 20000)
    x, it, r, rc = s.solve_pressure(b)
    lit, fb = s.unet_stats()
    rr = true_relres(s, b, x)
    say(f"  flexible {lit} it, fallback {fb}, total {it} it, relres {rr:.2e}")
    # the untrained network loses descent or hits the cap; Jacobi-PCG finishes
    assert rc == 0 and fb in (2, 6) and 1 <= lit <= 3
    assert rr <= 1.0001e-8
    x0, it0, r0, rc0 = s.solve_pressure(np.zeros_like(b))
    assert rc0 == 0 and it0 == 0 and not np.any(x0)
    s.close()


def test_learned_time_steps_match_jacobi(fsx_lib):
    say("learned time steps: cavity 8^3, NGP 8, 3 steps learned vs Jacobi (rtol 1e-10)")
    a = cavity_solver(fsx_lib, 8, rtol=1e-10, mode=0)
    a.run(3)
    ua, pa = a.download("Un"), a.download("Pn")
    a.close()
    s = cavity_solver(fsx_lib, 8, rtol=1e-10)
    s.unet_setup((9, 9, 9), seed=4)
    s.unet_capture(16)
    s.run(2)
    s.unet_capture(0)
    s.unet_prepare([0, 3, 8])
    s.unet_train(100, batch=4, lr=1e-3, seed=3)
    s.reset(mode=2, rtol=1e-10, maxit=20000)
    s.run(3)
    ub, pb = s.download("Un"), s.download("Pn")
    eu, ep = rel(ub, ua), rel(pb, pa)
    say(f"  rel diff U {eu:.2e} P {ep:.2e}")
    assert eu < 1e-6 and ep < 1e-6
    s.close()


def host_z(s):
    import scipy.sparse as sp
    n = s.sizes["NNp"]
    return sp.csr_matrix((s.download("Z"), s.topo.get("Zcol"), s.topo.get("Zrow")), shape=(n, n))


@pytest.mark.parametrize("N", [8, 16])
def test_two_level_preconditioner_matches_oracle(fsx_lib, N):
    """z = M r of the two-level form (the tiled Z products around the
    network) and of the additive form against the oracle's restatement in the
    device arithmetic model, on the context's Z: 2e-3 normwise."""
    say(f"two-level preconditioner: grid {N + 1}^3")
    s = cavity_solver(fsx_lib, N)
    g = (N + 1,) * 3
    s.unet_setup(g, seed=13)
    form, omega = s.unet_form_get()
    assert form == 1 and abs(omega - 0.8) < 1e-7
    w = s.unet_weights().astype(np.float64)
    gain = s.unet_gain()
    Z = host_z(s)
    dinv = (1.0 / Z.diagonal()).astype(np.float32).astype(np.float64)
    r = np.random.default_rng(N + 3).normal(size=(N + 1) ** 3).astype(np.float32)
    z = s.unet_precondition(r)
    ref = U.precondition_two_level(r.astype(np.float64), w, g, gain, float(np.float32(omega)), dinv, lambda v: Z @ v,
                                   exact=False)
    e = rel(z, ref)
    say(f"  two-level rel err {e:.2e}")
    assert e < 2e-3
    s.unet_form(0)
    assert s.unet_form_get()[0] == 0
    za = s.unet_precondition(r)
    ea = rel(za, U.precondition_additive(r.astype(np.float64), w, g, gain, dinv, exact=False))
    say(f"  additive rel err {ea:.2e}")
    assert ea < 2e-3
    s.close()


def test_two_level_solve_beats_additive_and_jacobi(fsx_lib):
    """On a 16^3 cavity right-hand side with the untrained (multilevel)
    network: the two-level flexible CG needs fewer iterations than the
    additive one, which needs fewer than Jacobi-PCG; every solve meets
    ||b - Z x|| <= 1e-8 ||b|| on the FP64 Z; a warm start from a nearby x0
    takes fewer iterations and meets the same test."""
    say("two-level solve: cavity 16^3, NGP 8")
    s = cavity_solver(fsx_lib, 16, steps=2)
    b = s.download("R2")
    Z = host_z(s)
    s.unet_setup((17, 17, 17), seed=7)
    s.set_solver(1, 1e-8, 20000)
    xj, itj, _, rc = s.solve_pressure(b)
    assert rc == 0
    res = {}
    for form in (0, 1):
        s.unet_form(form)
        s.set_solver(2, 1e-8, 20000)
        x, it, relres, rc = s.solve_pressure(b)
        lit, fb = s.unet_stats()
        tr = float(np.linalg.norm(b - Z @ x) / np.linalg.norm(b))
        say(f"  form {form}: {it} it (flexible {lit}, fallback {fb}), device {relres:.2e}, host {tr:.2e}")
        assert rc == 0 and fb == 0 and tr <= 1.0001e-8
        assert rel(x, xj) < 1e-5
        res[form] = it
    say(f"  jacobi {itj}")
    assert res[1] < res[0] < itj
    # warm start (the previous solve's answer, perturbed) on the device
    s.set_solver_x0(1)
    s.upload("Pdot", xj * (1.0 + 1e-3))
    s.set_solver(2, 1e-8, 20000)
    b2 = s.download("R2")
    x2, it2, rel2, rc2 = s.solve_pressure(b2)
    tr2 = float(np.linalg.norm(b2 - Z @ x2) / np.linalg.norm(b2))
    say(f"  warm start: {it2} it, host {tr2:.2e}")
    assert rc2 == 0 and tr2 <= 1.0001e-8 and it2 < res[1]
    s.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_graph_replay_matches_eager(fsx_lib, mode):
    """Solver iterations replayed from captured HIP graphs give bit-identical
    iterates to individual launches, on every replay (several graph launches
    per solve, three solves in a row with eager work in between)."""
    say(f"graph replay vs eager: cavity 16^3, solver mode {mode}")
    s = cavity_solver(fsx_lib, 16, steps=1)
    b = s.download("R2")
    if mode == 2:
        s.unet_setup((17, 17, 17), seed=5)
        s.unet_form(0)   # graphs are the additive form's (the tiled form launches directly)
    s.set_solver(mode, 1e-8, 20000)
    s.set_graphs(False)
    xe, ite, _, rce = s.solve_pressure(b)
    s.set_graphs(True)
    for rep in range(3):
        xg, itg, _, rcg = s.solve_pressure(b)
        assert rcg == rce == 0 and itg == ite, (rep, itg, ite)
        assert np.array_equal(xg, xe), rep
        s.spmv("Z", xg)   # eager work between replays
    s.close()


def test_random_pairs_append_and_train(fsx_lib):
    """fsx_unet_prepare_random appends pairs of the form in use, training on
    them gives a finite loss below the initial one, and pairs of another form
    are refused."""
    say("random training pairs: grid 17^3")
    s = cavity_solver(fsx_lib, 16)
    s.unet_setup((17, 17, 17), seed=5)
    assert s.unet_prepare_random(6, sweeps=4, seed=1) == 6
    assert s.unet_prepare_random(4, sweeps=0, seed=2) == 10
    # the loss over the same 64 seeded draws of the pairs before and after
    # training (an Adam step of lr 1e-12 leaves the weights as they are)
    l0 = s.unet_train(1, batch=64, lr=1e-12, seed=0)
    s.unet_train(100, batch=4, lr=1e-4, seed=1)
    l1 = s.unet_train(1, batch=64, lr=1e-12, seed=0)
    say(f"  loss {l0:.4f} -> {l1:.4f}")
    assert np.isfinite(l1) and 0 < l1 < l0
    s.unet_form(0)
    with pytest.raises(fsx_lib.FsxError):
        s.unet_prepare_random(2, sweeps=0, seed=3)
    s.close()


@pytest.mark.parametrize("case", ["uniform", "sinh"])
def test_tensor_tiles_against_class_tiles(fsx_lib, case):
    """On a box grid the tiles apply Z as Kronecker products of 1D stencils
    (fsx_set_z_tensor; the form reproduces Z's rows to 1e-12 here). The
    two-level preconditioner with it against the one with the FP32 row
    classes: 1e-5 normwise (FP32 sums in another order), and the learned
    solve meets ||b - Z x|| <= 1e-8 ||b|| on the FP64 Z in the same number
    of iterations give or take one. (The sinh-clustered 16^3 grid still
    forms row classes; at 64^3 it does not, and there only the tensor form
    gives the tiles.)"""
    N = 16
    mesh = fsx_lib.Mesh.cavity(N, cluster=2.0 if case == "sinh" else 0.0)
    mesh.set_params(NGP=27, dt=0.01, maxIter=3, tolerance=1e-3, converge=1e-6, viscosity=0.01)
    s = fsx_lib.Solver(mesh, device=0, NGP=27, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=1e-8,
                       pcg_maxit=20000, solver_mode=1)
    s.run(2)
    b = s.download("R2")
    Z = host_z(s)
    s.unet_setup((N + 1,) * 3, seed=7)
    used, err = s.z_tensor_info()
    say(f"  {case}: tensor form in use {used}, row mismatch {err:.1e}")
    assert used and 0.0 <= err < 1e-12
    r = np.random.default_rng(5).normal(size=(N + 1) ** 3).astype(np.float32)
    zt = s.unet_precondition(r)
    s.set_solver(2, 1e-8, 20000)
    xt, itt, _, rc = s.solve_pressure(b)
    assert rc == 0 and float(np.linalg.norm(b - Z @ xt) / np.linalg.norm(b)) <= 1.0001e-8
    s.set_z_tensor(False)
    assert not s.z_tensor_info()[0]
    s.set_solver(2, 1e-8, 20000)
    xc, itc, _, rc = s.solve_pressure(b)
    assert rc == 0 and float(np.linalg.norm(b - Z @ xc) / np.linalg.norm(b)) <= 1.0001e-8
    say(f"  iterations: tensor tiles {itt}, class tiles {itc}")
    zc = s.unet_precondition(r)
    e = rel(zt, zc.astype(np.float64))
    say(f"  preconditioner, tensor vs class tiles: {e:.2e}")
    assert e < 1e-5
    assert abs(itt - itc) <= 1
    s.set_z_tensor(True)
    assert s.z_tensor_info()[0]
    s.close()
    mesh.close()


def test_tensor_form_not_on_the_duct(fsx_lib):
    """The bent duct's pressure nodes form no box grid: no tensor form."""
    mesh = fsx_lib.Mesh.duct(n=4, n_in=6, n_arc=6, n_out=8, seed=3)
    mesh.set_params(NGP=27, dt=0.005, maxIter=3, tolerance=1e-3, converge=1e-6, viscosity=1.0 / 200)
    s = fsx_lib.Solver(mesh, device=0, NGP=27, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=1e-8,
                       pcg_maxit=20000, solver_mode=1)
    assert s.z_tensor_info() == (False, -1.0)
    s.close()
    mesh.close()
