"""Failure handling of the pressure solvers through the C-ABI, and uploads of
matrix values that the solvers' own forms must follow.

The reference handles neither failure: MKL's RCI CG prints and returns its
partial answer when it stops at maxit (fractionalStep/Cpp/
blascoCodinaHuerta.cpp:4041-4044) and CUSP's result is never checked
(fractionalStep/Cpp/CUDAcodes.cu:442-444). Here:

  * maxit = 1 returns FSX_ENOCONV from every iterative solver and keeps the
    partial answer, as the reference does: x is a descent step (the energy
    functional 0.5 x'Zx - b'x falls below its value 0 at x = 0) and the
    reported residual is the true ||b - Zx|| / ||b|| (computed by scipy on
    the host from the device's Z and the host's CSR pattern). The time loop
    keeps going with the partial answer and counts the unconverged solves in
    its step statistics instead of hiding them.
  * A NaN in the right-hand side returns FSX_ENUMERIC from every solver, and
    a NaN that reaches step2's operand returns it from fsx_step2; every kernel
    exits normally, so the context solves a clean right-hand side to its
    tolerance straight afterwards.
  * Uploading Z (scaled by 2) rebuilds every form the solvers read: each
    solver then meets ||b - 2Z x|| <= rtol ||b|| for the uploaded Z.
    Uploading M lumps it again (Md = row sums, BCH:3262-3295).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = {"fp64": 0, "mixed": 1, "learned": 2, "direct": 3}


def say(msg):
    print(msg, flush=True)


def make(fsx, mode, N=8):
    mesh = fsx.Mesh.cavity(N).set_params(NGP=27, dt=0.01, maxIter=3, tolerance=1e-3, converge=1e-6,
                                         viscosity=0.01)
    s = fsx.Solver(mesh, device=0, NGP=27, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=1e-10,
                   pcg_maxit=5000, solver_mode=0)
    if mode == MODES["learned"]:
        s.unet_setup((N + 1, N + 1, N + 1), seed=1)
    return mesh, s


def host_z(s):
    import scipy.sparse as sp
    n = s.sizes["NNp"]
    return sp.csr_matrix((s.download("Z"), s.topo.get("Zcol"), s.topo.get("Zrow")), shape=(n, n))


@pytest.mark.parametrize("name", ["fp64", "mixed", "learned"])
def test_maxit_returns_noconv_with_partial_answer(fsx_lib, name):
    mode = MODES[name]
    say(f"maxit = 1: {name} solver, cavity 8^3")
    mesh, s = make(fsx_lib, mode)
    Z = host_z(s)
    b = np.random.default_rng(11).standard_normal(s.sizes["NNp"])
    s.set_solver(mode, 1e-10, 1)
    x, it, relres, rc = s.solve_pressure(b)
    assert rc == fsx_lib.FSX_ENOCONV, rc
    true_rel = np.linalg.norm(b - Z @ x) / np.linalg.norm(b)
    energy = 0.5 * x @ (Z @ x) - b @ x
    say(f"  {it} iterations, reported {relres:.6e}, host {true_rel:.6e}, energy {energy:.3e}")
    assert np.all(np.isfinite(x)) and np.linalg.norm(x) > 0
    assert energy < 0.0
    assert 1 <= it <= 2
    assert abs(relres - true_rel) <= 1e-5 * true_rel
    # the time loop keeps the partial answers and counts them
    s.reset(mode=mode, rtol=1e-10, maxit=1)
    st = s.run(2)
    assert len(st) == 2
    for x_ in st:
        assert x_["noconv"] == x_["inner"] and x_["relres_max"] > 1e-10, x_
    s.close()
    mesh.close()


@pytest.mark.parametrize("name", ["fp64", "mixed", "learned", "direct"])
@pytest.mark.parametrize("x0", [0, 1])
def test_nan_rhs_returns_enumeric(fsx_lib, name, x0):
    mode = MODES[name]
    say(f"NaN right-hand side: {name} solver, x0 policy {x0}, cavity 8^3")
    mesh, s = make(fsx_lib, mode)
    Z = host_z(s)
    n = s.sizes["NNp"]
    s.set_solver(mode, 1e-10, 5000)
    s.set_solver_x0(x0)
    b = np.random.default_rng(12).standard_normal(n)
    bad = b.copy()
    bad[n // 3] = np.nan
    with pytest.raises(fsx_lib.FsxError) as e:
        s.solve_pressure(bad)
    assert e.value.code == fsx_lib.FSX_ENUMERIC, e.value
    # every kernel exited: the same context solves a clean system at once
    s.set_solver_x0(0)
    x, it, relres, rc = s.solve_pressure(b)
    true_rel = np.linalg.norm(b - Z @ x) / np.linalg.norm(b)
    say(f"  clean solve after the error: rc {rc}, {it} iterations, host residual {true_rel:.2e}")
    assert rc == fsx_lib.FSX_OK and true_rel <= 1.05e-10
    s.close()
    mesh.close()


@pytest.mark.parametrize("name", ["fp64", "mixed", "learned"])
def test_nan_in_step2_operand_returns_enumeric(fsx_lib, name):
    mode = MODES[name]
    say(f"NaN in step2's operand: {name} solver, cavity 8^3")
    mesh, s = make(fsx_lib, mode)
    s.reset(mode=mode, rtol=1e-10, maxit=5000)
    s.run(1)
    s.begin_step()
    s.convection()
    s.step1(1)
    u = s.download("UnpHalf")
    u[len(u) // 2] = np.nan
    s.upload("UnpHalf", u)
    with pytest.raises(fsx_lib.FsxError) as e:
        s.step2(1)
    assert e.value.code == fsx_lib.FSX_ENUMERIC, e.value
    # the context is usable: a fresh start runs a clean step
    s.reset(mode=mode, rtol=1e-10, maxit=5000)
    st = s.run(1)
    assert st[0]["relres"] <= 1e-10 and st[0]["noconv"] == 0
    s.close()
    mesh.close()


def test_z_and_m_uploads_rebuild_solver_forms(fsx_lib):
    say("Z upload (x2) and M upload: cavity 8^3, every solver")
    mesh, s = make(fsx_lib, MODES["learned"])
    Z = host_z(s)
    n = s.sizes["NNp"]
    b = np.random.default_rng(13).standard_normal(n)
    s.upload("Z", 2.0 * s.download("Z"))
    Z2 = 2.0 * Z
    for name in ("fp64", "mixed", "learned", "direct"):
        s.set_solver(MODES[name], 1e-10, 5000)
        x, it, relres, rc = s.solve_pressure(b)
        true_rel = np.linalg.norm(b - Z2 @ x) / np.linalg.norm(b)
        say(f"  {name}: rc {rc}, {it} iterations, reported {relres:.2e}, host {true_rel:.2e} on 2Z")
        assert rc == fsx_lib.FSX_OK
        assert true_rel <= 1.05e-10, true_rel
    # the solvers' FP64 form of Z is the uploaded one
    v = np.random.default_rng(14).standard_normal(n)
    assert np.linalg.norm(s.spmv("Zsolver64", v) - Z2 @ v) <= 1e-12 * np.linalg.norm(Z2 @ v)
    # M: the lumped masses follow an upload (row sums of 3M are 3 Md)
    m = s.download("M")
    md0 = s.download("MdOrigInv")
    s.upload("M", 3.0 * m)
    md3 = s.download("MdOrigInv")
    assert np.allclose(md3, md0 / 3.0, rtol=1e-14, atol=0.0)
    s.upload("M", m)
    assert np.array_equal(s.download("MdOrigInv"), md0)
    s.close()
    mesh.close()


@pytest.mark.parametrize("mass", ["lumped", "consistent"])
def test_time_loop_reports_failures(fsx_lib, mass):
    """A failure inside the time loop reaches the caller: with NaN in Un the
    first solve that meets it fails with FSX_ENUMERIC, and fsxd_run returns
    that as a negative status (the Python run() raises it) instead of a step
    count; the context then runs clean steps after a reset."""
    say(f"time loop failure: cavity 8^3, {mass} mass, NaN in Un")
    mesh, s = make(fsx_lib, MODES["fp64"])
    if mass == "consistent":
        s.set_mass(fsx_lib.MASS_CONSISTENT, rtol=1e-10, maxit=500)
    un = s.download("Un")
    un[len(un) // 2] = np.nan
    s.upload("Un", un)
    with pytest.raises(fsx_lib.FsxError) as e:
        s.run(1)
    assert e.value.code == fsx_lib.FSX_ENUMERIC, str(e.value)
    s.reset(mode=MODES["fp64"])
    st = s.run(2)
    assert len(st) == 2 and all(x["pcg"] > 0 and x["relres"] <= 1e-10 for x in st)
    assert np.isfinite(s.download("Un")).all()
    s.close()
