"""step1 of the mixed forms element by element on the matrix cores
(fsx_set_step1_elements, csrc/fsx_elem.hip) against the CSR form of the same
step (K's FP32 values, FP64 products) and against the FP64 reference
arithmetic (blascoCodinaHuerta.cpp:3746-3783 for step1, 3548-3700 for the
convection, K assembled as in 3176-3231).

The element form sums FP32 products of the shape tables with the FP32
shadows of Un, UnpHalf_prev and Acc_prev (and FP32 Jacobians from corner
coordinates relative to the element's first corner); the CSR form rounds K's
values to FP32 and sums FP64 products. Both are FP32-accurate forms of the
same operator, so each is compared with the FP64 step:
  * stage by stage (R1, UnpHalf, KAcc after one inner iteration without and
    one with K Acc_prev): normwise within 1e-5 of the FP64 step and within
    20 times the CSR form's distance from it (measured: R1 1.1-1.4e-7
    against the CSR form's 1.9-3.1e-8 on the cavities, 7.7e-9 against
    8.0e-9 on the duct; KAcc 4-5e-7 against 2-3e-7);
  * fields after 3 steps within the mixed forms' stated accuracy (velocity
    1e-6, pressure 1e-5 against the FP64 run; measured velocity 6.7e-9 to
    1.9e-8), and reproducible bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def say(msg):
    print(msg, flush=True)


CASES = {
    "cavity10": dict(mesh=("cavity", 10, 0.0), NGP=27),
    "cavity8_sinh_ngp8": dict(mesh=("cavity", 8, 1.5), NGP=8),
    "duct": dict(mesh=("duct",), NGP=27),
}


def make(fsx, case):
    c = CASES[case]
    if c["mesh"][0] == "cavity":
        mesh = fsx.Mesh.cavity(c["mesh"][1], cluster=c["mesh"][2])
        visc, dt = 0.01, 0.01
    else:
        mesh = fsx.Mesh.duct(n=4, n_in=6, n_arc=6, n_out=8, seed=3)
        visc, dt = 1.0 / 200, 0.005
    mesh.set_params(NGP=c["NGP"], dt=dt, maxIter=3, tolerance=1e-3, converge=1e-6, viscosity=visc)
    s = fsx.Solver(mesh, device=0, NGP=c["NGP"], maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=1e-10,
                   pcg_maxit=5000, solver_mode=0)
    return s, mesh


def forms(s, bits, elements):
    s.set_matrix_precision(bits)
    s.set_convection_precision(bits)
    s.set_step1_elements(elements)


def stage(s, bits, elements, warm=2):
    """The state after `warm` FP64 steps, then one time step's first two
    inner iterations stage by stage in the given forms: R1, UnpHalf and KAcc
    after the second step1 (the one with K Acc_prev), and R1 / UnpHalf after
    the first."""
    s.reset(mode=0)
    forms(s, 64, 0)
    s.run(warm)
    forms(s, bits, elements)
    s.begin_step()
    s.convection()
    s.step1(1)
    first = (s.download("R1"), s.download("UnpHalf"))
    s.step2(1)
    s.step3(1)
    s.next_iteration()
    s.kacc()
    s.convection()
    s.step1(2)
    return first + (s.download("R1"), s.download("UnpHalf"), s.download("KAcc"))


@pytest.mark.parametrize("case", list(CASES))
def test_step1_elements_stage(fsx_lib, case):
    """One inner iteration without and one with K Acc_prev: the element form
    against the FP64 step, next to the CSR form's distance from it."""
    s, mesh = make(fsx_lib, case)
    ref = stage(s, 64, 0)
    csr = stage(s, 32, 0)
    ele = stage(s, 32, 1)
    names = ("R1 (first)", "UnpHalf (first)", "R1", "UnpHalf", "KAcc")
    errs = []
    for name, r, c, e in zip(names, ref, csr, ele):
        ec, ee = rel(c, r), rel(e, r)
        say(f"  {case} {name}: CSR32 {ec:.2e}, elements {ee:.2e}, elements vs CSR32 {rel(e, c):.2e}")
        assert np.all(np.isfinite(e))
        errs.append((name, ee, ec))
    for name, ee, ec in errs:
        assert ee < 1e-5, (name, ee)
        assert ee <= 20 * ec + 1e-12, (name, ee, ec)
    s.close()
    mesh.close()


@pytest.mark.parametrize("case", list(CASES))
def test_step1_elements_fields(fsx_lib, case):
    """3 time steps in the mixed forms with step1 from the elements: the
    fields against the FP64 run within the stated accuracy of the mixed
    forms, and the runs reproducible bit for bit."""
    s, mesh = make(fsx_lib, case)
    out = {}
    for name, bits, el in (("fp64", 64, 0), ("csr32", 32, 0), ("elements", 32, 1), ("elements again", 32, 1)):
        s.reset(mode=0)
        forms(s, bits, el)
        st = s.run(3)
        assert all(x["relres"] <= 1e-10 for x in st)
        out[name] = (s.download("Un"), s.download("Pn"))
    for name in ("csr32", "elements"):
        eu, ep = rel(out[name][0], out["fp64"][0]), rel(out[name][1], out["fp64"][1])
        say(f"  {case} {name} vs FP64 after 3 steps: velocity {eu:.2e}, pressure {ep:.2e}")
        assert eu < 1e-6 and ep < 1e-5
    assert np.array_equal(out["elements"][0], out["elements again"][0])
    assert np.array_equal(out["elements"][1], out["elements again"][1])
    s.close()
    mesh.close()


def test_step1_elements_deferral(fsx_lib):
    """fsx_convection only records the convection when step1 will form it
    with the element products. Reading R1 in between forms it on its own:
    the R1 read is the element convection (bitwise that of the CSR forms'
    run), and the step1 that follows is then the CSR form's, bitwise. An
    uploaded K (the assembled values, uploaded back) has no element form:
    step1 takes the CSR products and the run is bitwise the CSR run."""
    s, mesh = make(fsx_lib, "cavity10")

    def one(elements, read_between=False, upload_k=False):
        s.reset(mode=0)
        forms(s, 64, 0)
        s.run(1)
        forms(s, 32, elements)
        if upload_k:
            s.upload("K", s.download("K"))
        s.begin_step()
        s.convection()
        r = s.download("R1") if read_between else None
        s.step1(1)
        return r, s.download("R1"), s.download("UnpHalf")

    _, r_csr, u_csr = one(0)
    s.reset(mode=0)
    forms(s, 64, 0)
    s.run(1)
    forms(s, 32, 0)
    s.begin_step()
    s.convection()
    conv_csr = s.download("R1")
    r_between, r_after, u_after = one(1, read_between=True)
    assert np.array_equal(r_between, conv_csr)
    assert np.array_equal(r_after, r_csr) and np.array_equal(u_after, u_csr)
    _, r_el, _ = one(1)
    assert not np.array_equal(r_el, r_csr)   # the element form does run without the read
    _, r_up, u_up = one(1, upload_k=True)
    assert np.array_equal(r_up, r_csr) and np.array_equal(u_up, u_csr)
    s.close()
    mesh.close()


@pytest.mark.parametrize("case", ["cavity10", "duct"])
def test_carried_gpn_equals_formed(fsx_lib, case):
    """With the element form, step3 carries G Pnp1 = G Pn + dt G Pdot into the
    next time step instead of a pass over G at its first step1 (k_pnp1 forms
    Pnp1 = Pn + dt Pdot). After 2 steps, the next step's first R1 with the
    carried G Pn and with G Pn formed afresh (the matrix copies refreshed,
    which drops the carried value) agree to FP64 rounding (bar 1e-13
    normwise)."""
    s, mesh = make(fsx_lib, case)
    out = []
    for fresh in (False, True):
        s.reset(mode=0)
        forms(s, 32, 1)
        s.run(2)
        if fresh:
            s.set_matrix_precision(32)
        s.begin_step()
        s.convection()
        s.step1(1)
        out.append(s.download("R1"))
    e = rel(out[0], out[1])
    say(f"  {case}: carried vs formed G Pn, R1 {e:.2e}")
    assert e < 1e-13
    s.close()
    mesh.close()


@pytest.mark.parametrize("case", list(CASES))
def test_g_classes_against_the_csr_copies(fsx_lib, case):
    """At matrix precision 32, step2's right-hand side and step3 read G and
    G^T as row classes where the rows compress: on the uniform cavity 64
    classes of G and 27 of G^T (the device numbers the velocity nodes in grid
    order, so G^T's classes carry the column offsets); on the stretched
    cavity and the curved duct
    the class lists would hold over an eighth of G's entries, the form is
    not taken and the runs are the CSR copies' bit for bit. One inner
    iteration's R2, Pdot and Acc against the CSR copies' within 1e-6
    normwise (the same FP32 values summed in another order, residue
    components below 2^-24 of their row's largest dropped), and the fields
    after 3 steps within the mixed forms' bar of the FP64 run."""
    s, mesh = make(fsx_lib, case)
    out = {}
    for on in (False, True):
        s.reset(mode=0)
        forms(s, 64, 0)
        s.run(2)
        forms(s, 32, 1)
        s.set_matrix_classes(on)
        info = s.g_class_info()
        s.begin_step()
        s.convection()
        s.step1(1)
        s.step2(1)
        r2, pdot = s.download("R2"), s.download("Pdot")
        s.step3(1)
        acc = s.download("Acc")
        s.reset(mode=0)
        st = s.run(3)
        assert all(x["relres"] <= 1e-10 for x in st)
        out[on] = (info, r2, pdot, acc, s.download("Un"), s.download("Pn"))
    s.reset(mode=0)
    forms(s, 64, 0)
    s.run(3)
    ref = (s.download("Un"), s.download("Pn"))
    s.set_matrix_classes(True)
    assert out[False][0] == (0, 0, 0, 0)
    if case == "cavity10":
        assert out[True][0][0] == 64 and out[True][0][2] == 27, out[True][0]
        for name, a, b in zip(("R2", "Pdot", "Acc"), out[True][1:4], out[False][1:4]):
            e = rel(a, b)
            say(f"  {case} {name}: classes vs CSR copies {e:.2e}")
            assert e < 1e-6, (name, e)
    else:
        assert out[True][0] == (0, 0, 0, 0)
        for a, b in zip(out[True][1:], out[False][1:]):
            assert np.array_equal(a, b)
    eu, ep = rel(out[True][4], ref[0]), rel(out[True][5], ref[1])
    say(f"  {case} classes vs FP64 after 3 steps: velocity {eu:.2e}, pressure {ep:.2e}")
    assert eu < 1e-6 and ep < 1e-5
    s.close()
    mesh.close()
