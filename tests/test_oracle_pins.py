"""Pins of the CPU oracle (oracle/fsx_oracle.cpp) by closed forms, invariants
and brute force, independent of the oracle's own code.

The reference ships no tests or golden outputs (its only fixture is the 30^3
input mesh, checked in test_host_topology.py), so every expected value here
comes from mathematics:
  * 1-D quadratic/linear Lagrange integrals in exact rational arithmetic; on
    an axis-aligned box element the hex27 matrices are Kronecker products of
    them (textbook P2 mass h/30[[4,2,-1],[2,16,2],[-1,2,4]] and stiffness
    1/(3h)[[7,-8,1],[-8,16,-8],[1,-8,7]]);
  * invariants: K 1 = 0, sum(M) = |Omega|, lumped mass = products of
    h/6[1,4,1], G^T 1 sums, Z symmetric with the penalty, A(u0) const = 0,
    convection of the x field by u0 = (1,0,0) = -lumped mass;
  * brute force: dense assembly from the element matrices, dense direct solves.
A plausible mistake (a dropped term, a sign, a transposed Jacobian, a wrong
Gauss weight, a wrong node order) changes one of these.
"""
from fractions import Fraction as F
import itertools

import numpy as np
import pytest

import meshgen

# parametric position of each hex27 node, BCH:2312-2343
POS = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1),
       (0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1),
       (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0),
       (0, -1, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1),
       (0, 0, -1), (0, -1, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, 0, 1), (0, 0, 0)]


# --- exact 1-D polynomial algebra on [-1, 1] (coefficient lists, low first)
def pmul(a, b):
    out = [F(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def pder(a):
    return [i * a[i] for i in range(1, len(a))] or [F(0)]


def pint(a):  # integral over [-1, 1]
    return sum(c * F(2, k + 1) for k, c in enumerate(a) if k % 2 == 0)


QUAD = {-1: [F(0), F(-1, 2), F(1, 2)], 0: [F(1), F(0), F(-1)], 1: [F(0), F(1, 2), F(1, 2)]}
LIN = {-1: [F(1, 2), F(-1, 2)], 1: [F(1, 2), F(1, 2)]}


def one_d(h):
    """Exact 1-D element matrices on an interval of length h (J = h/2)."""
    j = F(h) / 2
    keys = (-1, 0, 1)
    M = {(a, b): pint(pmul(QUAD[a], QUAD[b])) * j for a in keys for b in keys}
    K = {(a, b): pint(pmul(pder(QUAD[a]), pder(QUAD[b]))) / j for a in keys for b in keys}
    N = {(a, b): pint(pmul(QUAD[a], LIN[b])) * j for a in keys for b in (-1, 1)}          # int La Lb_lin
    D = {(a, b): pint(pmul(pder(QUAD[a]), LIN[b])) for a in keys for b in (-1, 1)}         # int La' Lb_lin
    return M, K, N, D


def test_textbook_1d_p2():
    M, K, _, _ = one_d(F(1))
    order = (-1, 0, 1)
    Mt = [[F(4), F(2), F(-1)], [F(2), F(16), F(2)], [F(-1), F(2), F(4)]]
    Kt = [[F(7), F(-8), F(1)], [F(-8), F(16), F(-8)], [F(1), F(-8), F(7)]]
    for i, a in enumerate(order):
        for k, b in enumerate(order):
            assert M[(a, b)] == Mt[i][k] / 30
            assert K[(a, b)] == Kt[i][k] / 3


def box_element_matrices(hx, hy, hz, nu=1.0, rho=1.0):
    """Exact hex27 Me, Ke, Ge_d of an axis-aligned box element (full integration)."""
    Mx, Kx, Nx, Dx = one_d(hx)
    My, Ky, Ny, Dy = one_d(hy)
    Mz, Kz, Nz, Dz = one_d(hz)
    Me = np.zeros((27, 27)); Ke = np.zeros((27, 27)); Ge = np.zeros((3, 27, 8))
    for i, (a, b, c) in enumerate(POS):
        for j, (d, e, f) in enumerate(POS):
            Me[i, j] = float(Mx[a, d] * My[b, e] * Mz[c, f])
            Ke[i, j] = float(F(nu) * (Kx[a, d] * My[b, e] * Mz[c, f] + Mx[a, d] * Ky[b, e] * Mz[c, f]
                                      + Mx[a, d] * My[b, e] * Kz[c, f]))
        for j, (d, e, f) in enumerate(POS[:8]):
            Ge[0, i, j] = float(-Dx[a, d] * Ny[b, e] * Nz[c, f] / F(rho))
            Ge[1, i, j] = float(-Nx[a, d] * Dy[b, e] * Nz[c, f] / F(rho))
            Ge[2, i, j] = float(-Nx[a, d] * Ny[b, e] * Dz[c, f] / F(rho))
    return Me, Ke, Ge


def dense_from_elements(o, Me, Ke, Ge):
    s = o.sizes()
    NN, NNp, NE = s["NN"], s["NNp"], s["NE"]
    L = o.i("LtoGnode").reshape(NE, 27)
    M = np.zeros((NN, NN)); K = np.zeros((NN, NN)); G = np.zeros((3, NN, NNp))
    for e in range(NE):
        M[np.ix_(L[e], L[e])] += Me
        K[np.ix_(L[e], L[e])] += Ke
        for d in range(3):
            G[d][np.ix_(L[e], L[e][:8])] += Ge[d]
    return M, K, G


def csr_dense(rowptr, col, val, shape):
    A = np.zeros(shape)
    for r in range(shape[0]):
        A[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    return A


@pytest.fixture(scope="module")
def box2(oracle_lib):
    """Uniform 2^3 cavity (h = 1/2), NGP 27, nu = 0.01, rho = 1."""
    return oracle_lib.case_from_mesh(meshgen.cavity(2), NGP=27, viscosity=0.01, dt=0.01)


def test_assembly_closed_form_ngp27(box2):
    o = box2
    s = o.sizes()
    Me, Ke, Ge = box_element_matrices(F(1, 2), F(1, 2), F(1, 2), nu=F(1, 100))
    M, K, G = dense_from_elements(o, Me, Ke, Ge)
    NN, NNp = s["NN"], s["NNp"]
    Mo = csr_dense(o.i("Mrow"), o.i("Mcol"), o.d("M"), (NN, NN))
    Ko = csr_dense(o.i("Mrow"), o.i("Mcol"), o.d("K"), (NN, NN))
    assert np.abs(Mo - M).max() < 1e-15 * np.abs(M).max() * 100
    assert np.abs(Ko - K).max() < 1e-13 * np.abs(K).max()
    for d, name in enumerate(("G1", "G2", "G3")):
        Go = csr_dense(o.i("Grow"), o.i("Gcol"), o.d(name), (NN, NNp))
        assert np.abs(Go - G[d]).max() < 1e-13 * np.abs(G[d]).max(), name


def test_assembly_sinh_box_elements(oracle_lib):
    """Non-uniform (sinh-clustered) box elements: per-element closed forms."""
    mesh = meshgen.cavity(4, cluster=1.7)
    o = oracle_lib.case_from_mesh(mesh, NGP=27, viscosity=0.02, density=2.0, dt=0.01)
    s = o.sizes()
    NN, NNp, NE = s["NN"], s["NNp"], s["NE"]
    L = o.i("LtoGnode").reshape(NE, 27)
    xyz = mesh["coord"]
    K = np.zeros((NN, NN)); G1 = np.zeros((NN, NNp))
    for e in range(NE):
        c = xyz[L[e][:8]]
        h = c[6] - c[0]
        _, Ke, Ge = box_element_matrices(F(h[0]), F(h[1]), F(h[2]), nu=F(0.02), rho=F(2))
        K[np.ix_(L[e], L[e])] += Ke
        G1[np.ix_(L[e], L[e][:8])] += Ge[0]
    Ko = csr_dense(o.i("Mrow"), o.i("Mcol"), o.d("K"), (NN, NN))
    G1o = csr_dense(o.i("Grow"), o.i("Gcol"), o.d("G1"), (NN, NNp))
    assert np.abs(Ko - K).max() < 1e-12 * np.abs(K).max()
    assert np.abs(G1o - G1).max() < 1e-12 * np.abs(G1).max()


def test_two_point_rule_is_under_integrated(oracle_lib, box2):
    """NGP = 8 differs from the exact matrices for K and M but gives the same
    lumped mass (the 1-D row sums are integrals of quadratics, exact for the
    2-point rule)."""
    o8 = oracle_lib.case_from_mesh(meshgen.cavity(2), NGP=8, viscosity=0.01, dt=0.01)
    assert np.abs(o8.d("M") - box2.d("M")).max() > 1e-6
    assert np.abs(o8.d("K") - box2.d("K")).max() > 1e-6
    assert np.allclose(o8.d("MdOrigInv"), box2.d("MdOrigInv"), rtol=1e-14, atol=0)
    # and NGP = 8 equals the Kronecker products of the 2-point 1-D matrices
    a = 1 / np.sqrt(3.0)
    pts, w = np.array([-a, a]), np.array([1.0, 1.0])
    q = {-1: lambda x: 0.5 * (x * x - x), 0: lambda x: 1 - x * x, 1: lambda x: 0.5 * (x * x + x)}
    m1 = {(i, j): 0.25 * np.sum(w * q[i](pts) * q[j](pts)) for i in (-1, 0, 1) for j in (-1, 0, 1)}
    Me = np.array([[m1[a_, d] * m1[b, e] * m1[c, f] for (d, e, f) in POS] for (a_, b, c) in POS])
    M, _, _ = dense_from_elements(o8, Me, np.zeros((27, 27)), np.zeros((3, 27, 8)))
    s = o8.sizes()
    Mo = csr_dense(o8.i("Mrow"), o8.i("Mcol"), o8.d("M"), (s["NN"], s["NN"]))
    assert np.abs(Mo - M).max() < 1e-15 * 100 * np.abs(M).max()


def test_invariants(box2):
    o = box2
    s = o.sizes()
    NN, NNp = s["NN"], s["NNp"]
    one = np.ones(NN)
    assert np.abs(o.spmv("K", one)).max() < 1e-14                        # K 1 = 0
    assert abs(o.d("M").sum() - 1.0) < 1e-14                              # sum M = |Omega|
    # lumped mass = products of the 1-D row sums h/6 [1, 4, 1] (summed over elements)
    xyz = o.d("coord").reshape(-1, 3)
    w1 = lambda t: np.where(np.isclose(t * 4 % 2, 1), 4 / 12, np.where(np.isclose(t, 0) | np.isclose(t, 1), 1 / 12, 2 / 12))
    md = w1(xyz[:, 0]) * w1(xyz[:, 1]) * w1(xyz[:, 2])
    assert np.allclose(1.0 / o.d("MdOrigInv")[:NN], md, rtol=1e-14, atol=0)
    # G_d^T 1 sums to -(1/rho) int d_d(sum Sv) Sp = 0 per pressure node; G_d 1 = -int d_d Sv_i:
    # its total over all rows is -int d_d (1) = 0, and sum_i x_i (G_1 1)_i = -|Omega|
    for d, name in enumerate(("G1", "G2", "G3")):
        g1 = o.spmv(name, np.ones(NNp))
        assert abs(g1.sum()) < 1e-14
        assert abs(np.dot(xyz[:, d], g1) + 1.0) < 1e-14
    # Z symmetric, with the penalty x1000 at the zero-pressure node only
    Z = csr_dense(o.i("Zrow"), o.i("Zcol"), o.d("Z"), (NNp, NNp))
    zp = s["zeroP"]
    Zs = Z.copy()
    Zs[zp, zp] /= 1000.0
    assert np.abs(Zs - Zs.T).max() < 1e-12 * np.abs(Zs).max()
    # Z = sum_d G_d^T MdOrigInv G_d (dense brute force)
    minv = o.d("MdOrigInv")[:NN]
    Zd = sum(csr_dense(o.i("Grow"), o.i("Gcol"), o.d(n), (NN, NNp)).T @ (minv[:, None] *
             csr_dense(o.i("Grow"), o.i("Gcol"), o.d(n), (NN, NNp))) for n in ("G1", "G2", "G3"))
    assert np.abs(Zs - Zd).max() < 1e-12 * np.abs(Zd).max()
    # positive definite with the pin
    assert np.linalg.eigvalsh(Z).min() > 0


def test_pattern_sizes_closed_form(oracle_lib):
    for N in (1, 2, 3, 5):
        o = oracle_lib.case_from_mesh(meshgen.cavity(N), NGP=8)
        s = o.sizes()
        assert s["NN"] == (2 * N + 1) ** 3
        assert s["nnzM"] == (8 * N + 1) ** 3
        assert s["nnzG"] == (5 * N + 1) ** 3
        if N >= 2:
            assert s["nnzZ"] == (5 * N - 1) ** 3
        # sorted, unique columns per row
        for rp, ci in ((o.i("Mrow"), o.i("Mcol")), (o.i("Grow"), o.i("Gcol")), (o.i("Zrow"), o.i("Zcol"))):
            for r in range(len(rp) - 1):
                c = ci[rp[r]:rp[r + 1]]
                assert np.all(np.diff(c) > 0)


def test_node_positions(oracle_lib):
    """Every element's 27 nodes sit at the parametric positions of BCH:2312-2343."""
    mesh = meshgen.cavity(4, cluster=1.2)
    o = oracle_lib.case_from_mesh(mesh, NGP=8)
    s = o.sizes()
    L = o.i("LtoGnode").reshape(s["NE"], 27)
    xyz = o.d("coord").reshape(-1, 3)
    for e in range(s["NE"]):
        c = xyz[L[e][:8]]
        lo, hi = c[0], c[6]
        for i, p in enumerate(POS):
            expect = lo + (np.array(p) + 1) / 2 * (hi - lo)
            assert np.allclose(xyz[L[e][i]], expect, atol=1e-14)
    assert np.array_equal(L[:, :8].ravel(), mesh["conn"].ravel())


def test_shape_functions(oracle_lib):
    o = oracle_lib.case_from_mesh(meshgen.cavity(1), NGP=27)
    Sv = o.d("Sv").reshape(27, 27)
    dSv = o.d("dSv").reshape(3, 27, 27)
    assert np.allclose(Sv.sum(axis=1), 1.0, atol=1e-15)
    assert np.abs(dSv.sum(axis=1)).max() < 1e-14
    w = o.d("gqw")
    p = o.d("gqp").reshape(27, 3)
    assert abs(w.sum() - 8.0) < 1e-14
    # the 3-point rule integrates x^4 y^2 exactly: (2/5)(2/3)(2)
    assert abs(np.sum(w * p[:, 0] ** 4 * p[:, 1] ** 2) - 2 / 5 * 2 / 3 * 2) < 1e-14


def test_convection_closed_forms(oracle_lib):
    mesh = meshgen.cavity(4, cluster=0.8)
    o = oracle_lib.case_from_mesh(mesh, NGP=27)
    s = o.sizes()
    NN = s["NN"]
    rng = np.random.default_rng(0)
    un = rng.standard_normal(3 * NN)
    # A(u0) applied to a constant field vanishes
    r = o.convection(un, np.ones(3 * NN))
    assert np.abs(r).max() < 1e-13
    # u0 = (1, 0, 0), uprev = (x, y, z): (u0 . grad) x = 1 -> R1_u = -int Sv_i = -Md
    xyz = o.d("coord").reshape(-1, 3)
    u0 = np.concatenate([np.ones(NN), np.zeros(NN), np.zeros(NN)])
    up = np.concatenate([xyz[:, 0], xyz[:, 1], xyz[:, 2]])
    r = o.convection(u0, up)
    md = 1.0 / o.d("MdOrigInv")[:NN]
    assert np.allclose(r[:NN], -md, rtol=1e-12, atol=1e-15)
    assert np.abs(r[NN:]).max() < 1e-14
    # bilinear in (u0, uprev): linear in uprev
    a, b = rng.standard_normal(3 * NN), rng.standard_normal(3 * NN)
    lhs = o.convection(un, 2 * a - 3 * b)
    rhs = 2 * o.convection(un, a) - 3 * o.convection(un, b)
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(rhs).max()


def test_convection_brute_force_single_element(oracle_lib):
    """Dense Ae by direct numpy quadrature on one box element."""
    mesh = meshgen.cavity(1)
    mesh["coord"] = mesh["coord"] * np.array([0.5, 0.8, 1.3])
    o = oracle_lib.case_from_mesh(mesh, NGP=27)
    NN = 27
    L = o.i("LtoGnode")
    rng = np.random.default_rng(3)
    un, up = rng.standard_normal(3 * NN), rng.standard_normal(3 * NN)
    g = np.array([-np.sqrt(0.6), 0, np.sqrt(0.6)])
    gw = np.array([5 / 9, 8 / 9, 5 / 9])
    h = np.array([0.5, 0.8, 1.3])
    q = {-1: (lambda x: 0.5 * (x * x - x), lambda x: x - 0.5), 0: (lambda x: 1 - x * x, lambda x: -2 * x),
         1: (lambda x: 0.5 * (x * x + x), lambda x: x + 0.5)}
    Ae = np.zeros((27, 27))
    for (i, x), (j, y), (k, z) in itertools.product(enumerate(g), enumerate(g), enumerate(g)):
        N = np.array([q[a][0](x) * q[b][0](y) * q[c][0](z) for (a, b, c) in POS])
        dN = np.array([[q[a][1](x) * q[b][0](y) * q[c][0](z) * 2 / h[0],
                        q[a][0](x) * q[b][1](y) * q[c][0](z) * 2 / h[1],
                        q[a][0](x) * q[b][0](y) * q[c][1](z) * 2 / h[2]] for (a, b, c) in POS])
        wk = gw[i] * gw[j] * gw[k] * np.prod(h) / 8
        u0 = np.array([N @ un[L + c * NN] for c in range(3)])
        Ae += wk * np.outer(N, dN @ u0)
    r = o.convection(un, up)
    for c in range(3):
        expect = np.zeros(NN)
        expect[L] = -(Ae @ up[L + c * NN])
        assert np.allclose(r[c * NN:(c + 1) * NN], expect, rtol=1e-12, atol=1e-14)


def test_pcg_matches_direct_solve(box2, oracle_lib):
    o = box2
    n = o.sizes()["NNp"]
    Z = csr_dense(o.i("Zrow"), o.i("Zcol"), o.d("Z"), (n, n))
    b = np.random.default_rng(5).standard_normal(n)
    x, it, rel = o.pcg_z(b, rtol=1e-12, maxit=1000)
    xd = np.linalg.solve(Z, b)
    assert np.linalg.norm(b - Z @ x) <= 1.01e-12 * np.linalg.norm(b)
    assert np.allclose(x, xd, rtol=1e-8, atol=1e-10 * np.abs(xd).max())
    # generic CSR entry on an SPD tridiagonal system: exact in n steps
    m = 20
    rp = np.arange(0, 3 * m - 1, 3).astype(np.int32)
    rows, cols, vals = [], [], []
    for i in range(m):
        for j in (i - 1, i, i + 1):
            if 0 <= j < m:
                rows.append(i); cols.append(j); vals.append(2.0 + i if i == j else -1.0)
    rp = np.searchsorted(rows, np.arange(m + 1)).astype(np.int32)
    bb = np.ones(m)
    x, it, rel = oracle_lib.pcg_csr(rp, cols, vals, bb, rtol=1e-14, maxit=100)
    A = csr_dense(rp, np.array(cols), np.array(vals), (m, m))
    assert np.allclose(A @ x, bb, atol=1e-12)
    # zero right-hand side: zero iterations, zero solution (BCH x0 = 0)
    x, it, rel = o.pcg_z(np.zeros(n))
    assert it == 0 and not x.any()


def test_zero_lid_gives_zero_fields(oracle_lib):
    o = oracle_lib.case_from_mesh(meshgen.cavity(3, lid=(0.0, 0.0, 0.0)), NGP=8)
    st = o.run(3)
    assert not o.d("Un").any() and not o.d("Pn").any()
    assert np.all(st[:, 0] == 3)     # 0/0 norms never satisfy the inner test (BCH:2958)


def test_cavity_mirror_symmetry(oracle_lib):
    """The lid drives along x, so the flow is mirror-symmetric about y = 1/2:
    u(x, 1-y, z) = u, v -> -v, w -> w, p -> p."""
    o = oracle_lib.case_from_mesh(meshgen.cavity(4), NGP=27, pcg_rtol=1e-13)
    o.run(3)
    s = o.sizes()
    NN = s["NN"]
    xyz = o.d("coord").reshape(-1, 3)
    key = {tuple(np.round(p, 9)): i for i, p in enumerate(xyz)}
    mirror = np.array([key[(round(p[0], 9), round(1 - p[1], 9), round(p[2], 9))] for p in xyz])
    u = o.d("Un")
    p = o.d("Pn")
    scale = np.abs(u).max()
    assert np.abs(u[:NN] - u[:NN][mirror]).max() < 1e-9 * scale
    assert np.abs(u[NN:2 * NN] + u[NN:2 * NN][mirror]).max() < 1e-9 * scale
    assert np.abs(u[2 * NN:] - u[2 * NN:][mirror]).max() < 1e-9 * scale
    pm = mirror[:s["NNp"]]
    assert np.abs(p - p[pm]).max() < 1e-7 * np.abs(p).max()


def test_step_invariants(oracle_lib):
    """One inner iteration: BC rows keep their values, the pressure solve
    satisfies Z Pdot = R2, and the velocity correction is -dt^2 MdInv G Pdot
    on the free rows. (Z is built with the BC-free lumped mass, BCH:3394, while
    the correction vanishes at BC rows, so G^T Unp1 is not driven to zero.)"""
    o = oracle_lib.case_from_mesh(meshgen.cavity(4), NGP=27, pcg_rtol=1e-12)
    o.set_d("UnpHalf_prev", o.d("Un"))
    o.step1(1)
    o.step2(1)
    o.step3(1)
    s = o.sizes()
    NN, n = s["NN"], s["NNp"]
    bcn = o.i("BCnodes")
    for name in ("UnpHalf", "Unp1"):
        v = o.d(name)
        for c in range(3):
            assert np.array_equal(v[c * NN + bcn], o.d("Un")[c * NN + bcn])
    Z = csr_dense(o.i("Zrow"), o.i("Zcol"), o.d("Z"), (n, n))
    assert np.linalg.norm(Z @ o.d("Pdot") - o.d("R2")) <= 1.01e-12 * np.linalg.norm(o.d("R2"))
    free = np.ones(NN, dtype=bool)
    free[bcn] = False
    minv = o.d("MdInv")[:NN]
    for c, g in enumerate(("G1", "G2", "G3")):
        corr = o.d("Unp1")[c * NN:(c + 1) * NN] - o.d("UnpHalf")[c * NN:(c + 1) * NN]
        gp = o.spmv(g, o.d("Pdot"))
        dt = 0.01
        assert np.allclose(corr[free], -dt * dt * minv[free] * gp[free], rtol=1e-12, atol=1e-18)


def test_consistent_mass_step(oracle_lib):
    """Consistent-mass mode (no reference counterpart; the reference lumps M):
    step1 sets UnpHalf = Un + dt a with M_bc a = R1 and step3 sets Acc with
    M_bc Acc = R3, where M_bc is M with BC rows and columns replaced by the
    identity. Checked against a dense direct solve built here from the
    oracle's CSR M and BC list, and at the BC nodes (a = 0, as with the
    lumped update)."""
    o = oracle_lib.case_from_mesh(meshgen.cavity(3), NGP=27, pcg_rtol=1e-12)
    o.set_mass(1, rtol=1e-13, maxit=500)
    s = o.sizes()
    NN = s["NN"]
    dt = 0.01
    M = csr_dense(o.i("Mrow"), o.i("Mcol"), o.d("M"), (NN, NN))
    bcn = o.i("BCnodes")
    Mbc = M.copy()
    Mbc[bcn, :] = 0.0
    Mbc[:, bcn] = 0.0
    Mbc[bcn, bcn] = 1.0
    o.set_d("UnpHalf_prev", o.d("Un"))
    o.step1(1)
    it, rel, _ = o.mass_stats()
    assert 0 < it < 500 and rel <= 1e-13
    R1 = o.d("R1")
    a = (o.d("UnpHalf") - o.d("Un")) / dt
    for c in range(3):
        ad = np.linalg.solve(Mbc, R1[c * NN:(c + 1) * NN])
        assert np.allclose(a[c * NN:(c + 1) * NN], ad, rtol=1e-9, atol=1e-9 * np.abs(ad).max())
        assert not ad[bcn].any()
    # the consistent update differs from the lumped one (M is not diagonal)
    lumped = R1 * o.d("MdInv")
    assert np.abs(a - lumped).max() > 1e-3 * np.abs(lumped).max()
    o.step2(1)
    o.step3(1)
    R3 = o.d("R3")
    acc = o.d("Acc")
    for c in range(3):
        ad = np.linalg.solve(Mbc, R3[c * NN:(c + 1) * NN])
        assert np.allclose(acc[c * NN:(c + 1) * NN], ad, rtol=1e-9, atol=1e-9 * np.abs(ad).max())
    assert np.allclose(o.d("Unp1"), o.d("UnpHalf") + dt * acc, rtol=0, atol=1e-15)


def test_warm_start_pcg(oracle_lib):
    """Pressure PCG from the previous Pdot (the reference's GPU path keeps it,
    CU:425-442): the same fields as the cold start (both to rtol 1e-12), fewer
    iterations in total, and each solve's true residual |R2 - Z Pdot| within
    rtol |R2| (recomputed here from the oracle's Z)."""
    cold = oracle_lib.case_from_mesh(meshgen.cavity(4), NGP=27, pcg_rtol=1e-12, pcg_maxit=5000)
    warm = oracle_lib.case_from_mesh(meshgen.cavity(4), NGP=27, pcg_rtol=1e-12, pcg_maxit=5000)
    warm.set_pcg_x0(1)
    sc = cold.run(3)
    sw = warm.run(3)
    for a in ("Un", "Pn"):
        x, y = warm.d(a), cold.d(a)
        assert np.linalg.norm(x - y) <= 1e-9 * np.linalg.norm(y), a
    assert sw[:, 1].sum() < sc[:, 1].sum()
    n = warm.sizes()["NNp"]
    Z = csr_dense(warm.i("Zrow"), warm.i("Zcol"), warm.d("Z"), (n, n))
    r2, pd = warm.d("R2"), warm.d("Pdot")
    assert np.linalg.norm(r2 - Z @ pd) <= 1.01e-12 * np.linalg.norm(r2)


def test_f32_operands(oracle_lib):
    """Operands rounded to FP32 (the oracle's model of the device's FP32
    shadows): step1's R1 moves by exactly -K (fl32(v) - v) per component,
    v = UnpHalf_prev, except at BC rows, which step1 zeroes; step2's R2 moves
    by exactly sum_d G_d^T (fl32(e_d) - e_d), e = UnpHalf / dt^2 - MdOrigInv
    KAcc, except at the pinned pressure row; with consistent mass the option
    does nothing."""
    dt = 0.01
    c = oracle_lib.case_from_mesh(meshgen.cavity(3), NGP=8, dt=dt)
    NN = c.sizes()["NN"]
    v = np.random.default_rng(5).standard_normal(3 * NN)
    c.set_d("UnpHalf_prev", v)
    c.step1(1)
    r0 = c.d("R1").copy()
    c.set_f32_operands(1)
    c.set_d("UnpHalf_prev", v)
    c.step1(1)
    r1 = c.d("R1").copy()
    K = csr_dense(c.i("Mrow"), c.i("Mcol"), c.d("K"), (NN, NN))
    dv = (v.astype(np.float32).astype(np.float64) - v).reshape(3, NN)
    want = -(K @ dv.T).T.reshape(-1)
    for n in c.i("BCnodes"):
        want[[n, n + NN, n + 2 * NN]] = 0.0
    assert np.linalg.norm(want) > 0
    assert np.linalg.norm((r1 - r0) - want) <= 1e-6 * np.linalg.norm(want)
    # step2: R2 = sum_d G_d^T e_d
    rng = np.random.default_rng(6)
    u, k = rng.standard_normal(3 * NN), rng.standard_normal(3 * NN)
    r2 = []
    for f in (0, 1):
        c.set_f32_operands(f)
        c.set_d("UnpHalf", u)
        c.set_d("KAcc", k)
        c.step2(2)
        r2.append(c.d("R2").copy())
    NNp = c.sizes()["NNp"]
    e = u / dt**2 - c.d("MdOrigInv") * k
    de = (e.astype(np.float32).astype(np.float64) - e).reshape(3, NN)
    want2 = sum(csr_dense(c.i("Grow"), c.i("Gcol"), c.d(f"G{d + 1}"), (NN, NNp)).T @ de[d] for d in range(3))
    pinned = (r2[0] == 0.0) & (r2[1] == 0.0)   # the zero-pressure row (BCH:4180-4184)
    assert pinned.sum() <= 1
    want2[pinned] = 0.0
    assert np.linalg.norm(want2) > 0
    assert np.linalg.norm((r2[1] - r2[0]) - want2) <= 1e-6 * np.linalg.norm(want2)
    c.set_f32_operands(1)
    c.set_mass(1, 1e-12, 1000)
    c.set_d("UnpHalf_prev", v)
    c.step1(1)
    rm1 = c.d("R1").copy()
    c.set_f32_operands(0)
    c.set_d("UnpHalf_prev", v)
    c.step1(1)
    assert np.array_equal(rm1, c.d("R1"))


def test_direct_pressure_solve(oracle_lib):
    """The oracle's Cholesky pressure solve (the MATLAB prototype's step,
    ML:1876-1885, 2021-2030) against a dense LAPACK solve of its own Z on a
    4^3 cavity; its half-bandwidth is the largest |i - j| of Z's pattern; a
    time step with it agrees with one whose PCG runs to rtol 1e-13."""
    c = oracle_lib.case_from_mesh(meshgen.cavity(4), NGP=8)
    n = c.sizes()["NNp"]
    Z = csr_dense(c.i("Zrow"), c.i("Zcol"), c.d("Z"), (n, n))
    b = np.random.default_rng(11).standard_normal(n)
    x, rel = c.direct_solve(b)
    want = np.linalg.solve(Z, b)
    assert np.linalg.norm(x - want) <= 1e-11 * np.linalg.norm(want)
    assert rel < 1e-13
    rows = np.repeat(np.arange(n), np.diff(c.i("Zrow")))
    assert c.chol_bandwidth() == np.abs(rows - c.i("Zcol")).max()
    a = oracle_lib.case_from_mesh(meshgen.cavity(4), NGP=8, pcg_rtol=1e-13, pcg_maxit=5000)
    d = oracle_lib.case_from_mesh(meshgen.cavity(4), NGP=8, pcg_rtol=1e-13, pcg_maxit=5000)
    d.set_direct(1)
    a.run(2)
    sd = d.run(2)
    assert (sd[:, 1] == 0).all()
    for name in ("Un", "Pn"):
        assert np.linalg.norm(d.d(name) - a.d(name)) <= 1e-10 * np.linalg.norm(a.d(name)), name


def _grid_pattern(nx, ny, nz, radius):
    """CSR of the (2 radius + 1)^3-point stencil on an nx*ny*nz grid (x fastest)"""
    import scipy.sparse as sp
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols = [], []
    rr = range(-radius, radius + 1)
    for dz in rr:
        for dy in rr:
            for dx in rr:
                a = idx[max(0, -dz):nz - max(0, dz), max(0, -dy):ny - max(0, dy), max(0, -dx):nx - max(0, dx)]
                b = idx[max(0, dz):nz - max(0, -dz) if dz < 0 else nz, :, :][:a.shape[0]]
                b = idx[max(0, dz):max(0, dz) + a.shape[0], max(0, dy):max(0, dy) + a.shape[1],
                        max(0, dx):max(0, dx) + a.shape[2]]
                rows.append(a.ravel())
                cols.append(b.ravel())
    m = sp.csr_matrix((np.ones(sum(r.size for r in rows)), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    m.sum_duplicates()
    m.sort_indices()
    return m


@pytest.mark.parametrize("order", ["grid", "nd", "random"])
def test_symbolic_cholesky_counts_the_numeric_fill(oracle_lib, order):
    """The oracle's symbolic Cholesky (column counts by row subtrees of the
    elimination tree) against the structural nonzeros of a numeric Cholesky
    of an SPD matrix with that pattern (random positive weights, diagonally
    dominant, so no entry of L cancels to zero), on a 6x5x4 grid with the
    pressure operator's radius-2 coupling, for three orders."""
    m = _grid_pattern(6, 5, 4, 2)
    n = m.shape[0]
    rng = np.random.default_rng(3)
    if order == "grid":
        perm = np.arange(n, dtype=np.int32)
    elif order == "nd":
        perm = oracle_lib.nested_dissection_grid(6, 5, 4, radius=2, leaf=8)
        assert np.array_equal(np.sort(perm), np.arange(n))
    else:
        perm = rng.permutation(n).astype(np.int32)
    a = m.toarray()
    w = np.triu(rng.uniform(0.1, 1.0, (n, n)), 1)
    a = a * (w + w.T)
    np.fill_diagonal(a, a.sum(axis=1) + 1.0)
    ap = a[np.ix_(perm, perm)]
    L = np.linalg.cholesky(ap)
    nnz_num = int(np.count_nonzero(np.abs(L) > 1e-300))
    nnz, flops, h = oracle_lib.symbolic_cholesky(m.indptr, m.indices, perm)
    assert nnz == nnz_num
    cc = np.count_nonzero(np.abs(L) > 1e-300, axis=0)
    assert flops == float(np.sum(cc.astype(np.float64) ** 2))
    assert 1 <= h <= n


def test_nested_dissection_fills_less_than_the_band(oracle_lib):
    """On a 17^3 grid with radius-2 coupling the nested-dissection order's
    factor is smaller and cheaper than the grid (band) order's, whose column
    counts are its half-bandwidth (2 nx ny + 2 nx + 2) plus one."""
    nx = 17
    m = _grid_pattern(nx, nx, nx, 2)
    n = m.shape[0]
    band = oracle_lib.symbolic_cholesky(m.indptr, m.indices, np.arange(n, dtype=np.int32))
    nd = oracle_lib.symbolic_cholesky(m.indptr, m.indices, oracle_lib.nested_dissection_grid(nx, nx, nx))
    w = 2 * nx * nx + 2 * nx + 2
    assert band[0] <= n * (w + 1)
    assert nd[0] < 0.75 * band[0] and nd[1] < 0.6 * band[1]   # measured 0.68 and 0.56
