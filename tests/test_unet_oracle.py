"""Pins of the learned-preconditioner oracle (oracle/unet_ref.py) on the CPU.

Each check is chosen so that a plausible slip in the restatement (a wrong
tap offset, a transposed weight block, a missing skip connection, a
mis-weighted prolongation) fails it: brute-force loops, exactness on affine
fields, dot-product (adjoint) identities, central finite differences, and
torch's own float32 -> bfloat16 rounding.
"""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402


def brute_conv(a, W, stride, out_shape):
    nz, ny, nx = out_shape
    gz, gy, gx, cin = a.shape
    out = np.zeros((nz, ny, nx, W.shape[2]))
    for k, j, i in itertools.product(range(nz), range(ny), range(nx)):
        for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
            t = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)
            z, y, x = stride * k + dz, stride * j + dy, stride * i + dx
            if 0 <= z < gz and 0 <= y < gy and 0 <= x < gx:
                out[k, j, i] += a[z, y, x] @ W[t]
    return out


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_matches_brute_force(stride):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(5, 5, 9, 2))
    W = rng.normal(size=(27, 2, 3))
    out_shape = (5, 5, 9) if stride == 1 else (3, 3, 5)
    np.testing.assert_allclose(U.conv(a, W, stride, out_shape), brute_conv(a, W, stride, out_shape), rtol=1e-12, atol=1e-12)


def test_upsample_exact_on_affine_fields():
    g1, g0 = (3, 5, 4), (5, 9, 7)
    zc, yc, xc = np.meshgrid(np.arange(3) * 2.0, np.arange(5) * 2.0, np.arange(4) * 2.0, indexing="ij")
    zf, yf, xf = np.meshgrid(np.arange(5) * 1.0, np.arange(9) * 1.0, np.arange(7) * 1.0, indexing="ij")
    f = lambda x, y, z: 0.3 + 1.5 * x - 0.7 * y + 2.25 * z  # noqa: E731
    c = f(xc, yc, zc)[..., None]
    np.testing.assert_allclose(U.up(c, g0)[..., 0], f(xf, yf, zf), rtol=0, atol=1e-12)
    # injection at even nodes, and a constant stays constant
    assert np.array_equal(U.up(c, g0)[::2, ::2, ::2], c)
    del g1


def test_upsample_trilinear_cell_centre():
    c = np.zeros((2, 2, 2, 1))
    c[1, 1, 1, 0] = 8.0
    f = U.up(c, (3, 3, 3))
    assert f[1, 1, 1, 0] == 1.0   # centre of the cell: 1/8 of each corner
    assert f[2, 2, 1, 0] == 4.0   # edge midpoint: 1/2
    assert f[2, 1, 1, 0] == 2.0   # face centre: 1/4


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_adjoint_dot_product(stride):
    rng = np.random.default_rng(2)
    gin = (5, 9, 5)
    gout = gin if stride == 1 else U.coarse(gin)
    a = rng.normal(size=gin + (3,))
    W = rng.normal(size=(27, 3, 4))
    g = rng.normal(size=gout + (4,))
    lhs = np.sum(U.conv(a, W, stride, gout) * g)
    rhs = np.sum(a * U.conv_adjoint(g, W, stride, gin))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # weight gradient: <conv(a, W), g> is linear in W with gradient wgrad(a, g)
    assert abs(lhs - np.sum(W * U.conv_wgrad(a, g, stride))) <= 1e-10 * abs(lhs)


def test_upsample_adjoint_dot_product():
    rng = np.random.default_rng(3)
    gc, gf = (3, 5, 2), (5, 9, 3)
    c = rng.normal(size=gc + (2,))
    f = rng.normal(size=gf + (2,))
    assert abs(np.sum(U.up(c, gf) * f) - np.sum(c * U.up_adjoint(f, gc))) < 1e-12 * np.sum(np.abs(f))


def test_pack_unpack_layout():
    offs, total = U.layer_offsets()
    assert total == sum(cout * U.kpad(cin) for cin, cout, _ in U.CHANNELS)
    w = U.init_weights(0)
    assert np.array_equal(U.pack(U.unpack(w)), w)
    # k = tap * cin + ci inside row co of layer 1 (cin 16, cout 32)
    mats = [np.zeros((27, cin, cout)) for cin, cout, _ in U.CHANNELS]
    mats[1][5, 3, 7] = 1.0
    w = U.pack(mats)
    assert np.flatnonzero(w).tolist() == [offs[1] + 7 * U.kpad(16) + 5 * 16 + 3]


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    a = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, size=5000),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0], dtype=np.float32)])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(U.bf16(a), ref)


def test_restriction_annihilates_checkerboard():
    """Full weighting (1, 1/2, 1/4, 1/8 for centre, face, edge, corner
    neighbours) maps the highest-frequency mode to zero at interior coarse
    nodes, so the learned correction leaves that mode to Jacobi."""
    shape = (9, 9, 9)
    k, j, i = np.meshgrid(*[np.arange(9)] * 3, indexing="ij")
    x = (1.0 - 2.0 * ((i + j + k) % 2)).reshape(shape + (1,))
    xc = U.up_adjoint(x, U.coarse(shape))
    assert np.abs(xc[1:-1, 1:-1, 1:-1]).max() < 1e-14
    assert xc[0, 0, 0, 0] == 1 - 3 * 0.5 + 3 * 0.25 - 0.125   # corner node: one octant of the stencil


def test_network_is_linear_and_scale_equivariant():
    rng = np.random.default_rng(5)
    shape = (9, 9, 17)
    w = U.init_weights(1)
    x1, x2 = rng.normal(size=1377), rng.normal(size=1377)
    y1, _ = U.forward(x1, w, shape)
    y2, _ = U.forward(x2, w, shape)
    y12, _ = U.forward(2.0 * x1 - 3.0 * x2, w, shape)
    np.testing.assert_allclose(y12, 2.0 * y1 - 3.0 * y2, rtol=1e-10, atol=1e-10 * np.abs(y12).max())


def test_network_is_symmetric_positive_semidefinite():
    """J^T J: <v, N u> = <u, N v> and <u, N u> = |J u|^2 >= 0 for any weights"""
    rng = np.random.default_rng(8)
    shape = (17, 9, 9)
    w = rng.normal(size=U.layer_offsets()[1])
    for _ in range(3):
        u, v = rng.normal(size=1377), rng.normal(size=1377)
        Nu, Nv = U.forward(u, w, shape, gain=0.7)[0], U.forward(v, w, shape, gain=0.7)[0]
        assert abs(v @ Nu - u @ Nv) <= 1e-11 * (abs(v @ Nu) + 1e-300)
        assert u @ Nu > 0


def _bpx(x, shape):
    """sum over the three coarse levels of 2^-(l-1) P_l P_l^T x with P_l^T the
    chain of (masked) full-weighting restrictions"""
    f = x.reshape(tuple(shape) + (1,))
    shapes, coarse_vals = [tuple(shape)], []
    for _ in range(3):
        sc = U.coarse(shapes[-1])
        f = U.mask(U.up_adjoint(f, sc))
        shapes.append(sc)
        coarse_vals.append(f)
    out = np.zeros_like(x)
    for lev in range(3, 0, -1):
        g = coarse_vals[lev - 1] * 0.5 ** (lev - 1)
        for k in range(lev, 0, -1):
            g = U.up(U.mask(g) if k < lev else g, shapes[k - 1])
        out += g.ravel()
    return out


def test_initial_weights_are_the_multilevel_preconditioner():
    rng = np.random.default_rng(9)
    shape = (17, 17, 9)
    x = rng.normal(size=17 * 17 * 9)
    y, _ = U.forward(x, U.init_weights(3, noise=0.0), shape)
    np.testing.assert_allclose(y, _bpx(x, shape), rtol=1e-12, atol=1e-12 * np.abs(y).max())


def test_initial_preconditioner_beats_jacobi_on_a_model_problem():
    """Dirichlet 7-point Laplacian on 17^3 (interior nodes, boundary rows
    pinned): PCG with dinv + gain J^T J at the initial weights needs fewer
    iterations than Jacobi-PCG"""
    import scipy.sparse as sp
    n1 = 17
    e = np.ones(n1)
    T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
    I1 = sp.identity(n1)
    A = (sp.kron(sp.kron(T, I1), I1) + sp.kron(sp.kron(I1, T), I1) + sp.kron(sp.kron(I1, I1), T)).tocsr()
    k, j, i = np.meshgrid(*[np.arange(n1)] * 3, indexing="ij")
    bnd = ((i == 0) | (j == 0) | (k == 0) | (i == n1 - 1) | (j == n1 - 1) | (k == n1 - 1)).ravel()
    A = A.tolil()
    for r in np.flatnonzero(bnd):
        A[r, :] = 0
        A[:, r] = 0
        A[r, r] = 6.0
    A = A.tocsr()
    dinv = 1.0 / A.diagonal()
    b = np.random.default_rng(10).normal(size=n1 ** 3)
    b[bnd] = 0
    w = U.init_weights(0, noise=0.0)
    gain = 0.25 * np.median(dinv)

    def pcg(M):
        x = np.zeros_like(b); r = b.copy(); z = M(r); p = z.copy(); rz = r @ z
        for it in range(1, 500):
            q = A @ p; a = rz / (p @ q); x += a * p; r -= a * q
            if np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b):
                return it
            z = M(r); rz2 = r @ z; p = z + rz2 / rz * p; rz = rz2
        return 500

    it_j = pcg(lambda r: dinv * r)
    it_l = pcg(lambda r: dinv * r + gain * U.forward(r, w, (n1,) * 3)[0])
    assert it_l < 0.75 * it_j, (it_l, it_j)


def _dirichlet_laplacian(n1):
    """7-point Laplacian on n1^3 with the boundary rows and columns replaced
    by 6 on the diagonal (an SPD model of the pressure operator)"""
    import scipy.sparse as sp
    e = np.ones(n1)
    T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
    I1 = sp.identity(n1)
    A = (sp.kron(sp.kron(T, I1), I1) + sp.kron(sp.kron(I1, T), I1) + sp.kron(sp.kron(I1, I1), T)).tocsr()
    k, j, i = np.meshgrid(*[np.arange(n1)] * 3, indexing="ij")
    bnd = ((i == 0) | (j == 0) | (k == 0) | (i == n1 - 1) | (j == n1 - 1) | (k == n1 - 1)).ravel()
    keep = sp.diags((~bnd).astype(float))
    A = (keep @ A @ keep + sp.diags(6.0 * bnd)).tocsr()
    return A, bnd


def test_two_level_error_propagator_is_the_product_of_three():
    """I - M Z = (I - S Z)(I - C Z)(I - S Z) for the two-level form, evaluated
    as the product of the three propagators on random vectors (a dropped
    sweep, a swapped order or a wrong sign breaks it)"""
    n1 = 9
    A, _ = _dirichlet_laplacian(n1)
    dinv = 1.0 / A.diagonal()
    shape = (n1,) * 3
    rng = np.random.default_rng(21)
    w = U.init_weights(4, noise=0.3)
    gain, om = 0.07, 0.8
    zop = lambda v: A @ v  # noqa: E731
    M = lambda v: U.precondition_two_level(v, w, shape, gain, om, dinv, zop)  # noqa: E731
    E = lambda v: v - om * dinv * (A @ v)  # noqa: E731
    F = lambda v: v - U.forward(A @ v, w, shape, gain)[0]  # noqa: E731
    for _ in range(3):
        v = rng.normal(size=n1 ** 3)
        lhs = v - M(A @ v)
        rhs = E(F(E(v)))
        np.testing.assert_allclose(lhs, rhs, rtol=0, atol=1e-12 * np.abs(rhs).max())
    # no coarse correction: two Jacobi sweeps, M = 2S - S Z S
    r = rng.normal(size=n1 ** 3)
    z0 = U.precondition_two_level(r, w, shape, 0.0, om, dinv, zop)
    S = om * dinv
    np.testing.assert_allclose(z0, 2 * S * r - S * (A @ (S * r)), rtol=1e-13, atol=1e-13)


def test_two_level_is_symmetric_positive_definite():
    """M is symmetric and positive definite for any network weights while the
    smoothing converges (omega lambda_max(D^-1 Z) < 2; about 1.6 here)"""
    n1 = 9
    A, _ = _dirichlet_laplacian(n1)
    dinv = 1.0 / A.diagonal()
    shape = (n1,) * 3
    rng = np.random.default_rng(22)
    zop = lambda v: A @ v  # noqa: E731
    for seed, noise, gain in ((1, 0.3, 0.05), (2, 1.0, 0.5)):
        w = U.init_weights(seed, noise=noise)
        M = lambda v: U.precondition_two_level(v, w, shape, gain, 0.8, dinv, zop)  # noqa: E731
        u, v = rng.normal(size=n1 ** 3), rng.normal(size=n1 ** 3)
        a, b = u @ M(v), M(u) @ v
        assert abs(a - b) <= 1e-12 * (abs(a) + abs(b)), (a, b)
        for _ in range(4):
            x = rng.normal(size=n1 ** 3)
            assert x @ M(x) > 0


def test_two_level_beats_additive_on_a_model_problem():
    """Dirichlet 7-point Laplacian on 17^3: PCG with the two-level form at the
    initial weights needs fewer iterations than the additive form, which needs
    fewer than Jacobi"""
    n1 = 17
    A, bnd = _dirichlet_laplacian(n1)
    dinv = 1.0 / A.diagonal()
    b = np.random.default_rng(23).normal(size=n1 ** 3)
    b[bnd] = 0
    w = U.init_weights(0, noise=0.0)
    gain = 0.25 * np.median(dinv)
    shape = (n1,) * 3

    def pcg(M):
        x = np.zeros_like(b); r = b.copy(); z = M(r); p = z.copy(); rz = r @ z
        for it in range(1, 500):
            q = A @ p; a = rz / (p @ q); x += a * p; r -= a * q
            if np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b):
                return it
            z = M(r); rz2 = r @ z; p = z + rz2 / rz * p; rz = rz2
        return 500

    it_j = pcg(lambda r: dinv * r)
    it_a = pcg(lambda r: U.precondition_additive(r, w, shape, gain, dinv))
    it_t = pcg(lambda r: U.precondition_two_level(r, w, shape, gain, 0.8, dinv, lambda v: A @ v))
    assert it_t < it_a < it_j, (it_t, it_a, it_j)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    shape = (9, 17, 9)
    n = 1377
    w = U.init_weights(2)
    x = rng.normal(size=n)
    t = rng.normal(size=n)
    gain = 0.7
    L, g, _ = U.loss_grad(x, t, w, shape, gain)
    offs, total = U.layer_offsets()
    h = 1e-5
    for li, ((cin, cout, _), off) in enumerate(zip(U.CHANNELS, offs)):
        kp = U.kpad(cin)
        for _ in range(4):
            co, k = rng.integers(cout), rng.integers(27 * cin)
            idx = off + co * kp + k
            wp, wm = w.copy(), w.copy()
            wp[idx] += h
            wm[idx] -= h
            fd = (U.loss_grad(x, t, wp, shape, gain)[0] - U.loss_grad(x, t, wm, shape, gain)[0]) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1e-3, abs(g).max()), (U.LAYER_NAMES[li], fd, g[idx])
    # padding entries carry no gradient
    for (cin, cout, _), off in zip(U.CHANNELS, offs):
        kp = U.kpad(cin)
        assert not np.any(g[off:off + cout * kp].reshape(cout, kp)[:, 27 * cin:])


def test_device_model_close_to_exact():
    """the BF16 model differs from float64 by BF16-sized relative errors only"""
    rng = np.random.default_rng(7)
    shape = (17, 17, 17)
    w = U.init_weights(3)
    x = rng.normal(size=17 ** 3)
    ye, _ = U.forward(x, w, shape, exact=True)
    yd, _ = U.forward(x, w, shape, exact=False)
    rel = np.linalg.norm(yd - ye) / np.linalg.norm(ye)
    assert 1e-5 < rel < 3e-2


def test_grid_check():
    with pytest.raises(ValueError):
        U.forward(np.zeros(13 * 13 * 13), U.init_weights(0), (13, 13, 13))
