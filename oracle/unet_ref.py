"""CPU oracle of the learned pressure preconditioner (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU baseline may import this
module; the product path never does.

The reference has no learned preconditioner: its pressure solver is
Jacobi-PCG only (CU:418-451 CUSP cg with diagonal preconditioner,
BCH:3981-4055 MKL RCI CG with the Jacobi step). This module restates the
network the GPU library trains and applies, so that its kernels can be
checked. The preconditioner is z = dinv r + gain J^T J r with

    f0 = B0 C0 x           conv 1 -> 16, stride 2, pressure grid -> G0
    f1 = B1 C1 f0          conv 16 -> 32, stride 2, G0 -> G1
    f2 = B2 C2 f1          conv 32 -> 64, stride 2, G1 -> G2
    J^T J x = C0^T B0 (f0 + C1^T B1 (f1 + C2^T f2))

B_l zeroes the boundary nodes of grid l. All convolutions are 3x3x3 with
zero padding and no bias;
grids are (nz, ny, nx) with x fastest, each coarse grid keeps every other
node of the one above. Weights of a layer are stored [cout][kp] with
k = tap * cin + ci, tap = (dx+1) + 3 (dy+1) + 9 (dz+1), kp = 27 cin rounded
up to 32.

Two arithmetic models:
  * exact=True: float64 everywhere (used for finite-difference checks of the
    hand-written gradient);
  * exact=False: the device's rounding points -- BF16 weights, BF16 copies of
    the activations that feed matrix-core convolutions, FP32 elsewhere --
    with float64 sums (the device's FP32 summation order differs, so
    comparisons use a tolerance).

Parity pins (tests/test_unet_oracle.py): the convolution against a
brute-force loop, the prolongation against linear interpolation of an affine
field (exact), every transposed operator against its forward by the
dot-product test <A x, y> = <x, A^T y>, symmetry and positivity of J^T J,
the initial weights against the additive multilevel formula, the gradient
against central finite differences, and BF16 rounding against torch.
"""
from __future__ import annotations

import numpy as np

CHANNELS = [(1, 16, 2), (16, 32, 2), (32, 64, 2)]
LAYER_NAMES = ["C0", "C1", "C2"]


def kpad(cin):
    return (27 * cin + 31) // 32 * 32


def layer_offsets():
    offs, off = [], 0
    for cin, cout, _ in CHANNELS:
        offs.append(off)
        off += cout * kpad(cin)
    return offs, off


def coarse(shape):
    return tuple((n + 1) // 2 for n in shape)


def bf16(a):
    """float32 -> bfloat16 (round to nearest even) -> float32, finite inputs."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(w):
    """flat parameter vector -> list of (27, cin, cout) tap matrices"""
    offs, total = layer_offsets()
    assert w.size == total
    mats = []
    for (cin, cout, _), off in zip(CHANNELS, offs):
        kp = kpad(cin)
        W = w[off:off + cout * kp].reshape(cout, kp)[:, :27 * cin]
        mats.append(W.reshape(cout, 27, cin).transpose(1, 2, 0))
    return mats


def pack(mats):
    offs, total = layer_offsets()
    w = np.zeros(total, dtype=np.float64)
    for (cin, cout, _), off, M in zip(CHANNELS, offs, mats):
        kp = kpad(cin)
        blk = np.zeros((cout, kp))
        blk[:, :27 * cin] = M.transpose(2, 0, 1).reshape(cout, 27 * cin)
        w[off:off + cout * kp] = blk.ravel()
    return w


def init_weights(seed, noise=0.1):
    """Channel 0 carries the additive multilevel preconditioner (C0: full
    weighting; C1, C2: full weighting x 1/sqrt(2)), so that
    J^T J = sum_l 2^-(l-1) P_l P_l^T over the three coarse levels; the other
    entries are seeded noise of `noise` x a fan-in-scaled normal (the
    device's own generator is not restated)."""
    rng = np.random.default_rng(seed)
    mats = [rng.normal(0.0, noise * np.sqrt(1.0 / (27 * cin)), size=(27, cin, cout)) for cin, cout, _ in CHANNELS]
    for li, M in enumerate(mats):
        a = 1.0 if li == 0 else np.sqrt(0.5)
        for t, dx, dy, dz in _taps():
            M[t, 0, 0] = a * (0.5 if dx else 1.0) * (0.5 if dy else 1.0) * (0.5 if dz else 1.0)
    return pack(mats)


def _taps():
    for t in range(27):
        yield t, t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def conv(a, W, stride, out_shape):
    """out[o] = sum_t a[stride*o + d_t] @ W[t]; a: (nz,ny,nx,cin)"""
    nz, ny, nx = out_shape
    p = np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nz, ny, nx, W.shape[2]), dtype=np.float64)
    for t, dx, dy, dz in _taps():
        s = p[1 + dz:1 + dz + stride * (nz - 1) + 1:stride,
              1 + dy:1 + dy + stride * (ny - 1) + 1:stride,
              1 + dx:1 + dx + stride * (nx - 1) + 1:stride]
        out += s @ W[t]
    return out


def conv_adjoint(g, W, stride, in_shape):
    """transpose of conv: gin[stride*o + d_t] += g[o] @ W[t]^T"""
    nz, ny, nx = g.shape[:3]
    gz, gy, gx = in_shape
    p = np.zeros((gz + 2, gy + 2, gx + 2, W.shape[1]), dtype=np.float64)
    for t, dx, dy, dz in _taps():
        p[1 + dz:1 + dz + stride * (nz - 1) + 1:stride,
          1 + dy:1 + dy + stride * (ny - 1) + 1:stride,
          1 + dx:1 + dx + stride * (nx - 1) + 1:stride] += g @ W[t].T
    return p[1:-1, 1:-1, 1:-1]


def conv_wgrad(a, g, stride):
    """dW[t][ci][co] = sum_o a[stride*o + d_t][ci] g[o][co]"""
    nz, ny, nx = g.shape[:3]
    p = np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))
    cin, cout = a.shape[3], g.shape[3]
    dW = np.zeros((27, cin, cout))
    g2 = g.reshape(-1, cout)
    for t, dx, dy, dz in _taps():
        s = p[1 + dz:1 + dz + stride * (nz - 1) + 1:stride,
              1 + dy:1 + dy + stride * (ny - 1) + 1:stride,
              1 + dx:1 + dx + stride * (nx - 1) + 1:stride]
        dW[t] = s.reshape(-1, cin).T @ g2
    return dW


def _interp_axis(c, axis, nf):
    """node-aligned linear interpolation along one axis: fine 2i = coarse i,
    fine 2i+1 = (coarse i + coarse i+1) / 2"""
    c = np.moveaxis(c, axis, 0)
    f = np.zeros((nf,) + c.shape[1:], dtype=np.float64)
    f[0::2] = c
    f[1::2] = 0.5 * (c[:-1] + c[1:])
    return np.moveaxis(f, 0, axis)


def up(c, fine_shape):
    f = _interp_axis(c, 0, fine_shape[0])
    f = _interp_axis(f, 1, fine_shape[1])
    return _interp_axis(f, 2, fine_shape[2])


def _restrict_axis(f, axis, nc):
    f = np.moveaxis(f, axis, 0)
    c = f[0::2].astype(np.float64).copy()
    c[:-1] += 0.5 * f[1::2]
    c[1:] += 0.5 * f[1::2]
    assert c.shape[0] == nc
    return np.moveaxis(c, 0, axis)


def up_adjoint(f, coarse_shape):
    c = _restrict_axis(f, 2, coarse_shape[2])
    c = _restrict_axis(c, 1, coarse_shape[1])
    return _restrict_axis(c, 0, coarse_shape[0])


def check_grid(shape):
    for n in shape:
        if n < 9 or (n - 1) % 8:
            raise ValueError("pressure grid needs 8m+1 nodes per axis")


def mask(a):
    """zero the boundary nodes (B_l)"""
    a = a.copy()
    a[0] = a[-1] = 0.0
    a[:, 0] = a[:, -1] = 0.0
    a[:, :, 0] = a[:, :, -1] = 0.0
    return a


def _models(exact):
    r = (lambda a: a) if exact else (lambda a: bf16(a).astype(np.float64))
    f32 = (lambda a: a) if exact else (lambda a: a.astype(np.float32).astype(np.float64))
    return r, f32


def apply_jtj(x3, W, grids, exact):
    """encoder and transpose pass on x3 (pressure grid, 1 channel); returns
    the activations the device keeps (dict)"""
    gf, g0, g1, g2 = grids
    r, f32 = _models(exact)
    A = {"x": x3}
    A["f0"] = f32(mask(conv(x3, W[0], 2, g0)))
    A["f1"] = f32(mask(conv(r(A["f0"]), W[1], 2, g1)))
    A["f2"] = f32(mask(conv(r(A["f1"]), W[2], 2, g2)))
    A["g1"] = f32(mask(conv_adjoint(r(A["f2"]), W[2], 2, g1) + A["f1"]))
    A["g0"] = f32(mask(conv_adjoint(r(A["g1"]), W[1], 2, g0) + A["f0"]))
    return A


def forward(x, w, shape, gain=1.0, exact=True):
    """y = gain J^T J x for x flat on the pressure grid `shape` (x fastest).
    Returns (y flat, cache for backward)."""
    check_grid(shape)
    gf = tuple(shape)
    g0 = coarse(gf)
    g1 = coarse(g0)
    g2 = coarse(g1)
    r, f32 = _models(exact)
    W = [r(M) for M in unpack(np.asarray(w, dtype=np.float64))]
    x3 = f32(np.asarray(x, dtype=np.float64).reshape(gf + (1,)))
    A = apply_jtj(x3, W, (gf, g0, g1, g2), exact)
    y = f32(gain * conv_adjoint(A["g0"], W[0], 2, gf))
    return y.ravel(), {"A": A, "W": W, "grids": (gf, g0, g1, g2), "gain": gain, "exact": exact}


def backward(gy, c):
    """Parameter gradient (flat, layer layout) of <gy, y> for the forward
    cached in c: with y = gain J^T J x, the gradient of gain <J gy, J x> is
    the encoder backpropagation at x with output gradients J (gain gy) plus
    the one at gain gy with output gradients J x, and both sets of output-side
    gradients are values of the transpose pass."""
    r, f32 = _models(c["exact"])
    gf = c["grids"][0]
    A = c["A"]
    B = apply_jtj(f32(c["gain"] * np.asarray(gy, dtype=np.float64).reshape(gf + (1,))), c["W"], c["grids"], c["exact"])
    pairs = [("x", "g0"), ("f0", "g1"), ("f1", "f2")]
    dW = [conv_wgrad(A[i], B[o], 2) + conv_wgrad(B[i], A[o], 2) for i, o in pairs]
    return pack(dW)


def loss_grad(x, t, w, shape, gain=1.0, exact=True):
    """relative loss |y - t|^2 / |t|^2 and its parameter gradient"""
    y, c = forward(x, w, shape, gain, exact)
    t = np.asarray(t, dtype=np.float64)
    tt = float(t @ t)
    d = y - t
    return float(d @ d) / tt, backward(2.0 * d / tt, c), y


def precondition_additive(r, w, shape, gain, dinv, exact=True):
    """The additive form the flexible PCG can apply: z = dinv r + gain J^T J r."""
    return dinv * r + forward(r, w, shape, gain, exact)[0]


def precondition_two_level(r, w, shape, gain, omega, dinv, zop, exact=True):
    """The two-level form: S = omega diag(dinv) (one Jacobi sweep) before and
    after the coarse correction C = gain J^T J (the network),

        z1 = S r,  r1 = r - Z z1,  z2 = z1 + C r1,  z = z2 + S (r - Z z2),

    so that I - M Z = (I - S Z)(I - C Z)(I - S Z). zop(v) applies Z."""
    z1 = omega * dinv * r
    r1 = r - zop(z1)
    z2 = z1 + forward(r1, w, shape, gain, exact)[0]
    return z2 + omega * dinv * (r - zop(z2))
