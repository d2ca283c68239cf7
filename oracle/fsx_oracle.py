"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module. It loads ``oracle/_build/libfsx_oracle.so``, built by
``make -C oracle`` from ``fsx_oracle.cpp``, a C++ restatement of the
reference CPU path (fractionalStep/Cpp/blascoCodinaHuerta.cpp).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FSX_ORACLE_LIB names another build (the sanitizer build of `make -C oracle sanitize`)
_LIB_PATH = os.environ.get("FSX_ORACLE_LIB") or os.path.join(_HERE, "_build", "libfsx_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_new.restype = ctypes.c_void_p
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_error.restype = ctypes.c_char_p
        L.orc_error.argtypes = [ctypes.c_void_p]
        L.orc_read_inp.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_set_mesh.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, _I, ctypes.c_int, _D,
                                   ctypes.c_int, _I, ctypes.c_int, _D]
        L.orc_set_params.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int]
        L.orc_get_header.argtypes = [ctypes.c_void_p, _D]
        L.orc_setup.argtypes = [ctypes.c_void_p]
        L.orc_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.orc_len.restype = ctypes.c_longlong
        L.orc_len.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_get_d.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _D]
        L.orc_set_d.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _D]
        L.orc_get_i.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _I]
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        for f in ("orc_step1", "orc_step2", "orc_step3"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_kacc.argtypes = [ctypes.c_void_p]
        L.orc_convection.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.orc_pcg_z.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_double, ctypes.c_int, _D]
        L.orc_pcg_csr.argtypes = [ctypes.c_int, _I, _I, _D, _D, _D, ctypes.c_double, ctypes.c_int, _D]
        L.orc_spmv.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double]
        L.orc_timings.argtypes = [ctypes.c_void_p, _D]
        L.orc_set_mass.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_set_pcg_x0.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_f32_operands.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_direct.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_direct_solve.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.orc_chol_bandwidth.argtypes = [ctypes.c_void_p]
        L.orc_symbolic_cholesky.argtypes = [ctypes.c_int, _I, _I, _I, ctypes.POINTER(ctypes.c_longlong), _D, _I]
        L.orc_mass_stats.argtypes = [ctypes.c_void_p, _D]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


SIZE_KEYS = ("NE", "NCN", "NN", "NNp", "nnzM", "nnzG", "nnzZ", "nBCnodes", "nColors", "monPoint", "zeroP",
             "nVelFaces", "nBC")


class Oracle:
    """One oracle case: mesh + parameters -> setup -> steps."""

    def __init__(self):
        self.L = lib()
        self.h = ctypes.c_void_p(self.L.orc_new())

    def close(self):
        if self.h:
            self.L.orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what} failed: {self.L.orc_error(self.h).decode()}")

    def read_inp(self, path):
        self._check(self.L.orc_read_inp(self.h, path.encode()), "read_inp")

    def header(self):
        out = np.zeros(9)
        self.L.orc_get_header(self.h, _dp(out))
        keys = ("NGP", "density", "viscosity", "dt", "t_ini", "t_final", "maxIter", "tolerance", "converge")
        return dict(zip(keys, out.tolist()))

    def set_mesh(self, mesh):
        coord = np.ascontiguousarray(mesh["coord"], dtype=np.float64)
        conn = np.ascontiguousarray(mesh["conn"], dtype=np.int32)
        bcstr = np.ascontiguousarray(mesh["BCstr"], dtype=np.float64)
        faces = np.ascontiguousarray(mesh["faces"], dtype=np.int32)
        mon = np.ascontiguousarray(mesh.get("mon", [0.5, 0.5, 0.5]), dtype=np.float64)
        self._keep = (coord, conn, bcstr, faces, mon)
        self._check(self.L.orc_set_mesh(self.h, conn.shape[0], coord.shape[0], _dp(coord), _ip(conn),
                                        bcstr.shape[0], _dp(bcstr), faces.shape[0], _ip(faces),
                                        int(mesh["zeroP"]), _dp(mon)), "set_mesh")

    def set_params(self, NGP=27, density=1.0, viscosity=0.01, dt=0.01, maxIter=3, tolerance=1e-3,
                   converge=1e-6, pcg_rtol=1e-8, pcg_maxit=5000):
        self._check(self.L.orc_set_params(self.h, NGP, density, viscosity, dt, maxIter, tolerance, converge,
                                          pcg_rtol, pcg_maxit), "set_params")

    def setup(self):
        self._check(self.L.orc_setup(self.h), "setup")

    def sizes(self):
        out = (ctypes.c_longlong * 13)()
        self.L.orc_sizes(self.h, out)
        return dict(zip(SIZE_KEYS, list(out)))

    def d(self, name):
        n = self.L.orc_len(self.h, name.encode(), 0)
        if n < 0:
            raise KeyError(name)
        a = np.empty(n, dtype=np.float64)
        self.L.orc_get_d(self.h, name.encode(), _dp(a))
        return a

    def set_d(self, name, arr):
        n = self.L.orc_len(self.h, name.encode(), 0)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if n < 0 or a.size != n:
            raise ValueError(f"{name}: size {a.size} != {n}")
        self.L.orc_set_d(self.h, name.encode(), _dp(a))

    def i(self, name):
        n = self.L.orc_len(self.h, name.encode(), 1)
        if n < 0:
            raise KeyError(name)
        a = np.empty(n, dtype=np.int32)
        self.L.orc_get_i(self.h, name.encode(), _ip(a))
        return a

    def run(self, nsteps):
        stats = np.zeros((nsteps, 8))
        done = self.L.orc_run(self.h, nsteps, _dp(stats))
        return stats[:done]

    def set_pcg_x0(self, policy):
        """Pressure PCG initial guess: 0 zero (the CPU path), 1 the previous
        Pdot (the GPU path's warm start)."""
        self._check(self.L.orc_set_pcg_x0(self.h, policy), "set_pcg_x0")

    def set_direct(self, on):
        """Pressure solve by Cholesky of Z (1, the MATLAB prototype's step)
        or Jacobi-PCG (0, the reference)."""
        self._check(self.L.orc_set_direct(self.h, int(on)), "set_direct")

    def direct_solve(self, b):
        """x = Z^-1 b by the Cholesky factor of Z; returns (x, |b - Zx|/|b|)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        rel = np.zeros(1)
        self._check(self.L.orc_direct_solve(self.h, _dp(b), _dp(x), _dp(rel)), "direct_solve")
        return x, float(rel[0])

    def chol_bandwidth(self):
        return self.L.orc_chol_bandwidth(self.h)

    def set_f32_operands(self, f32):
        """Round step1's K operands UnpHalf_prev and Acc_prev and step2's G^T
        operand UnpHalf/dt^2 - MdOrigInv KAcc to FP32 (1) or not (0): a model
        of the device's FP32 operand shadows, lumped mass only."""
        self._check(self.L.orc_set_f32_operands(self.h, int(f32)), "set_f32_operands")

    def set_mass(self, mode, rtol=1e-10, maxit=1000):
        """Velocity mass: 0 lumped (the reference), 1 consistent M solved by
        Jacobi-PCG per component."""
        self._check(self.L.orc_set_mass(self.h, mode, rtol, maxit), "set_mass")

    def mass_stats(self):
        """(iterations, relative residual) of the last mass solve, largest over
        the components, and the total iterations so far."""
        out = np.zeros(3)
        self.L.orc_mass_stats(self.h, _dp(out))
        return int(out[0]), float(out[1]), int(out[2])

    def step1(self, it):
        self.L.orc_step1(self.h, it)

    def step2(self, it):
        self.L.orc_step2(self.h, it)

    def step3(self, it):
        self.L.orc_step3(self.h, it)

    def kacc(self):
        self.L.orc_kacc(self.h)

    def convection(self, Un, Uprev):
        Un = np.ascontiguousarray(Un, dtype=np.float64)
        Uprev = np.ascontiguousarray(Uprev, dtype=np.float64)
        R1 = np.empty_like(Un)
        self.L.orc_convection(self.h, _dp(Un), _dp(Uprev), _dp(R1))
        return R1

    def pcg_z(self, b, rtol=1e-8, maxit=5000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        rel = ctypes.c_double()
        it = self.L.orc_pcg_z(self.h, _dp(b), _dp(x), rtol, maxit, ctypes.byref(rel))
        return x, it, rel.value

    def spmv(self, which, x, y=None, alpha=1.0, beta=0.0):
        """which: 'K', 'G1'..'G3', 'Z', 'G1T'..'G3T'."""
        code = {"K": 0, "G1": 1, "G2": 2, "G3": 3, "Z": 4, "G1T": 5, "G2T": 6, "G3T": 7}[which]
        s = self.sizes()
        nout = s["NNp"] if which in ("Z", "G1T", "G2T", "G3T") else s["NN"]
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(nout) if y is None else np.array(y, dtype=np.float64)
        self.L.orc_spmv(self.h, code, _dp(x), _dp(y), alpha, beta)
        return y

    def timings(self):
        out = np.zeros(4)
        self.L.orc_timings(self.h, _dp(out))
        return dict(step_s=out[0], pcg_s=out[1], pcg_iters=int(out[2]), solves=int(out[3]))


def pcg_csr(rowptr, col, val, b, rtol=1e-8, maxit=5000):
    L = lib()
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    rel = ctypes.c_double()
    it = L.orc_pcg_csr(b.size, _ip(rowptr), _ip(col), _dp(val), _dp(b), _dp(x), rtol, maxit, ctypes.byref(rel))
    return x, it, rel.value


def symbolic_cholesky(rowptr, col, perm):
    """(nnz(L) with the diagonal, sum of squared column counts, elimination
    tree height) of the Cholesky factor of the symmetric pattern under perm
    (perm[new] = old)"""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    nnz = ctypes.c_longlong()
    fl = np.zeros(1)
    h = np.zeros(1, dtype=np.int32)
    rc = lib().orc_symbolic_cholesky(perm.size, _ip(rowptr), _ip(col), _ip(perm), ctypes.byref(nnz), _dp(fl), _ip(h))
    if rc != 0:
        raise ValueError(f"symbolic_cholesky: bad input ({rc})")
    return nnz.value, float(fl[0]), int(h[0])


def nested_dissection_grid(nx, ny, nz, radius=2, leaf=64):
    """Nested-dissection order of an nx*ny*nz grid (x fastest) whose nodes
    couple within `radius` per axis: each box is split across its longest
    axis by a slab of `radius` planes, the two halves ordered first
    (recursively), the slab last; boxes of at most `leaf` nodes keep the
    grid order. Returns perm (perm[new] = old)."""
    out = []

    def rec(x0, x1, y0, y1, z0, z1):
        dims = (x1 - x0, y1 - y0, z1 - z0)
        if dims[0] * dims[1] * dims[2] <= leaf or max(dims) <= 2 * radius + 1:
            for z in range(z0, z1):
                for y in range(y0, y1):
                    out.extend(range(z * nx * ny + y * nx + x0, z * nx * ny + y * nx + x1))
            return
        a = int(np.argmax(dims))
        lo, hi = (x0, x1) if a == 0 else ((y0, y1) if a == 1 else (z0, z1))
        m = (lo + hi - radius) // 2
        box = [x0, x1, y0, y1, z0, z1]
        left, right, sep = list(box), list(box), list(box)
        left[2 * a + 1] = m
        sep[2 * a], sep[2 * a + 1] = m, m + radius
        right[2 * a] = m + radius
        rec(*left)
        rec(*right)
        rec_sep = sep
        for z in range(rec_sep[4], rec_sep[5]):
            for y in range(rec_sep[2], rec_sep[3]):
                out.extend(range(z * nx * ny + y * nx + rec_sep[0], z * nx * ny + y * nx + rec_sep[1]))

    rec(0, nx, 0, ny, 0, nz)
    return np.asarray(out, dtype=np.int32)


def set_threads(n):
    lib().orc_set_threads(int(n))


def case_from_mesh(mesh, **params):
    o = Oracle()
    o.set_mesh(mesh)
    o.set_params(**params)
    o.setup()
    return o
