"""The allocation-free pressure solve (fsx_solve_pressure_noalloc) against
the allocating one (fsx_solve_pressure with the FP64 Jacobi-PCG from x0 = 0):
both run the same kernels in the same order on the same Z, so x, the
iteration count and the reported residual must be identical, bit for bit.
Two calls in a row on one GPU, the second on a smaller system (another
context), share the library's scratch array; the second must not see
anything of the first. b and x are device arrays that the test allocates
and copies with the HIP runtime (libamdhip64) itself."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make(fsx, N):
    mesh = fsx.Mesh.cavity(N).set_params(NGP=27, dt=0.01, maxIter=3, tolerance=1e-3, converge=1e-6, viscosity=0.01)
    s = fsx.Solver(mesh, device=0, NGP=27, maxIter=3, tolerance=1e-3, converge=1e-6, pcg_rtol=1e-10,
                   pcg_maxit=5000, solver_mode=0)
    return mesh, s


class DeviceArray:
    """n doubles of device memory from hipMalloc, with host copies in and out."""

    def __init__(self, hip, n):
        self.hip, self.n, self.ptr = hip, n, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n), 1) == 0

    def get(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), 2) == 0
        return a

    def free(self):
        assert self.hip.hipFree(self.ptr) == 0


def test_noalloc_pressure_solve_matches_allocating_solve(fsx_lib):
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    cases = []
    for N, seed in ((12, 21), (6, 22)):   # the second call is the smaller system
        say(f"no-alloc solve: cavity {N}^3")
        mesh, s = make(fsx_lib, N)
        n = s.sizes["NNp"]
        b = np.random.default_rng(seed).standard_normal(n)
        # the allocating entry point in the same configuration
        s.set_solver(fsx_lib.SOLVER_FP64, 1e-10, 5000)
        s.set_solver_x0(fsx_lib.X0_ZERO)
        x_ref, it_ref, rel_ref, rc_ref = s.solve_pressure(b)
        assert rc_ref == fsx_lib.FSX_OK
        # the allocation-free one on device arrays; x starts as garbage
        bd, xd = DeviceArray(hip, n), DeviceArray(hip, n)
        bd.put(b)
        xd.put(np.full(n, np.nan))
        it, rel, rc = s.solve_pressure_noalloc(bd.ptr.value, xd.ptr.value)
        x = xd.get()
        bd.free()
        xd.free()
        say(f"  {n} rows: allocating {it_ref} it {rel_ref:.3e}, no-alloc {it} it {rel:.3e}")
        assert rc == fsx_lib.FSX_OK
        assert it == it_ref and rel == rel_ref
        assert np.array_equal(x, x_ref)
        cases.append((mesh, s, n))
    assert cases[1][2] < cases[0][2]
    for mesh, s, _ in cases:
        s.close()
        mesh.close()
