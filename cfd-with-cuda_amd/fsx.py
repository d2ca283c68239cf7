"""Python bindings (ctypes) of libfsx, the MI355X fractional-step core.

The library is built in-tree by ``make -C cfd-with-cuda_amd`` into
``cfd-with-cuda_amd/_build/libfsx.so``. Loading fails loudly when it is
missing: there is no fallback path. All numerics run in the library (HIP
kernels for gfx950 behind the C-ABI of include/fsx.h, host topology and the
time-loop driver of include/fsx_host.h).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FSX_LIB names another build of the library (the sanitizer build of `make sanitize`)
LIB_PATH = os.environ.get("FSX_LIB") or os.path.join(HERE, "_build", "libfsx.so")

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

FSX_OK, FSX_EINVAL, FSX_EHIP, FSX_ESTATE, FSX_ENOCONV, FSX_ENOMEM, FSX_ENUMERIC = range(7)
CSR_M, CSR_G, CSR_Z = 0, 1, 2
SOLVER_FP64, SOLVER_MIXED, SOLVER_LEARNED, SOLVER_DIRECT = 0, 1, 2, 3
KERNEL_STEP1, KERNEL_KACC, KERNEL_STEP2_RHS, KERNEL_STEP3, KERNEL_STEP1_ELEMENTS, KERNEL_STEP1_NODES = 0, 1, 2, 3, 4, 5
MASS_LUMPED, MASS_CONSISTENT, MASS_CONSISTENT_CSR = 0, 1, 2
X0_ZERO, X0_PREVIOUS, X0_PROJECTION = 0, 1, 2
DIRECT_BAND, DIRECT_ND = 0, 1

VEC = dict(Un=0, Unp1=1, Unp1_prev=2, UnpHalf=3, UnpHalf_prev=4, Acc=5, KAcc=6, R1=7, Pn=8, Pnp1=9,
           Pnp1_prev=10, Pdot=11, R2=12, MdInv=13, MdOrigInv=14, K=15, M=16, G1=17, G2=18, G3=19, Z=20)

# Exported C symbols declared in include/fsx.h and include/fsx_host.h
DEVICE_SYMBOLS = (
    "fsx_create", "fsx_destroy", "fsx_last_error", "fsx_device_bytes", "fsx_upload_mesh", "fsx_upload_csr",
    "fsx_upload_bc", "fsx_set_params", "fsx_set_solver", "fsx_set_graphs", "fsx_assemble_static", "fsx_build_Z", "fsx_init_state",
    "fsx_begin_step", "fsx_convection", "fsx_step1", "fsx_step2", "fsx_step3", "fsx_check_convergence",
    "fsx_next_iteration", "fsx_kacc", "fsx_end_step", "fsx_monitor", "fsx_download", "fsx_upload_vec", "fsx_sync",
    "fsx_solve_pressure", "fsx_solve_pressure_noalloc", "fsx_spmv", "fsx_bench_spmv", "fsx_bench_kernel", "fsx_bench_copy", "fsx_set_mass", "fsx_mass_stats", "fsx_set_mass_precision", "fsx_set_matrix_precision", "fsx_set_matrix_classes", "fsx_g_class_info", "fsx_set_z_tensor", "fsx_z_tensor_info", "fsx_set_step1_window", "fsx_set_step1_elements", "fsx_set_convection_form",
    "fsx_set_convection_precision", "fsx_set_solver_x0", "fsx_direct_stats", "fsx_set_direct_order", "fsx_direct_info", "fsx_z_patterns", "fsx_z_forms", "fsx_timing_reset", "fsx_timing_get",
    "fsx_unet_setup", "fsx_unet_params", "fsx_unet_get_weights", "fsx_unet_set_weights", "fsx_unet_gain",
    "fsx_unet_apply", "fsx_unet_loss_grad", "fsx_unet_capture", "fsx_unet_captured", "fsx_unet_prepare", "fsx_unet_prepare_random",
    "fsx_unet_train", "fsx_unet_limit", "fsx_unet_stats", "fsx_unet_form", "fsx_unet_form_get", "fsx_unet_precondition",
)
HOST_SYMBOLS = (
    "fsxh_mesh_cavity", "fsxh_mesh_duct", "fsxh_mesh_read_inp", "fsxh_mesh_write_inp", "fsxh_mesh_free",
    "fsxh_mesh_info", "fsxh_mesh_params", "fsxh_mesh_set_params", "fsxh_mesh_get", "fsxh_mesh_profile",
    "fsxh_topology", "fsxh_topo_free", "fsxh_topo_info", "fsxh_topo_len", "fsxh_topo_get",
    "fsxd_default_options", "fsxd_create", "fsxd_free", "fsxd_run", "fsxd_reset", "fsxd_ctx", "fsxd_topo", "fsxd_error",
    "fsxd_setup_times", "fsxh_write_tecplot", "fsxh_read_restart", "fsxd_write_tecplot", "fsxd_read_restart",
    "fsxh_nd_order", "fsxh_g_classes", "fsxh_z_tensor",
)


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("NGP", ctypes.c_int), ("maxIter", ctypes.c_int),
                ("tolerance", ctypes.c_double), ("converge", ctypes.c_double), ("solver_mode", ctypes.c_int),
                ("pcg_rtol", ctypes.c_double), ("pcg_maxit", ctypes.c_int), ("timing", ctypes.c_int)]


class StepStats(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int), ("inner_iters", ctypes.c_int), ("pcg_iters", ctypes.c_int),
                ("steady", ctypes.c_int), ("time", ctypes.c_double), ("max_acc", ctypes.c_double),
                ("mon", ctypes.c_double * 4), ("wall_s", ctypes.c_double), ("pcg_relres", ctypes.c_double),
                ("pcg_noconv", ctypes.c_int), ("pad_", ctypes.c_int), ("pcg_relres_max", ctypes.c_double)]


class Timing(ctypes.Structure):
    _fields_ = [("pcg_ms", ctypes.c_double), ("pcg_spmv_ms", ctypes.c_double),
                ("pcg_spmv_launches", ctypes.c_longlong), ("other_ms", ctypes.c_double),
                ("pcg_iters", ctypes.c_longlong)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libfsx.so not built: {LIB_PATH} (run `make -C cfd-with-cuda_amd` or "
                           f"__graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "fsx_create": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
        "fsx_destroy": ([_vp], None),
        "fsx_last_error": ([_vp], ctypes.c_char_p),
        "fsx_device_bytes": ([_vp], ctypes.c_size_t),
        "fsx_upload_mesh": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, _D, ctypes.c_int, _I, _I], ctypes.c_int),
        "fsx_upload_csr": ([_vp, ctypes.c_int, ctypes.c_int, _I, _I], ctypes.c_int),
        "fsx_upload_bc": ([_vp, ctypes.c_int, _I, _D, ctypes.c_int], ctypes.c_int),
        "fsx_set_params": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "fsx_set_solver": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int], ctypes.c_int),
        "fsx_set_graphs": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_assemble_static": ([_vp], ctypes.c_int),
        "fsx_build_Z": ([_vp], ctypes.c_int),
        "fsx_init_state": ([_vp], ctypes.c_int),
        "fsx_begin_step": ([_vp], ctypes.c_int),
        "fsx_convection": ([_vp], ctypes.c_int),
        "fsx_step1": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_step2": ([_vp, ctypes.c_int, _I, _D], ctypes.c_int),
        "fsx_step3": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_check_convergence": ([_vp, ctypes.c_double, _I, _D, _D], ctypes.c_int),
        "fsx_next_iteration": ([_vp], ctypes.c_int),
        "fsx_kacc": ([_vp], ctypes.c_int),
        "fsx_end_step": ([_vp, ctypes.c_double, _I, _D], ctypes.c_int),
        "fsx_monitor": ([_vp, ctypes.c_int, _D], ctypes.c_int),
        "fsx_download": ([_vp, ctypes.c_int, _D, ctypes.c_size_t], ctypes.c_int),
        "fsx_upload_vec": ([_vp, ctypes.c_int, _D, ctypes.c_size_t], ctypes.c_int),
        "fsx_sync": ([_vp], ctypes.c_int),
        "fsx_solve_pressure": ([_vp, _D, _D, _I, _D], ctypes.c_int),
        "fsx_solve_pressure_noalloc": ([_vp, _vp, _vp, _I, _D], ctypes.c_int),
        "fsx_spmv": ([_vp, ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "fsx_bench_spmv": ([_vp, ctypes.c_int, ctypes.c_int, _D, _D], ctypes.c_int),
        "fsx_bench_kernel": ([_vp, ctypes.c_int, ctypes.c_int, _D, _D], ctypes.c_int),
        "fsx_set_mass": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int], ctypes.c_int),
        "fsx_set_step1_window": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_set_step1_elements": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_set_convection_form": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_set_matrix_precision": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_set_convection_precision": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_set_solver_x0": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_direct_stats": ([_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "fsx_mass_stats": ([_vp, _I, _D, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "fsx_bench_copy": ([_vp, ctypes.c_size_t, ctypes.c_int, _D], ctypes.c_int),
        "fsx_z_patterns": ([_vp, _I, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "fsx_z_forms": ([_vp, ctypes.POINTER(ctypes.c_longlong), _I, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "fsx_timing_reset": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_timing_get": ([_vp, ctypes.POINTER(Timing)], ctypes.c_int),
        "fsx_unet_setup": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64], ctypes.c_int),
        "fsx_unet_params": ([_vp, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "fsx_unet_get_weights": ([_vp, _F, ctypes.c_size_t], ctypes.c_int),
        "fsx_unet_set_weights": ([_vp, _F, ctypes.c_size_t], ctypes.c_int),
        "fsx_unet_gain": ([_vp, ctypes.c_float, _F], ctypes.c_int),
        "fsx_unet_apply": ([_vp, _F, _F], ctypes.c_int),
        "fsx_unet_loss_grad": ([_vp, ctypes.c_int, _F, _F, _D, _F], ctypes.c_int),
        "fsx_unet_capture": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_unet_captured": ([_vp, _I], ctypes.c_int),
        "fsx_unet_prepare": ([_vp, _I, ctypes.c_int, ctypes.c_int, _I], ctypes.c_int),
        "fsx_unet_prepare_random": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _I], ctypes.c_int),
        "fsx_unet_train": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, _D], ctypes.c_int),
        "fsx_unet_limit": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_unet_form": ([_vp, ctypes.c_int, ctypes.c_float], ctypes.c_int),
        "fsx_unet_form_get": ([_vp, _I, _F], ctypes.c_int),
        "fsx_unet_precondition": ([_vp, _F, _F], ctypes.c_int),
        "fsx_unet_stats": ([_vp, _I, _I], ctypes.c_int),
        "fsxh_mesh_cavity": ([ctypes.c_int, ctypes.c_double, ctypes.c_double], _vp),
        "fsxh_mesh_duct": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_int], _vp),
        "fsxh_mesh_read_inp": ([ctypes.c_char_p, ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        "fsxh_mesh_write_inp": ([_vp, ctypes.c_char_p], ctypes.c_int),
        "fsxh_mesh_free": ([_vp], None),
        "fsxh_mesh_info": ([_vp, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "fsxh_mesh_params": ([_vp, _D], ctypes.c_int),
        "fsxh_mesh_set_params": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "fsxh_mesh_get": ([_vp, ctypes.c_char_p, _vp], ctypes.c_int),
        "fsxh_mesh_profile": ([_vp, _D], ctypes.c_int),
        "fsxh_topology": ([_vp, ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        "fsxh_topo_free": ([_vp], None),
        "fsxh_topo_info": ([_vp, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "fsxh_topo_len": ([_vp, ctypes.c_char_p], ctypes.c_longlong),
        "fsxh_topo_get": ([_vp, ctypes.c_char_p, _vp], ctypes.c_int),
        "fsxd_default_options": ([ctypes.POINTER(Options)], None),
        "fsxd_create": ([_vp, ctypes.POINTER(Options), ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        "fsxd_free": ([_vp], None),
        "fsxd_run": ([_vp, ctypes.c_int, ctypes.POINTER(StepStats)], ctypes.c_int),
        "fsxd_reset": ([_vp], ctypes.c_int),
        "fsxd_ctx": ([_vp], _vp),
        "fsxd_topo": ([_vp], _vp),
        "fsxd_error": ([_vp], ctypes.c_char_p),
        "fsxd_setup_times": ([_vp, _D], ctypes.c_int),
        "fsxh_write_tecplot": ([_vp, ctypes.c_char_p, _D, _D, ctypes.c_char_p], ctypes.c_int),
        "fsxh_read_restart": ([_vp, ctypes.c_char_p, _D, _D], ctypes.c_int),
        "fsxd_write_tecplot": ([_vp, ctypes.c_char_p, ctypes.c_char_p], ctypes.c_int),
        "fsxd_read_restart": ([_vp, ctypes.c_char_p], ctypes.c_int),
        "fsxh_nd_order": ([ctypes.c_int, _I, _I, _D, ctypes.c_int, _I, _D], ctypes.c_int),
        "fsx_set_direct_order": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_set_mass_precision": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_direct_info": ([_vp, _D, _D, _I, _I], ctypes.c_int),
        "fsx_set_matrix_classes": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_set_z_tensor": ([_vp, ctypes.c_int], ctypes.c_int),
        "fsx_z_tensor_info": ([_vp, _I, _D], ctypes.c_int),
        "fsx_g_class_info": ([_vp, _I, ctypes.POINTER(ctypes.c_longlong), _I, ctypes.POINTER(ctypes.c_longlong)],
                             ctypes.c_int),
        "fsxh_g_classes": ([ctypes.c_int, ctypes.c_int, _I, _I, _D, ctypes.c_double, _I, _I, _I, _F], ctypes.c_int),
        "fsxh_z_tensor": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.c_double, ctypes.c_int, _I,
                           _I, _D, ctypes.c_int, _D, _D], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def g_classes(rp, ci, g3, ncols, prune_rel=2.0**-24):
    """Row classes of G (fsxh_g_classes): G's CSR pattern (rp, ci) with
    g3 = (G1, G2, G3) a row of an (nnz, 3) array over ncols columns. Returns
    None when the rows do not compress, else a dict: counts (G's classes,
    their entries, G^T's classes, theirs, 1 if G^T's classes carry column
    offsets), row_cls, trow_cls and g3_f32, the (nnz, 3) FP32 values the
    class form represents (NaN where G^T's classes disagree with G's)."""
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    g3 = np.ascontiguousarray(g3, dtype=np.float64).reshape(-1, 3)
    n = rp.size - 1
    cnt = np.zeros(5, dtype=np.int32)
    rc_ = np.zeros(n, dtype=np.int32)
    tc = np.zeros(int(ncols), dtype=np.int32)
    g32 = np.zeros(g3.shape, dtype=np.float32)
    rc = lib().fsxh_g_classes(n, int(ncols), _ip(rp), _ip(ci), _dp(g3), float(prune_rel), _ip(cnt), _ip(rc_),
                              _ip(tc), _fp(g32))
    if rc != FSX_OK:
        raise FsxError(rc, "fsxh_g_classes: inconsistent pattern")
    if cnt[0] < 0:
        return None
    return dict(counts=tuple(int(c) for c in cnt), row_cls=rc_, trow_cls=tc, g3_f32=g32)


def z_tensor(shape, xc, ngp, rho, rp, ci, v, zero_p):
    """Z of a tensor-product grid as Kronecker products of 1D stencils
    (fsxh_z_tensor): shape (nx, ny, nz), corner coordinates xc (n, 3), Z's
    rows (rp, ci, v). Returns dict(holds, err, pen, coef) with coef the
    ((nx + ny + nz), 10) tables (K's five coefficients, then M's), None when
    the coordinates do not separate."""
    nx, ny, nz = (int(a) for a in shape)
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    coef = np.zeros((nx + ny + nz, 10))
    info = np.zeros(3)
    rc = lib().fsxh_z_tensor(nx, ny, nz, _dp(xc), int(ngp), float(rho), rp.size - 1, _ip(rp), _ip(ci), _dp(v),
                             int(zero_p), _dp(coef), _dp(info))
    if rc != FSX_OK:
        raise FsxError(rc, "fsxh_z_tensor: inconsistent sizes")
    return dict(holds=bool(info[2]), err=float(info[0]), pen=float(info[1]), coef=coef if info[0] >= 0 else None)


def nd_order(rp, ci, xyz, leaf=64):
    """The direct solve's nested-dissection order of a symmetric pattern
    (fsxh_nd_order): (perm new -> old, dict of nnz_l, flops, fronts, depth)."""
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = rp.size - 1
    perm = np.zeros(n, np.int32)
    st = np.zeros(4)
    rc = lib().fsxh_nd_order(n, _ip(rp), _ip(ci), _dp(xyz), int(leaf), _ip(perm), _dp(st))
    if rc != FSX_OK:
        raise FsxError(rc, "fsxh_nd_order: inconsistent pattern or coordinates")
    return perm, dict(nnz_l=st[0], flops=st[1], fronts=int(st[2]), depth=int(st[3]))


class FsxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fsx error {code}: {msg}")
        self.code = code


# ------------------------------------------------------------------ mesh
class Mesh:
    """Mesh input (the content of a reference .inp file)."""

    def __init__(self, handle):
        if not handle:
            raise FsxError(FSX_EINVAL, "mesh construction failed")
        self.L = lib()
        self.h = _vp(handle)

    @classmethod
    def cavity(cls, N, cluster=0.0, lid=1.0):
        return cls(lib().fsxh_mesh_cavity(N, cluster, lid))

    @classmethod
    def duct(cls, n=16, n_in=40, n_arc=40, n_out=80, L_in=5.0, R=2.0, L_out=10.0, peak=1.0, seed=12345,
             permute=True):
        return cls(lib().fsxh_mesh_duct(n, n_in, n_arc, n_out, L_in, R, L_out, peak, seed, 1 if permute else 0))

    @classmethod
    def read_inp(cls, path):
        out = _vp()
        err = ctypes.create_string_buffer(512)
        rc = lib().fsxh_mesh_read_inp(path.encode(), ctypes.byref(out), err, 512)
        if rc != FSX_OK:
            raise FsxError(rc, err.value.decode())
        return cls(out.value)

    def write_inp(self, path):
        rc = self.L.fsxh_mesh_write_inp(self.h, path.encode())
        if rc != FSX_OK:
            raise FsxError(rc, "write_inp")

    def close(self):
        if self.h:
            self.L.fsxh_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = (ctypes.c_longlong * 8)()
        self.L.fsxh_mesh_info(self.h, out)
        return dict(zip(("NE", "NCN", "nVelFaces", "nBC", "zeroP", "NGP", "maxIter", "has_profile"), list(out)))

    def params(self):
        out = np.zeros(10)
        self.L.fsxh_mesh_params(self.h, _dp(out))
        keys = ("dt", "t_ini", "t_final", "tolerance", "converge", "density", "viscosity", "mon_x", "mon_y", "mon_z")
        return dict(zip(keys, out.tolist()))

    def set_params(self, NGP=27, dt=0.01, maxIter=3, tolerance=1e-3, converge=1e-6, density=1.0, viscosity=0.01):
        rc = self.L.fsxh_mesh_set_params(self.h, NGP, dt, maxIter, tolerance, converge, density, viscosity)
        if rc != FSX_OK:
            raise FsxError(rc, "set_params")
        return self

    def profile(self):
        out = np.zeros(4)
        self.L.fsxh_mesh_profile(self.h, _dp(out))
        return dict(bc=int(out[0]), peak=out[1], y0=out[2], z0=out[3])

    def arrays(self):
        i = self.info()
        coord = np.empty((i["NCN"], 3))
        conn = np.empty((i["NE"], 8), dtype=np.int32)
        faces = np.empty((i["nVelFaces"], 3), dtype=np.int32)
        bcstr = np.empty((i["nBC"], 3))
        for name, a in (("coord", coord), ("conn", conn), ("faces", faces), ("BCstr", bcstr)):
            self.L.fsxh_mesh_get(self.h, name.encode(), a.ctypes.data_as(_vp))
        p = self.params()
        return dict(coord=coord, conn=conn, faces=faces, BCstr=bcstr, zeroP=i["zeroP"],
                    mon=np.array([p["mon_x"], p["mon_y"], p["mon_z"]]))


class Topology:
    INT_ARRAYS = ("LtoGnode", "Mrow", "Mcol", "Grow", "Gcol", "Zrow", "Zcol", "BCnodes", "BCwhich", "color_ptr",
                  "color_order")

    def __init__(self, mesh=None, handle=None, owner=True):
        self.L = lib()
        self.owner = owner
        if handle is not None:
            self.h = _vp(handle)
        else:
            out = _vp()
            err = ctypes.create_string_buffer(512)
            rc = self.L.fsxh_topology(mesh.h, ctypes.byref(out), err, 512)
            if rc != FSX_OK:
                raise FsxError(rc, err.value.decode())
            self.h = out

    def close(self):
        if self.h and self.owner:
            self.L.fsxh_topo_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = (ctypes.c_longlong * 11)()
        self.L.fsxh_topo_info(self.h, out)
        keys = ("NE", "NCN", "NN", "NNp", "nnzM", "nnzG", "nnzZ", "nBCnodes", "ncolors", "monPoint", "zeroP")
        return dict(zip(keys, list(out)))

    def write_tecplot(self, path, Un, Pn, title=""):
        """Fields ([u|v|w] and corner pressures) to a Tecplot file in the
        reference's layout (createTecplot, BCH:4249-4482)."""
        un = np.ascontiguousarray(Un, dtype=np.float64)
        pn = np.ascontiguousarray(Pn, dtype=np.float64)
        rc = self.L.fsxh_write_tecplot(self.h, title.encode(), _dp(un), _dp(pn), path.encode())
        if rc != FSX_OK:
            raise FsxError(rc, f"cannot write {path}")

    def read_restart(self, path):
        """(Un, Pn) from a file in that layout (readRestartFile, BCH:4214-4242)."""
        i = self.info()
        un = np.zeros(3 * i["NN"])
        pn = np.zeros(i["NNp"])
        rc = self.L.fsxh_read_restart(self.h, path.encode(), _dp(un), _dp(pn))
        if rc != FSX_OK:
            raise FsxError(rc, f"cannot read restart file {path}")
        return un, pn

    def get(self, name):
        n = self.L.fsxh_topo_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        a = np.empty(n, dtype=np.float64 if name in ("coord", "BCvalues") else np.int32)
        self.L.fsxh_topo_get(self.h, name.encode(), a.ctypes.data_as(_vp))
        return a


# ------------------------------------------------------------------ solver
def default_options(**kw):
    o = Options()
    lib().fsxd_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Solver:
    """Driver: topology + device context + the reference's time loop."""

    def __init__(self, mesh, **opts):
        self.L = lib()
        self.opt = default_options(**opts)
        out = _vp()
        err = ctypes.create_string_buffer(1024)
        rc = self.L.fsxd_create(mesh.h, ctypes.byref(self.opt), ctypes.byref(out), err, 1024)
        if rc != FSX_OK:
            raise FsxError(rc, err.value.decode())
        self.h = out
        self.ctx = _vp(self.L.fsxd_ctx(self.h))
        self.topo = Topology(handle=self.L.fsxd_topo(self.h), owner=False)
        self.sizes = self.topo.info()
        self.mesh_params = mesh.params()

    def close(self):
        if getattr(self, "h", None):
            self.L.fsxd_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def err(self):
        return self.L.fsx_last_error(self.ctx).decode()

    def check(self, rc, what=""):
        if rc != FSX_OK:
            raise FsxError(rc, f"{what}: {self.err()}")

    def run(self, nsteps):
        stats = (StepStats * max(1, nsteps))()
        n = self.L.fsxd_run(self.h, nsteps, stats)
        if n < 0:
            raise FsxError(-n, self.L.fsxd_error(self.h).decode())
        return [dict(step=s.step, inner=s.inner_iters, pcg=s.pcg_iters, steady=s.steady, time=s.time,
                     max_acc=s.max_acc, mon=list(s.mon), wall_s=s.wall_s, relres=s.pcg_relres,
                     noconv=s.pcg_noconv, relres_max=s.pcg_relres_max) for s in stats[:n]]

    def set_case(self, viscosity, lid=1.0):
        """Another viscosity and lid speed on the same mesh and context: the
        BC values of the mesh scaled by `lid` are uploaded, the static
        matrices are assembled again with the new viscosity (K scales with it;
        M, G and Z do not change) and the state goes back to the initial
        fields (fsx_init_state via fsxd_reset)."""
        p = self.mesh_params
        self.check(self.L.fsx_set_params(self.ctx, self.opt.NGP, p["density"], float(viscosity), p["dt"]), "set_params")
        nodes = np.ascontiguousarray(self.topo.get("BCnodes"), dtype=np.int32)
        vals = np.ascontiguousarray(self.topo.get("BCvalues"), dtype=np.float64) * float(lid)
        self.check(self.L.fsx_upload_bc(self.ctx, nodes.size, _ip(nodes), _dp(vals), int(self.sizes["zeroP"])),
                   "upload_bc")
        self.check(self.L.fsx_assemble_static(self.ctx), "assemble_static")
        self.reset()

    def reset(self, mode=None, rtol=None, maxit=None):
        """Back to the initial fields; optionally switch the pressure solver."""
        rc = self.L.fsxd_reset(self.h)
        if rc != FSX_OK:
            raise FsxError(rc, self.L.fsxd_error(self.h).decode())
        if mode is not None:
            self.set_solver(mode, self.opt.pcg_rtol if rtol is None else rtol,
                            self.opt.pcg_maxit if maxit is None else maxit)

    def bench_spmv(self, precision=0, reps=50):
        ms = ctypes.c_double()
        by = ctypes.c_double()
        self.check(self.L.fsx_bench_spmv(self.ctx, precision, reps, ctypes.byref(ms), ctypes.byref(by)), "bench_spmv")
        return ms.value, by.value

    def write_tecplot(self, path, title=""):
        """Current Un, Pn to a Tecplot file (createTecplot, BCH:4249-4482)."""
        rc = self.L.fsxd_write_tecplot(self.h, title.encode(), path.encode())
        if rc != FSX_OK:
            raise FsxError(rc, self.L.fsxd_error(self.h).decode())

    def read_restart(self, path):
        """Restart from a Tecplot file instead of the initial BCs (BCH:2795-2796)."""
        rc = self.L.fsxd_read_restart(self.h, path.encode())
        if rc != FSX_OK:
            raise FsxError(rc, self.L.fsxd_error(self.h).decode())

    def set_matrix_precision(self, bits):
        """64 (the reference) or 32: FP32 copies of K, G, G^T in the step
        kernels, the mixed-precision forms' FP32 assembly."""
        self.check(self.L.fsx_set_matrix_precision(self.ctx, bits), "set_matrix_precision")

    def set_matrix_classes(self, on):
        """At matrix precision 32: G and G^T as row classes where the rows
        compress (True, the default) or the FP32 CSR copies (False)."""
        self.check(self.L.fsx_set_matrix_classes(self.ctx, int(bool(on))), "set_matrix_classes")

    def set_z_tensor(self, on):
        """The learned solver's tiles apply Z as Kronecker products of 1D
        stencils where that form holds (True, the default) or the FP32 row
        classes (False)."""
        self.check(self.L.fsx_set_z_tensor(self.ctx, int(bool(on))), "set_z_tensor")

    def z_tensor_info(self):
        """(in use, largest relative mismatch against Z's rows or -1)."""
        a = ctypes.c_int()
        e = ctypes.c_double()
        self.check(self.L.fsx_z_tensor_info(self.ctx, ctypes.byref(a), ctypes.byref(e)), "z_tensor_info")
        return bool(a.value), e.value

    def g_class_info(self):
        """(G's classes, their entries, G^T's classes, theirs) of the class
        form in use; zeros when the step kernels read the CSR copies."""
        a, c = ctypes.c_int(), ctypes.c_int()
        b, d = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.L.fsx_g_class_info(self.ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                           ctypes.byref(d)), "g_class_info")
        return a.value, b.value, c.value, d.value

    def set_convection_form(self, elements):
        self.check(self.L.fsx_set_convection_form(self.ctx, int(bool(elements))), "set_convection_form")

    def set_step1_window(self, enable):
        self.check(self.L.fsx_set_step1_window(self.ctx, int(bool(enable))), "set_step1_window")

    def set_step1_elements(self, enable):
        """step1 of the mixed forms from the element forms of the convection
        and K on the matrix cores (the default) or from K's CSR values."""
        self.check(self.L.fsx_set_step1_elements(self.ctx, int(bool(enable))), "set_step1_elements")

    def set_solver_x0(self, policy):
        """X0_ZERO (the CPU path), X0_PREVIOUS (warm start from the previous
        Pdot, the reference's GPU path) or X0_PROJECTION (from the Z-norm
        projection onto the last 16 solutions)."""
        self.check(self.L.fsx_set_solver_x0(self.ctx, policy), "set_solver_x0")

    def direct_stats(self):
        """(half-bandwidth, factorization ms, band bytes) of the direct
        solver's Cholesky factor of Z; half-bandwidth -1 before the first
        direct solve."""
        w = ctypes.c_int()
        ms = ctypes.c_double()
        by = ctypes.c_double()
        self.check(self.L.fsx_direct_stats(self.ctx, ctypes.byref(w), ctypes.byref(ms), ctypes.byref(by)), "direct_stats")
        return w.value, ms.value, by.value

    def set_mass_precision(self, bits):
        """64 (default) or 32: the consistent-mass solves in FP32 segments
        with FP64 residuals and stop test (the element form only)."""
        self.check(self.L.fsx_set_mass_precision(self.ctx, bits), "set_mass_precision")

    def set_direct_order(self, order):
        """DIRECT_ND (the default: nested dissection, multifrontal factor) or
        DIRECT_BAND (band Cholesky) for the direct pressure solve."""
        self.check(self.L.fsx_set_direct_order(self.ctx, order), "set_direct_order")

    def direct_info(self):
        """(factor bytes, factor flops, fronts, depth) of the nested-dissection
        factor; depth -1 before one is formed."""
        by = ctypes.c_double()
        fl = ctypes.c_double()
        nf = ctypes.c_int()
        dp = ctypes.c_int()
        self.check(self.L.fsx_direct_info(self.ctx, ctypes.byref(by), ctypes.byref(fl), ctypes.byref(nf),
                                          ctypes.byref(dp)), "direct_info")
        return by.value, fl.value, nf.value, dp.value

    def set_convection_precision(self, bits):
        """64 (the reference) or 32: the convection residual's element
        arithmetic in FP32 (the mixed forms)."""
        self.check(self.L.fsx_set_convection_precision(self.ctx, bits), "set_convection_precision")

    def set_mass(self, mode, rtol=1e-10, maxit=1000):
        """Velocity mass of the step1/step3 updates: MASS_LUMPED (the
        reference) or MASS_CONSISTENT (3-RHS Jacobi-PCG on M with BC rows and
        columns replaced by the identity)."""
        self.check(self.L.fsx_set_mass(self.ctx, mode, rtol, maxit), "set_mass")

    def mass_stats(self):
        """(iterations, relative residual) of the last mass solve and the
        total mass iterations so far."""
        it = ctypes.c_int()
        rel = ctypes.c_double()
        tot = ctypes.c_longlong()
        self.check(self.L.fsx_mass_stats(self.ctx, ctypes.byref(it), ctypes.byref(rel), ctypes.byref(tot)), "mass_stats")
        return it.value, rel.value, tot.value

    def bench_kernel(self, kernel, reps=20):
        """(ms per launch, algorithmic bytes per launch) of one step row kernel
        (KERNEL_STEP1, KERNEL_KACC, KERNEL_STEP2_RHS, KERNEL_STEP3, and the two
        passes of step1's element form, KERNEL_STEP1_ELEMENTS and
        KERNEL_STEP1_NODES); clobbers
        the step vectors (reset() before stepping again)."""
        ms = ctypes.c_double()
        by = ctypes.c_double()
        self.check(self.L.fsx_bench_kernel(self.ctx, kernel, reps, ctypes.byref(ms), ctypes.byref(by)), "bench_kernel")
        return ms.value, by.value

    def z_patterns(self):
        n = ctypes.c_int()
        ln = ctypes.c_longlong()
        self.check(self.L.fsx_z_patterns(self.ctx, ctypes.byref(n), ctypes.byref(ln)), "z_patterns")
        return n.value, ln.value

    def z_forms(self):
        """(entries of Z kept by the solvers, FP32 row classes, class-table
        entries); 0 classes when the FP32 copy did not compress"""
        nz = ctypes.c_longlong()
        nc = ctypes.c_int()
        ln = ctypes.c_longlong()
        self.check(self.L.fsx_z_forms(self.ctx, ctypes.byref(nz), ctypes.byref(nc), ctypes.byref(ln)), "z_forms")
        return nz.value, nc.value, ln.value

    def bench_copy(self, nbytes, reps=20):
        ms = ctypes.c_double()
        self.check(self.L.fsx_bench_copy(self.ctx, nbytes, reps, ctypes.byref(ms)), "bench_copy")
        return ms.value

    def setup_times(self):
        out = np.zeros(3)
        self.L.fsxd_setup_times(self.h, _dp(out))
        return dict(topology=out[0], upload=out[1], assembly=out[2])

    def vec_len(self, name):
        s = self.sizes
        if name in ("Pn", "Pnp1", "Pnp1_prev", "Pdot", "R2"):
            return s["NNp"]
        if name in ("K", "M"):
            return s["nnzM"]
        if name in ("G1", "G2", "G3"):
            return s["nnzG"]
        if name == "Z":
            return s["nnzZ"]
        return 3 * s["NN"]

    def download(self, name):
        n = self.vec_len(name)
        a = np.empty(n)
        self.check(self.L.fsx_download(self.ctx, VEC[name], _dp(a), n), f"download {name}")
        return a

    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.check(self.L.fsx_upload_vec(self.ctx, VEC[name], _dp(a), a.size), f"upload {name}")

    def solve_pressure(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        it = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.L.fsx_solve_pressure(self.ctx, _dp(b), _dp(x), ctypes.byref(it), ctypes.byref(rel))
        if rc not in (FSX_OK, FSX_ENOCONV):
            self.check(rc, "solve_pressure")
        return x, it.value, rel.value, rc

    def solve_pressure_noalloc(self, b_dev, x_dev):
        """The allocation-free FP64 Jacobi-PCG on device arrays of NNp doubles
        (raw device addresses b_dev, x_dev, ready for the context's stream);
        returns (iterations, relative residual, status)."""
        it = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.L.fsx_solve_pressure_noalloc(self.ctx, _vp(b_dev), _vp(x_dev), ctypes.byref(it), ctypes.byref(rel))
        if rc not in (FSX_OK, FSX_ENOCONV):
            self.check(rc, "solve_pressure_noalloc")
        return it.value, rel.value, rc

    def spmv(self, which, x, y=None, alpha=1.0, beta=0.0):
        """which: K, G1..G3, Z, G1T..G3T; Zsolver64 / Zsolver32 apply Z as the
        FP64 / FP32 solvers read it (alpha = 1, beta = 0 only)"""
        code = {"K": 0, "G1": 1, "G2": 2, "G3": 3, "Z": 4, "G1T": 5, "G2T": 6, "G3T": 7,
                "Zsolver64": 8, "Zsolver32": 9}[which]
        s = self.sizes
        ny = s["NNp"] if which in ("Z", "G1T", "G2T", "G3T", "Zsolver64", "Zsolver32") else s["NN"]
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(ny) if y is None else np.array(y, dtype=np.float64)
        self.check(self.L.fsx_spmv(self.ctx, code, _dp(x), _dp(y), alpha, beta), "spmv")
        return y

    def set_solver(self, mode, rtol, maxit):
        self.check(self.L.fsx_set_solver(self.ctx, mode, rtol, maxit), "set_solver")

    def set_graphs(self, enable):
        self.check(self.L.fsx_set_graphs(self.ctx, 1 if enable else 0), "set_graphs")

    # single-stage calls (the fsx_* C-ABI, for stage-by-stage parity tests)
    def begin_step(self):
        self.check(self.L.fsx_begin_step(self.ctx), "begin_step")

    def convection(self):
        self.check(self.L.fsx_convection(self.ctx), "convection")

    def step1(self, it):
        self.check(self.L.fsx_step1(self.ctx, it), "step1")

    def step2(self, it):
        n = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.L.fsx_step2(self.ctx, it, ctypes.byref(n), ctypes.byref(rel))
        if rc not in (FSX_OK, FSX_ENOCONV):
            self.check(rc, "step2")
        return n.value, rel.value

    def step3(self, it):
        self.check(self.L.fsx_step3(self.ctx, it), "step3")

    def check_convergence(self, tol):
        c = ctypes.c_int()
        ru = ctypes.c_double()
        rp = ctypes.c_double()
        self.check(self.L.fsx_check_convergence(self.ctx, tol, ctypes.byref(c), ctypes.byref(ru), ctypes.byref(rp)))
        return bool(c.value), ru.value, rp.value

    def next_iteration(self):
        self.check(self.L.fsx_next_iteration(self.ctx))

    def kacc(self):
        self.check(self.L.fsx_kacc(self.ctx))

    def end_step(self, converge):
        s = ctypes.c_int()
        m = ctypes.c_double()
        self.check(self.L.fsx_end_step(self.ctx, converge, ctypes.byref(s), ctypes.byref(m)))
        return bool(s.value), m.value

    def timing_reset(self, enable=True):
        self.check(self.L.fsx_timing_reset(self.ctx, 1 if enable else 0))

    def timing(self):
        t = Timing()
        self.check(self.L.fsx_timing_get(self.ctx, ctypes.byref(t)))
        return dict(pcg_ms=t.pcg_ms, pcg_spmv_ms=t.pcg_spmv_ms, pcg_spmv_launches=t.pcg_spmv_launches,
                    other_ms=t.other_ms, pcg_iters=t.pcg_iters)

    def sync(self):
        self.check(self.L.fsx_sync(self.ctx))

    # learned preconditioner (fsx_unet_* of include/fsx.h)
    def unet_setup(self, grid, seed=0):
        nx, ny, nz = grid
        self.check(self.L.fsx_unet_setup(self.ctx, nx, ny, nz, seed), "unet_setup")
        self.unet_grid = (nz, ny, nx)

    def unet_nparams(self):
        n = ctypes.c_longlong()
        self.check(self.L.fsx_unet_params(self.ctx, ctypes.byref(n)), "unet_params")
        return n.value

    def unet_weights(self):
        w = np.empty(self.unet_nparams(), dtype=np.float32)
        self.check(self.L.fsx_unet_get_weights(self.ctx, _fp(w), w.size), "unet_get_weights")
        return w

    def unet_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self.check(self.L.fsx_unet_set_weights(self.ctx, _fp(w), w.size), "unet_set_weights")

    def unet_gain(self, set_to=0.0):
        g = ctypes.c_float()
        self.check(self.L.fsx_unet_gain(self.ctx, set_to, ctypes.byref(g)), "unet_gain")
        return g.value

    def unet_apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty_like(x)
        self.check(self.L.fsx_unet_apply(self.ctx, _fp(x), _fp(y)), "unet_apply")
        return y

    def unet_loss_grad(self, x, t, kind=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.unet_nparams(), dtype=np.float32)
        loss = ctypes.c_double()
        self.check(self.L.fsx_unet_loss_grad(self.ctx, kind, _fp(x), _fp(t), ctypes.byref(loss), _fp(g)), "unet_loss_grad")
        return loss.value, g

    def unet_capture(self, capacity):
        self.check(self.L.fsx_unet_capture(self.ctx, capacity), "unet_capture")

    def unet_captured(self):
        n = ctypes.c_int()
        self.check(self.L.fsx_unet_captured(self.ctx, ctypes.byref(n)), "unet_captured")
        return n.value

    def unet_prepare(self, ks, learned=False, append=False):
        k = np.ascontiguousarray(ks, dtype=np.int32)
        n = ctypes.c_int()
        flags = (1 if learned else 0) | (2 if append else 0)
        self.check(self.L.fsx_unet_prepare(self.ctx, _ip(k), k.size, flags, ctypes.byref(n)), "unet_prepare")
        return n.value

    def unet_prepare_random(self, count, sweeps=0, seed=0):
        """count training pairs from seeded drawn errors, appended; returns the pair count"""
        n = ctypes.c_int()
        self.check(self.L.fsx_unet_prepare_random(self.ctx, int(count), int(sweeps), int(seed), ctypes.byref(n)),
                   "unet_prepare_random")
        return n.value

    def unet_train(self, steps, batch=4, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        self.check(self.L.fsx_unet_train(self.ctx, steps, batch, lr, seed, ctypes.byref(loss)), "unet_train")
        return loss.value

    def unet_form(self, form, omega=0.8):
        self.check(self.L.fsx_unet_form(self.ctx, form, omega), "unet_form")

    def unet_form_get(self):
        f = ctypes.c_int()
        om = ctypes.c_float()
        self.check(self.L.fsx_unet_form_get(self.ctx, ctypes.byref(f), ctypes.byref(om)), "unet_form_get")
        return f.value, om.value

    def unet_precondition(self, r):
        r = np.ascontiguousarray(r, dtype=np.float32)
        z = np.zeros_like(r)
        self.check(self.L.fsx_unet_precondition(self.ctx, _fp(r), _fp(z)), "unet_precondition")
        return z

    def unet_limit(self, learned_maxit):
        self.check(self.L.fsx_unet_limit(self.ctx, learned_maxit), "unet_limit")

    def unet_stats(self):
        it = ctypes.c_int()
        fb = ctypes.c_int()
        self.check(self.L.fsx_unet_stats(self.ctx, ctypes.byref(it), ctypes.byref(fb)), "unet_stats")
        return it.value, fb.value
