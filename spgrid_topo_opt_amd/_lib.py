"""ctypes binding of libspgrid_topo_opt_amd.so (include/spgrid_topo_opt_amd.h).

There is no fallback: if the HIP library is not built, importing the solver
raises.  Build it with ``python -m spgrid_topo_opt_amd.build`` (or
``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspgrid_topo_opt_amd.so")

SPG_OK = 0
SPG_E_INVALID = -1
SPG_E_VALIDATION = -2
SPG_E_SINGULAR = -3
SPG_E_NOT_CONVERGED = -4
SPG_E_HIP = -5
SPG_E_NAN = -6
SPG_E_VERSION = -7

OP_APPLY, OP_RESIDUAL, OP_GS_SWEEP, OP_GS_BOUNDARY, OP_RESTRICT, OP_PROLONGATE, OP_STENCIL_RES, OP_VCYCLE, OP_BOTTOM, \
    OP_OPERATOR = range(10)

# exported C symbols (kept in sync with include/spgrid_topo_opt_amd.h)
EXPORTS = [
    "fem_solve", "fem_solve_test", "spg_last_error", "spg_version", "spg_device_count", "spg_create",
    "spg_destroy", "spg_update_density", "spg_set_forces", "spg_solve", "spg_get_displacement_device",
    "spg_solve_adjoint_device", "spg_sensitivity_device", "spg_num_levels", "spg_level_size", "spg_level_nodes",
    "spg_kernel", "spg_level_stencils", "spg_bench_apply", "spg_bench_kernel", "spg_get_info",
    "spg_update_density_device", "spg_filter_sensitivity", "spg_oc_trial", "spg_set_initial_guess_device",
    "spg_oc_trial_noalloc",
]


class SpgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[spg error {code}] {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("res", C.c_int32 * 3),
        ("dx", C.c_double),
        ("mu", C.c_double),
        ("lam", C.c_double),
        ("penalty", C.c_double),
        ("min_stiffness", C.c_double),
        ("min_density", C.c_double),
        ("mg_levels", C.c_int32),
        ("smooth_iters", C.c_int32),
        ("boundary_iters", C.c_int32),
        ("device", C.c_int32),
        ("bottom_dense_max", C.c_int32),
    ]


class SolveParams(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double),
        ("inner_tolerance", C.c_double),
        ("max_iterations", C.c_int32),
        ("max_passes", C.c_int32),
        ("precision", C.c_int32),
        ("print_residuals", C.c_int32),
        ("warm_start", C.c_int32),
        ("reserved0", C.c_int32),
        ("reference_linf", C.c_double),
    ]


class Info(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("passes", C.c_int32),
        ("converged", C.c_int32),
        ("levels", C.c_int32),
        ("r0_linf", C.c_double),
        ("r_linf", C.c_double),
        ("r_l2", C.c_double),
        ("setup_seconds", C.c_double),
        ("solve_seconds", C.c_double),
        ("active_nodes", C.c_int64),
        ("active_cells", C.c_int64),
        ("gpu_blocks", C.c_int64),
        ("bottom_dofs", C.c_int64),
        ("hbm_bytes", C.c_double),
        ("r_start_linf", C.c_double),
        ("warm_reused", C.c_int32),
        ("bottom_kind", C.c_int32),
        ("bottom_bytes", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        # One HIP runtime per process: when PyTorch is installed, load it first so
        # that this library binds (by soname) to the HIP runtime, rocBLAS and
        # rocSOLVER that torch already mapped; tensors' device pointers then
        # belong to the same runtime as our kernels.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"HIP library {LIB_PATH} is not built; run `python -m spgrid_topo_opt_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.spg_last_error.restype = C.c_char_p
        L.spg_version.restype = C.c_char_p
        L.spg_device_count.argtypes = [ip]
        L.spg_create.argtypes = [C.POINTER(Config), C.c_int64, ip, dp, C.c_int64, ip, ip, dp, C.POINTER(vp)]
        L.spg_destroy.argtypes = [vp]
        L.spg_destroy.restype = None
        L.spg_update_density.argtypes = [vp, dp]
        L.spg_set_forces.argtypes = [vp, C.c_int64, ip, dp]
        L.spg_solve.argtypes = [vp, C.POINTER(SolveParams), dp, C.POINTER(Info)]
        L.spg_get_displacement_device.argtypes = [vp, vp]
        L.spg_set_initial_guess_device.argtypes = [vp, vp]
        L.spg_solve_adjoint_device.argtypes = [vp, C.POINTER(SolveParams), vp, vp, C.POINTER(Info)]
        L.spg_sensitivity_device.argtypes = [vp, vp, vp, vp, dp]
        L.spg_num_levels.argtypes = [vp, ip]
        L.spg_level_size.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64)]
        L.spg_level_nodes.argtypes = [vp, C.c_int32, ip, C.POINTER(C.c_uint8)]
        L.spg_kernel.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp, dp, dp]
        L.spg_level_stencils.argtypes = [vp, C.c_int32, dp]
        L.spg_bench_apply.argtypes = [vp, C.c_int32, C.c_int32, dp]
        L.spg_bench_kernel.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp]
        L.spg_get_info.argtypes = [vp, C.POINTER(Info)]
        L.spg_update_density_device.argtypes = [vp, vp]
        L.spg_filter_sensitivity.argtypes = [C.c_int64, vp, vp, vp, vp, C.c_double, vp]
        L.spg_oc_trial.argtypes = [C.c_int64, vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32,
                                   C.c_double, vp, dp, dp]
        L.spg_oc_trial_noalloc.argtypes = [C.c_int64, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int32, C.c_double, vp, vp, vp]
        L.fem_solve.argtypes = [vp, vp]
        L.fem_solve.restype = None
        L.fem_solve_test.argtypes = [vp, vp]
        L.fem_solve_test.restype = None
        _lib = L
    return _lib


def _check(rc):
    if rc != SPG_OK:
        raise SpgError(rc, lib().spg_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


def device_count():
    n = C.c_int32(0)
    _check(lib().spg_device_count(C.byref(n)))
    return n.value


def make_config(prob, device=0):
    c = Config()
    c.res[:] = [int(v) for v in prob.res]
    c.dx = prob.dx
    c.mu = prob.mu
    c.lam = prob.lam
    c.penalty = prob.penalty
    c.min_stiffness = prob.min_stiffness
    c.min_density = prob.min_density
    c.mg_levels = prob.mg_levels
    c.smooth_iters = prob.smooth_iters
    c.boundary_iters = prob.boundary_iters
    c.device = device
    return c


class Solver:
    """GPU solver for one problem (spgrid_topo_opt_amd.workloads.Problem)."""

    def __init__(self, prob, device=0, bottom_dense_max=0):
        """bottom_dense_max: the largest bottom (unknowns) factored as a dense
        inverse, 0 for the library's default (spg_config.bottom_dense_max)"""
        self.prob = prob
        self.cfg = make_config(prob, device)
        self.cfg.bottom_dense_max = int(bottom_dense_max)
        bb = np.ascontiguousarray(prob.block_base, np.int32)
        de = np.ascontiguousarray(prob.density, np.float64)
        dc = np.ascontiguousarray(prob.dir_coord, np.int32).reshape(-1, 3)
        da = np.ascontiguousarray(prob.dir_axis, np.int32)
        dv = np.ascontiguousarray(prob.dir_value, np.float64)
        self.nblocks = bb.shape[0]
        h = C.c_void_p()
        _check(lib().spg_create(C.byref(self.cfg), bb.shape[0], _ip(bb), _dp(de), da.shape[0], _ip(dc), _ip(da),
                                _dp(dv), C.byref(h)))
        self.h = h
        self.set_forces(prob.force_coord, prob.force_value)

    def close(self):
        # at interpreter shutdown the module globals may already be gone
        L = globals().get("_lib")
        if getattr(self, "h", None) and L is not None:
            L.spg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_forces(self, coord, value):
        fc = np.ascontiguousarray(coord, np.int32).reshape(-1, 3)
        fv = np.ascontiguousarray(value, np.float64).reshape(-1, 3)
        _check(lib().spg_set_forces(self.h, fc.shape[0], _ip(fc), _dp(fv)))

    def update_density(self, density):
        d = np.ascontiguousarray(density, np.float64).reshape(self.nblocks, 64)
        _check(lib().spg_update_density(self.h, _dp(d)))

    def update_density_device(self, ptr):
        """densities from a device array (input block order, nblocks*64 float64)"""
        _check(lib().spg_update_density_device(self.h, C.c_void_p(ptr)))

    def solve(self, tolerance=1e-6, inner_tolerance=0.0, max_iterations=500, max_passes=1, precision=0,
              want_output=True, raise_on_error=True, print_residuals=0, warm_start=0, reference_linf=0.0):
        """print_residuals: 1 prints ||r||_inf of every PCG iteration to stdout, 2 to stderr.
        warm_start: 0 from u_D, 1 from the last solution / set guess if its
        residual is below reference_linf (or the zero-guess residual), 2 always."""
        p = SolveParams(tolerance, inner_tolerance, max_iterations, max_passes, precision, print_residuals,
                        warm_start, 0, reference_linf)
        info = Info()
        out = np.zeros((self.nblocks, 64, 4)) if want_output else None
        rc = lib().spg_solve(self.h, C.byref(p), _dp(out), C.byref(info))
        if raise_on_error:
            _check(rc)
        d = info.as_dict()
        d["rc"] = rc
        return out, d

    def solve_mixed(self, tolerance=1e-6, **kw):
        """FP32 inner MG-PCG with FP64 defect correction.  With inner_tolerance 0
        each pass aims at the target itself; a further pass runs only if the
        FP64 residual after a pass still misses it."""
        kw.setdefault("inner_tolerance", 0.0)
        kw.setdefault("max_passes", 30)
        return self.solve(tolerance=tolerance, precision=0, **kw)

    # --- device-pointer API (torch tensors' data_ptr())
    def displacement_device(self, ptr):
        _check(lib().spg_get_displacement_device(self.h, C.c_void_p(ptr)))

    def set_initial_guess_device(self, ptr):
        """starting guess of the next warm-started solve (input block order, device)"""
        _check(lib().spg_set_initial_guess_device(self.h, C.c_void_p(ptr)))

    def solve_adjoint_device(self, g_ptr, lam_ptr, tolerance=1e-6, inner_tolerance=0.0, max_iterations=500,
                             max_passes=30, precision=0):
        p = SolveParams(tolerance, inner_tolerance, max_iterations, max_passes, precision, 0, 0, 0, 0.0)
        info = Info()
        _check(lib().spg_solve_adjoint_device(self.h, C.byref(p), C.c_void_p(g_ptr), C.c_void_p(lam_ptr),
                                              C.byref(info)))
        return info.as_dict()

    def sensitivity_device(self, dc_ptr, lam_ptr=None, u_ptr=None):
        obj = C.c_double(0.0)
        _check(lib().spg_sensitivity_device(self.h, C.c_void_p(u_ptr) if u_ptr else None,
                                            C.c_void_p(lam_ptr) if lam_ptr else None, C.c_void_p(dc_ptr),
                                            C.byref(obj)))
        return obj.value

    # --- kernel-level hooks
    @property
    def levels(self):
        n = C.c_int32(0)
        _check(lib().spg_num_levels(self.h, C.byref(n)))
        return n.value

    def info(self):
        """solve-independent sizes: levels, active nodes/cells, blocks, bottom unknowns, device bytes"""
        info = Info()
        _check(lib().spg_get_info(self.h, C.byref(info)))
        return info.as_dict()

    def level_blocks(self, level):
        n = C.c_int64(0)
        _check(lib().spg_level_size(self.h, level, C.byref(n)))
        return n.value

    def level_nodes(self, level):
        nb = self.level_blocks(level)
        xyz = np.zeros((nb * 64, 3), np.int32)
        fl = np.zeros(nb * 64, np.uint8)
        _check(lib().spg_level_nodes(self.h, level, _ip(xyz), fl.ctypes.data_as(C.POINTER(C.c_uint8))))
        return xyz, fl

    def kernel(self, op, level, inp=None, b=None, precision=1, forward=True, out_level=None, sweeps=1):
        """one operator on host vectors; `sweeps` boundary sweeps for OP_GS_BOUNDARY"""
        nout = self.level_blocks(level if out_level is None else out_level) * 64
        out = np.zeros((nout, 3))
        inp = None if inp is None else np.ascontiguousarray(inp, np.float64)
        b = None if b is None else np.ascontiguousarray(b, np.float64)
        arg = int(bool(forward)) | (int(sweeps) << 8)
        _check(lib().spg_kernel(self.h, op, level, precision, arg, _dp(inp), _dp(b), _dp(out)))
        return out

    def stencils(self, level):
        nb = self.level_blocks(level)
        out = np.zeros((nb * 64, 243))
        _check(lib().spg_level_stencils(self.h, level, _dp(out)))
        return out.reshape(-1, 3, 3, 3, 3, 3)

    def bench_apply(self, precision=0, reps=20):
        s = C.c_double(0.0)
        _check(lib().spg_bench_apply(self.h, precision, reps, C.byref(s)))
        return s.value

    def bench_kernel(self, op, level=0, precision=0, arg=0, reps=10):
        """seconds per run of one operator on device-resident vectors"""
        s = C.c_double(0.0)
        _check(lib().spg_bench_kernel(self.h, op, level, precision, arg, reps, C.byref(s)))
        return s.value


def filter_sensitivity(nblocks, nbr_ptr, inside_ptr, dc_ptr, rho_ptr, filter_size, out_ptr):
    """HIP sensitivity filter over device arrays (see spg_filter_sensitivity)"""
    _check(lib().spg_filter_sensitivity(nblocks, C.c_void_p(nbr_ptr), C.c_void_p(inside_ptr), C.c_void_p(dc_ptr),
                                        C.c_void_p(rho_ptr), filter_size, C.c_void_p(out_ptr)))


def oc_trial(n, mask_ptr, rho_ptr, dc_ptr, mid, step, min_density, out_ptr=None, fixed_density=1.0,
             exclude_fixed=True, dc_shift=0.0):
    """one OC trial over device arrays (mask bit 0 updated, bit 1 fixed);
    returns (volume, max change)"""
    vol, ch = C.c_double(0.0), C.c_double(0.0)
    _check(lib().spg_oc_trial(n, C.c_void_p(mask_ptr), C.c_void_p(rho_ptr), C.c_void_p(dc_ptr), mid, step,
                              min_density, fixed_density, 1 if exclude_fixed else 0, dc_shift,
                              C.c_void_p(out_ptr) if out_ptr else None, C.byref(vol), C.byref(ch)))
    return vol.value, ch.value


def oc_trial_noalloc(n, mask_ptr, rho_ptr, dc_ptr, mid, step, min_density, result_ptr, out_ptr=None,
                     fixed_density=1.0, exclude_fixed=True, dc_shift=0.0, stream=0):
    """spg_oc_trial_noalloc: one OC trial enqueued on `stream` (a raw
    hipStream_t value, 0 = null stream) that allocates nothing; writes
    {volume, change} to the device array at result_ptr and does not
    synchronize"""
    _check(lib().spg_oc_trial_noalloc(n, C.c_void_p(mask_ptr), C.c_void_p(rho_ptr), C.c_void_p(dc_ptr), mid, step,
                                      min_density, fixed_density, 1 if exclude_fixed else 0, dc_shift,
                                      C.c_void_p(out_ptr) if out_ptr else None, C.c_void_p(result_ptr),
                                      C.c_void_p(stream) if stream else None))


def fem_solve_bytes(data: bytes, test=False) -> bytes:
    """Call the C entry point `fem_solve` (or `fem_solve_test`) on a serialized
    FEMInputs stream and return the serialized FEMOutputs stream."""
    data = bytes(data)
    buf = C.c_char_p(data)  # the bytes object's own storage: read, not copied
    out = C.c_void_p()
    fn = lib().fem_solve_test if test else lib().fem_solve
    fn(C.cast(buf, C.c_void_p), C.byref(out))
    if not out.value:
        raise SpgError(SPG_E_INVALID, "fem_solve produced no output")
    n = C.c_uint64.from_address(out.value).value
    res = C.string_at(out.value, n)
    _free_output(out.value)
    return res


_libc = None


def _free_output(ptr):
    """Outputs are allocated with new char[]; release them with delete[]
    (same allocator as malloc/free for char arrays in libstdc++)."""
    global _libc
    if _libc is None:
        _libc = C.CDLL("libstdc++.so.6")
        _libc._ZdaPv.argtypes = [C.c_void_p]
        _libc._ZdaPv.restype = None
    _libc._ZdaPv(C.c_void_p(ptr))
