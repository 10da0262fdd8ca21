"""TEST INFRASTRUCTURE ONLY: ctypes binding of the C oracle (spgrid_oracle.c).

The oracle is a CPU FP64 restatement of the reference's narrow-band SPGrid
elasticity solve.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this module, and only as the checker / CPU
baseline -- never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libspgrid_oracle.so")

ACTIVE_NODE = 0x10
DIRICHLET_NODE_X = 0x20
DIRICHLET_NODE = 0xE0
ACTIVE_CELL = 0x100
BOUNDARY_NODE = 0x1000
MODE_FP64 = 0
MODE_REF = 1


def build() -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class _Problem(C.Structure):
    _fields_ = [
        ("res", C.c_int32 * 3),
        ("dx", C.c_double),
        ("mu0", C.c_double),
        ("lambda0", C.c_double),
        ("penalty", C.c_double),
        ("min_stiffness", C.c_double),
        ("min_density", C.c_double),
        ("mg_levels", C.c_int32),
        ("smooth_iters", C.c_int32),
        ("boundary_iters", C.c_int32),
        ("nblocks", C.c_int64),
        ("block_base", C.POINTER(C.c_int32)),
        ("density", C.POINTER(C.c_double)),
        ("ndir", C.c_int64),
        ("dir_coord", C.POINTER(C.c_int32)),
        ("dir_axis", C.POINTER(C.c_int32)),
        ("dir_value", C.POINTER(C.c_double)),
        ("nforce", C.c_int64),
        ("force_coord", C.POINTER(C.c_int32)),
        ("force_value", C.POINTER(C.c_double)),
    ]


class _SolveParams(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double),
        ("inner_tolerance", C.c_double),
        ("max_iterations", C.c_int32),
        ("max_passes", C.c_int32),
        ("mode", C.c_int32),
    ]


class _SolveInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("passes", C.c_int32),
        ("converged", C.c_int32),
        ("r0_linf", C.c_double),
        ("r_linf", C.c_double),
        ("setup_seconds", C.c_double),
        ("solve_seconds", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_build.restype = vp
        L.orc_build.argtypes = [C.POINTER(_Problem), C.c_char_p, C.c_int]
        L.orc_free.argtypes = [vp]
        L.orc_num_levels.argtypes = [vp]
        L.orc_num_levels.restype = C.c_int
        L.orc_num_nodes.argtypes = [vp, C.c_int]
        L.orc_num_nodes.restype = C.c_int64
        L.orc_node_xyz.argtypes = [vp, C.c_int, C.POINTER(C.c_int32)]
        L.orc_node_flags.argtypes = [vp, C.c_int, C.POINTER(C.c_uint32)]
        L.orc_cell_moduli.argtypes = [vp, dp, dp]
        L.orc_apply.argtypes = [vp, dp, dp]
        L.orc_gs_sweep.argtypes = [vp, C.c_int, dp, dp, C.c_int, C.c_int]
        L.orc_restrict.argtypes = [vp, C.c_int, dp, dp]
        L.orc_prolongate.argtypes = [vp, C.c_int, dp, dp]
        L.orc_stencil_residual.argtypes = [vp, C.c_int, dp, dp, dp]
        L.orc_get_stencils.argtypes = [vp, C.c_int, dp]
        L.orc_get_stencils.restype = C.c_int
        L.orc_vcycle.argtypes = [vp, dp, dp, C.c_int]
        L.orc_solve.argtypes = [vp, C.POINTER(_SolveParams), dp, dp, C.POINTER(_SolveInfo), C.c_char_p, C.c_int]
        L.orc_solve.restype = C.c_int
        L.orc_time_pcg_iterations.argtypes = [vp, C.c_int, C.c_int, dp]
        L.orc_time_pcg_iterations.restype = C.c_int
        L.orc_sensitivity.argtypes = [vp, dp, dp, dp]
        L.orc_sensitivity.restype = C.c_double
        L.orc_adjoint_sensitivity.argtypes = [vp, dp, dp, dp]
        L.orc_pack_displacements.argtypes = [vp, dp, dp, dp]
        L.orc_element_tables.argtypes = [dp, dp]
        L.orc_stiffness_mapping.argtypes = [C.c_double] * 4
        L.orc_stiffness_mapping.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def element_tables():
    kmu = np.zeros(576)
    kla = np.zeros(576)
    lib().orc_element_tables(_dp(kmu), _dp(kla))
    return kmu.reshape(8, 8, 3, 3), kla.reshape(8, 8, 3, 3)


def stiffness_mapping(penalty, min_stiffness, min_density, d):
    return lib().orc_stiffness_mapping(penalty, min_stiffness, min_density, d)


class Oracle:
    """Built oracle system for one problem (see spgrid_topo_opt_amd.workloads.Problem)."""

    def __init__(self, prob):
        self._keep = {
            "bb": np.ascontiguousarray(prob.block_base, dtype=np.int32),
            "de": np.ascontiguousarray(prob.density, dtype=np.float64),
            "dc": np.ascontiguousarray(prob.dir_coord, dtype=np.int32).reshape(-1, 3),
            "da": np.ascontiguousarray(prob.dir_axis, dtype=np.int32),
            "dv": np.ascontiguousarray(prob.dir_value, dtype=np.float64),
            "fc": np.ascontiguousarray(prob.force_coord, dtype=np.int32).reshape(-1, 3),
            "fv": np.ascontiguousarray(prob.force_value, dtype=np.float64).reshape(-1, 3),
        }
        k = self._keep
        p = _Problem()
        p.res[:] = [int(v) for v in prob.res]
        p.dx = prob.dx
        p.mu0 = prob.mu
        p.lambda0 = prob.lam
        p.penalty = prob.penalty
        p.min_stiffness = prob.min_stiffness
        p.min_density = prob.min_density
        p.mg_levels = prob.mg_levels
        p.smooth_iters = prob.smooth_iters
        p.boundary_iters = prob.boundary_iters
        p.nblocks = k["bb"].shape[0]
        p.block_base = _ip(k["bb"])
        p.density = _dp(k["de"])
        p.ndir = k["da"].shape[0]
        p.dir_coord = _ip(k["dc"])
        p.dir_axis = _ip(k["da"])
        p.dir_value = _dp(k["dv"])
        p.nforce = k["fc"].shape[0]
        p.force_coord = _ip(k["fc"])
        p.force_value = _dp(k["fv"])
        self._p = p
        err = C.create_string_buffer(512)
        self.h = lib().orc_build(C.byref(p), err, 512)
        if not self.h:
            raise RuntimeError("oracle build failed: " + err.value.decode())
        self.prob = prob

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    @property
    def levels(self):
        return lib().orc_num_levels(self.h)

    def nnodes(self, level=0):
        return int(lib().orc_num_nodes(self.h, level))

    def xyz(self, level=0):
        a = np.zeros((self.nnodes(level), 3), np.int32)
        lib().orc_node_xyz(self.h, level, _ip(a))
        return a

    def flags(self, level=0):
        a = np.zeros(self.nnodes(level), np.uint32)
        lib().orc_node_flags(self.h, level, a.ctypes.data_as(C.POINTER(C.c_uint32)))
        return a

    def moduli(self):
        n = self.nnodes(0)
        mu = np.zeros(n)
        la = np.zeros(n)
        lib().orc_cell_moduli(self.h, _dp(mu), _dp(la))
        return mu, la

    def apply(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        f = np.zeros_like(u)
        lib().orc_apply(self.h, _dp(u), _dp(f))
        return f

    def gs_sweep(self, level, u, b, forward=True, boundary=False):
        u = np.array(u, dtype=np.float64).reshape(-1, 3)
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
        lib().orc_gs_sweep(self.h, level, _dp(u), _dp(b), int(forward), int(boundary))
        return u

    def restrict(self, coarse_level, fine):
        fine = np.ascontiguousarray(fine, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((self.nnodes(coarse_level), 3))
        lib().orc_restrict(self.h, coarse_level, _dp(fine), _dp(out))
        return out

    def prolongate(self, fine_level, fine, coarse):
        fine = np.array(fine, dtype=np.float64).reshape(-1, 3)
        coarse = np.ascontiguousarray(coarse, dtype=np.float64).reshape(-1, 3)
        lib().orc_prolongate(self.h, fine_level, _dp(fine), _dp(coarse))
        return fine

    def stencil_residual(self, level, u, b):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
        r = np.zeros_like(u)
        lib().orc_stencil_residual(self.h, level, _dp(u), _dp(b), _dp(r))
        return r

    def stencils(self, level):
        out = np.zeros((self.nnodes(level), 243))
        if lib().orc_get_stencils(self.h, level, _dp(out)) != 0:
            raise ValueError("level %d has no stencils" % level)
        return out.reshape(-1, 3, 3, 3, 3, 3)

    def vcycle(self, b, mode=MODE_FP64):
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
        z = np.zeros_like(b)
        lib().orc_vcycle(self.h, _dp(b), _dp(z), mode)
        return z

    def solve(self, tolerance=1e-6, inner_tolerance=0.0, max_iterations=500, max_passes=1, mode=MODE_FP64):
        n = self.nnodes(0)
        u = np.zeros((n, 3))
        r = np.zeros((n, 3))
        prm = _SolveParams(tolerance, inner_tolerance, max_iterations, max_passes, mode)
        info = _SolveInfo()
        err = C.create_string_buffer(512)
        rc = lib().orc_solve(self.h, C.byref(prm), _dp(u), _dp(r), C.byref(info), err, 512)
        out = {k: getattr(info, k) for k, _ in _SolveInfo._fields_}
        out["rc"] = rc
        out["error"] = err.value.decode()
        return u, r, out

    def time_pcg_iterations(self, iters, mode=MODE_FP64):
        sec = C.c_double(0)
        it = lib().orc_time_pcg_iterations(self.h, iters, mode, C.byref(sec))
        return it, sec.value

    def sensitivity(self, u):
        n = self.nnodes(0)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        dc = np.zeros(n)
        se = np.zeros(n)
        obj = lib().orc_sensitivity(self.h, _dp(u), _dp(dc), _dp(se))
        return obj, dc, se

    def adjoint_sensitivity(self, u, lam):
        n = self.nnodes(0)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1, 3)
        g = np.zeros(n)
        lib().orc_adjoint_sensitivity(self.h, _dp(u), _dp(lam), _dp(g))
        return g

    def pack(self, u, r=None):
        nb = self._keep["bb"].shape[0]
        out = np.zeros((nb, 64, 4))
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        rr = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3) if r is not None else None
        lib().orc_pack_displacements(self.h, _dp(u), _dp(rr) if rr is not None else None, _dp(out))
        return out
