"""Python mirror of the reference's FEM interface (topo_opt/fem_interface.h).

``FEMInputs`` / ``FEMOutputs`` hold the same fields as the C++ structs
(fem_interface.h:165-390) and (de)serialize to the same byte stream
(topo_opt/serialization.h BinarySerializer: u64 total length, PODs raw,
vectors/strings/maps length-prefixed, maps in key order).  ``FEMInterface.run``
plays the caller side of fem_interface.h:464-478: serialize, call the C entry
point ``fem_solve``, parse the outputs, release the callee's buffer.
"""
from __future__ import annotations

import dataclasses
import struct
from typing import Dict, List

import numpy as np

VERSION_MAJOR = 3
VERSION_MINOR = 1


class _W:
    def __init__(self):
        self.parts = [b"\0" * 8]

    def pod(self, fmt, *v):
        self.parts.append(struct.pack("<" + fmt, *v))

    def raw(self, b):
        self.parts.append(bytes(b))

    def string(self, s: str):
        b = s.encode()
        self.pod("Q", len(b))
        self.raw(b)

    def done(self) -> bytes:
        out = b"".join(self.parts)
        return struct.pack("<Q", len(out)) + out[8:]


class _R:
    def __init__(self, data: bytes):
        self.d = memoryview(data)
        (self.total,) = struct.unpack_from("<Q", self.d, 0)
        if self.total > len(data):
            raise ValueError("stream truncated")
        self.h = 8

    def pod(self, fmt):
        n = struct.calcsize("<" + fmt)
        if self.h + n > self.total:
            raise ValueError("stream truncated")
        v = struct.unpack_from("<" + fmt, self.d, self.h)
        self.h += n
        return v if len(v) > 1 else v[0]

    def raw(self, n):
        if self.h + n > self.total:
            raise ValueError("stream truncated")
        b = bytes(self.d[self.h:self.h + n])
        self.h += n
        return b

    def string(self):
        n = self.pod("Q")
        return self.raw(n).decode()


@dataclasses.dataclass
class KrylovSolverParameters:
    tolerance: float = 1e-4
    max_iterations: int = 500
    restart_iterations: int = 0
    print_residuals: bool = False


def _w_scalar_grid(w, resolution, base, data):
    w.pod("3i", *[int(v) for v in resolution])
    n = 0 if base is None else len(base)
    w.pod("Q", n)
    if n:
        rec = np.zeros(n, dtype=[("b", "<i4", 3), ("d", "<f8", 64)])
        rec["b"] = base
        rec["d"] = np.asarray(data, np.float64).reshape(n, 64)
        w.raw(rec.tobytes())


def _r_scalar_grid(r):
    res = r.pod("3i")
    n = r.pod("Q")
    rec = np.frombuffer(r.raw(n * (12 + 512)), dtype=[("b", "<i4", 3), ("d", "<f8", 64)], count=n)
    return list(res), rec["b"].copy(), rec["d"].copy()


def _w_base(w, o):
    w.pod("ii", o.version_major, o.version_minor)
    w.pod("Q", len(o.extra_strings))
    for k in sorted(o.extra_strings):
        w.string(k)
        w.string(o.extra_strings[k])
    w.pod("Q", len(o.extra_ints))
    for k in sorted(o.extra_ints):
        w.string(k)
        w.pod("i", o.extra_ints[k])
    w.pod("Q", len(o.extra_doubles))
    for k in sorted(o.extra_doubles):
        w.string(k)
        w.pod("d", o.extra_doubles[k])


def _r_base(r, o):
    o.version_major, o.version_minor = r.pod("ii")
    if o.version_major != VERSION_MAJOR or o.version_minor > VERSION_MINOR:
        raise ValueError(f"FEM interface version {o.version_major}.{o.version_minor} not supported")
    o.extra_strings = {}
    for _ in range(r.pod("Q")):
        k = r.string()
        o.extra_strings[k] = r.string()
    o.extra_ints = {}
    for _ in range(r.pod("Q")):
        k = r.string()
        o.extra_ints[k] = r.pod("i")
    o.extra_doubles = {}
    for _ in range(r.pod("Q")):
        k = r.string()
        o.extra_doubles[k] = r.pod("d")


class _StateMixin:
    def set_solver_state_ptr(self, addr: int):
        """fem_interface.h:193-200: split a 64-bit pointer into two ints."""
        lo = addr & 0xFFFFFFFF
        hi = (addr >> 32) & 0xFFFFFFFF
        self.extra_ints["state_ptr_lower"] = lo - (1 << 32) if lo >= (1 << 31) else lo
        self.extra_ints["state_ptr_higher"] = hi - (1 << 32) if hi >= (1 << 31) else hi

    def get_solver_state_ptr(self):
        if "state_ptr_lower" not in self.extra_ints:
            return None
        lo = self.extra_ints["state_ptr_lower"] & 0xFFFFFFFF
        hi = self.extra_ints["state_ptr_higher"] & 0xFFFFFFFF
        return (hi << 32) + lo


@dataclasses.dataclass
class FEMInputs(_StateMixin):
    version_major: int = VERSION_MAJOR
    version_minor: int = VERSION_MINOR
    extra_strings: Dict[str, str] = dataclasses.field(default_factory=dict)
    extra_ints: Dict[str, int] = dataclasses.field(default_factory=dict)
    extra_doubles: Dict[str, float] = dataclasses.field(default_factory=dict)
    resolution: List[int] = dataclasses.field(default_factory=lambda: [0, 0, 0])
    dx: float = 0.0
    preserved_buffer_size: int = 0
    use_density_only: bool = True
    global_mu: float = 0.0
    global_lambda: float = 0.0
    density_resolution: List[int] = dataclasses.field(default_factory=lambda: [0, 0, 0])
    density_base: np.ndarray = None      # (nb,3) int32
    density: np.ndarray = None           # (nb,64) float64
    mg_level: int = 0
    explicit_mg_level: int = 0
    pre_and_post_smoothing_iter: int = 1
    bottom_smoothing_iter: int = 0
    jacobi_damping: float = 0.0
    forces_coord: np.ndarray = None      # (nf,3) int32
    forces_value: np.ndarray = None      # (nf,3) float64
    dirichlet_coord: np.ndarray = None   # (nd,3) int32
    dirichlet_axis: np.ndarray = None    # (nd,) int32
    dirichlet_value: np.ndarray = None   # (nd,) float64
    dirichlet_cells: list = dataclasses.field(default_factory=list)
    krylov: KrylovSolverParameters = dataclasses.field(default_factory=KrylovSolverParameters)
    caller_method: str = ""

    # named extras with the reference's defaults (fem_interface.h:278-328)
    def get_int(self, k, d):
        return self.extra_ints.get(k, d)

    def get_double(self, k, d):
        return self.extra_doubles.get(k, d)

    def serialize(self) -> bytes:
        """FEMInputs::io (fem_interface.h:252-276)."""
        w = _W()
        _w_base(w, self)
        w.pod("d", self.dx)
        w.pod("Q", self.preserved_buffer_size)
        w.pod("B", 1 if self.use_density_only else 0)
        if not self.use_density_only:
            raise NotImplementedError("mu/lambda grids are not produced by this driver")
        w.pod("dd", self.global_mu, self.global_lambda)
        _w_scalar_grid(w, self.density_resolution, self.density_base, self.density)
        w.pod("iiii", self.mg_level, self.explicit_mg_level, self.pre_and_post_smoothing_iter, self.bottom_smoothing_iter)
        w.pod("d", self.jacobi_damping)
        w.pod("3i", *[int(v) for v in self.resolution])
        fc = np.zeros((0, 3), np.int32) if self.forces_coord is None else np.asarray(self.forces_coord, np.int32)
        fv = np.zeros((0, 3)) if self.forces_value is None else np.asarray(self.forces_value, np.float64)
        w.pod("Q", fc.shape[0])
        if fc.shape[0]:
            rec = np.zeros(fc.shape[0], dtype=[("c", "<i4", 3), ("f", "<f8", 3)])
            rec["c"] = fc
            rec["f"] = fv
            w.raw(rec.tobytes())
        dc = np.zeros((0, 3), np.int32) if self.dirichlet_coord is None else np.asarray(self.dirichlet_coord, np.int32)
        w.pod("Q", dc.shape[0])
        if dc.shape[0]:
            rec = np.zeros(dc.shape[0], dtype=[("c", "<i4", 3), ("a", "<i4"), ("v", "<f8")])
            rec["c"] = dc
            rec["a"] = self.dirichlet_axis
            rec["v"] = self.dirichlet_value
            w.raw(rec.tobytes())
        w.pod("Q", len(self.dirichlet_cells))
        for c, a, v in self.dirichlet_cells:
            w.pod("3iid", *c, a, v)
        k = self.krylov
        w.pod("diiB", k.tolerance, k.max_iterations, k.restart_iterations, 1 if k.print_residuals else 0)
        w.string(self.caller_method)
        return w.done()

    @classmethod
    def deserialize(cls, data: bytes) -> "FEMInputs":
        r = _R(data)
        o = cls()
        _r_base(r, o)
        o.dx = r.pod("d")
        o.preserved_buffer_size = r.pod("Q")
        o.use_density_only = bool(r.pod("B"))
        if not o.use_density_only:
            raise NotImplementedError("mu/lambda grids")
        o.global_mu, o.global_lambda = r.pod("dd")
        o.density_resolution, o.density_base, o.density = _r_scalar_grid(r)
        o.mg_level, o.explicit_mg_level, o.pre_and_post_smoothing_iter, o.bottom_smoothing_iter = r.pod("iiii")
        o.jacobi_damping = r.pod("d")
        o.resolution = list(r.pod("3i"))
        n = r.pod("Q")
        rec = np.frombuffer(r.raw(n * 36), dtype=[("c", "<i4", 3), ("f", "<f8", 3)], count=n)
        o.forces_coord, o.forces_value = rec["c"].copy(), rec["f"].copy()
        n = r.pod("Q")
        rec = np.frombuffer(r.raw(n * 24), dtype=[("c", "<i4", 3), ("a", "<i4"), ("v", "<f8")], count=n)
        o.dirichlet_coord, o.dirichlet_axis, o.dirichlet_value = rec["c"].copy(), rec["a"].copy(), rec["v"].copy()
        n = r.pod("Q")
        o.dirichlet_cells = []
        for _ in range(n):
            v = r.pod("3iid")
            o.dirichlet_cells.append((list(v[:3]), v[3], v[4]))
        t, mi, ri, pr = r.pod("diiB")
        o.krylov = KrylovSolverParameters(t, mi, ri, bool(pr))
        o.caller_method = r.string()
        if r.h != r.total:
            raise ValueError("trailing bytes after FEMInputs")
        return o

    @classmethod
    def from_problem(cls, p, tolerance=1e-4, max_iterations=500, extra_ints=None, extra_doubles=None):
        """Pack a workloads.Problem the way the optimizer does
        (topo_opt/spgrid_topo_opt.cpp:989-1200)."""
        ei = {"boundary_smoothing": p.boundary_iters}
        ei.update(extra_ints or {})
        ed = {"penalty": p.penalty, "minimum_stiffness": p.min_stiffness, "minimum_density": p.min_density}
        ed.update(extra_doubles or {})
        return cls(extra_ints=ei, extra_doubles=ed, resolution=list(p.res), dx=p.dx, use_density_only=True,
                   global_mu=p.mu, global_lambda=p.lam, density_resolution=list(p.res), density_base=p.block_base,
                   density=p.density, mg_level=p.mg_levels, pre_and_post_smoothing_iter=p.smooth_iters,
                   forces_coord=p.force_coord, forces_value=p.force_value, dirichlet_coord=p.dir_coord,
                   dirichlet_axis=p.dir_axis, dirichlet_value=p.dir_value,
                   krylov=KrylovSolverParameters(tolerance, max_iterations, 0, False), caller_method="spgrid_topo_opt_amd")


@dataclasses.dataclass
class FEMOutputs(_StateMixin):
    version_major: int = VERSION_MAJOR
    version_minor: int = VERSION_MINOR
    extra_strings: Dict[str, str] = dataclasses.field(default_factory=dict)
    extra_ints: Dict[str, int] = dataclasses.field(default_factory=dict)
    extra_doubles: Dict[str, float] = dataclasses.field(default_factory=dict)
    resolution: List[int] = dataclasses.field(default_factory=lambda: [0, 0, 0])
    displacement_base: np.ndarray = None   # (nb,3)
    displacements: np.ndarray = None       # (nb,64,4)
    residual_field_resolution: List[int] = dataclasses.field(default_factory=lambda: [0, 0, 0])
    residual_field_base: np.ndarray = None
    residual_field: np.ndarray = None
    info: str = ""
    residuals: List[float] = dataclasses.field(default_factory=list)
    error_code: int = 0
    success: bool = False

    def serialize(self) -> bytes:
        w = _W()
        _w_base(w, self)
        w.pod("3i", *self.resolution)
        n = 0 if self.displacement_base is None else len(self.displacement_base)
        w.pod("Q", n)
        if n:
            rec = np.zeros(n, dtype=[("b", "<i4", 3), ("d", "<f8", (64, 4))])
            rec["b"] = self.displacement_base
            rec["d"] = self.displacements
            w.raw(rec.tobytes())
        _w_scalar_grid(w, self.residual_field_resolution, self.residual_field_base, self.residual_field)
        w.string(self.info)
        w.pod("Q", len(self.residuals))
        for v in self.residuals:
            w.pod("d", v)
        w.pod("iB", self.error_code, 1 if self.success else 0)
        return w.done()

    @classmethod
    def deserialize(cls, data: bytes) -> "FEMOutputs":
        """FEMOutputs::io (fem_interface.h:373-381)."""
        r = _R(data)
        o = cls()
        _r_base(r, o)
        o.resolution = list(r.pod("3i"))
        n = r.pod("Q")
        rec = np.frombuffer(r.raw(n * (12 + 2048)), dtype=[("b", "<i4", 3), ("d", "<f8", (64, 4))], count=n)
        o.displacement_base, o.displacements = rec["b"].copy(), rec["d"].copy()
        o.residual_field_resolution, o.residual_field_base, o.residual_field = _r_scalar_grid(r)
        o.info = r.string()
        n = r.pod("Q")
        o.residuals = [r.pod("d") for _ in range(n)]
        o.error_code, s = r.pod("iB")
        o.success = bool(s)
        if r.h != r.total:
            raise ValueError("trailing bytes after FEMOutputs")
        return o


class FEMInterface:
    """Caller side of fem_interface.h:392-487."""

    def __init__(self, param: FEMInputs = None):
        self.param = param or FEMInputs()
        self.outputs = None

    def preserve_output(self, num_active_blocks: int):
        # fem_interface.h:445-454: 1 MiB + sizeof(VectorGrid::Block) per active block;
        # the struct is int[3] + double[64][4], padded to 8-byte alignment: 2064 bytes
        self.param.preserved_buffer_size = (1 << 20) + num_active_blocks * 2064

    def run(self, test: bool = False) -> FEMOutputs:
        from . import _lib
        out = _lib.fem_solve_bytes(self.param.serialize(), test=test)
        self.outputs = FEMOutputs.deserialize(out)
        return self.outputs
