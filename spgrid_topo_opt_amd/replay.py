"""Replay a dumped FEM solve: the `fem_solve` and `convert_fem_solve` tasks
(reference: topo_opt/misc.cpp:14-99).

The optimizer dumps every solve's serialized FEMInputs
(topo_opt/spgrid_topo_opt.cpp:1206-1222).  This tool reads one dump, applies
the same command-line overrides as the reference task (`cg_max_iterations=N`,
`minimum_stiffness=X`; any other field is an error; like the reference
solver, `fem_solve` ignores the krylov.max_iterations that
`cg_max_iterations` sets), clears the solver-state pointer and runs the
drop-in `fem_solve` of this library, or, with
`--convert`, reports the density range and NaN/Inf checks of
`convert_fem_solve` and writes a human-readable parameter dump.

  python -m spgrid_topo_opt_amd.replay FILE [field=value ...] [--out OUT] [--test]
  python -m spgrid_topo_opt_amd.replay FILE --convert [--keep-density] [--text OUT.txt]

FILE is either the raw serializer stream (`.tcb`) or a zip archive whose
first member is that stream (`.tcb.zip`).
"""
from __future__ import annotations

import argparse
import io
import math
import sys
import time
import zipfile

import numpy as np

from . import fem_interface as FI


def read_stream(path: str) -> bytes:
    """The serialized FEMInputs in `path` (raw stream or first zip member)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] == b"PK\x03\x04":
        with zipfile.ZipFile(io.BytesIO(data)) as z:
            names = z.namelist()
            if not names:
                raise ValueError(f"{path}: empty zip archive")
            data = z.read(names[0])
    return data


def read_inputs(path: str) -> FI.FEMInputs:
    return FI.FEMInputs.deserialize(read_stream(path))


def apply_overrides(param: FI.FEMInputs, overrides) -> None:
    """`field=value` overrides of the reference task (misc.cpp:29-42)."""
    for p in overrides:
        if "=" not in p:
            continue
        field, value = p.split("=", 1)
        if field == "cg_max_iterations":
            param.krylov.max_iterations = int(value)
        elif field == "minimum_stiffness":
            param.extra_doubles["minimum_stiffness"] = float(value)
        else:
            raise ValueError(f"Unknown field {field!r}")


def replay(path: str, overrides=(), test: bool = False):
    """Run one dumped solve; returns (FEMOutputs, seconds)."""
    iface = FI.FEMInterface(read_inputs(path))
    param = iface.param
    iface.preserve_output(len(param.density_base))
    param.set_solver_state_ptr(0)
    param.extra_ints.pop("keep_state", None)
    apply_overrides(param, overrides)
    t0 = time.perf_counter()
    out = iface.run(test=test)
    return out, time.perf_counter() - t0


def density_summary(param: FI.FEMInputs) -> dict:
    """min / max of the non-zero densities and NaN / Inf flags (misc.cpp:63-87)."""
    d = np.asarray(param.density, np.float64).reshape(-1)
    nz = d[d != 0]
    finite = nz[np.isfinite(nz)]
    return {
        "blocks": int(len(param.density_base)),
        "min_density": float(min(1.0, finite.min())) if finite.size else 1.0,
        "max_density": float(max(0.0, finite.max())) if finite.size else 0.0,
        "has_nan": bool(np.isnan(nz).any()),
        "has_inf": bool(np.isinf(nz).any()),
    }


def human_readable(param: FI.FEMInputs, keep_density: bool = False) -> str:
    """A plain-text dump of the parameters; the density blocks only with
    keep_density (the reference clears them unless asked, misc.cpp:89-91)."""
    lines = ["FEM Solve Parameters"]
    for k in ("version_major", "version_minor", "dx", "preserved_buffer_size", "use_density_only", "global_mu",
              "global_lambda", "mg_level", "explicit_mg_level", "pre_and_post_smoothing_iter",
              "bottom_smoothing_iter", "jacobi_damping", "caller_method"):
        lines.append(f"  {k} = {getattr(param, k)}")
    lines.append(f"  resolution = {list(param.resolution)}")
    for name in ("extra_strings", "extra_ints", "extra_doubles"):
        for k, v in sorted(getattr(param, name).items()):
            lines.append(f"  {name}[{k}] = {v}")
    k = param.krylov
    lines.append(f"  krylov = tolerance {k.tolerance}, max_iterations {k.max_iterations}, "
                 f"restart_iterations {k.restart_iterations}, print_residuals {k.print_residuals}")
    fc = np.zeros((0, 3)) if param.forces_coord is None else param.forces_coord
    lines.append(f"  forces: {len(fc)}")
    for c, f in zip(fc, param.forces_value if len(fc) else []):
        lines.append(f"    {list(map(int, c))} {list(map(float, f))}")
    dc = np.zeros((0, 3)) if param.dirichlet_coord is None else param.dirichlet_coord
    lines.append(f"  dirichlet_nodes: {len(dc)}")
    for c, a, v in zip(dc, param.dirichlet_axis if len(dc) else [], param.dirichlet_value if len(dc) else []):
        lines.append(f"    {list(map(int, c))} axis {int(a)} value {float(v)}")
    lines.append(f"  dirichlet_cells: {len(param.dirichlet_cells)}")
    lines.append(f"  density: resolution {list(param.density_resolution)}, {len(param.density_base)} blocks")
    if keep_density:
        for b, d in zip(param.density_base, param.density):
            lines.append(f"    {list(map(int, b))} " + " ".join(repr(float(x)) for x in d))
    return "\n".join(lines) + "\n"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("file")
    ap.add_argument("overrides", nargs="*", help="field=value (cg_max_iterations, minimum_stiffness)")
    ap.add_argument("--convert", action="store_true", help="report densities and write a text dump instead of solving")
    ap.add_argument("--keep-density", action="store_true", help="include the density blocks in the text dump")
    ap.add_argument("--text", default="human_readable.txt", help="text dump path for --convert")
    ap.add_argument("--out", help="write the serialized FEMOutputs here")
    ap.add_argument("--test", action="store_true", help="use the fake solver fem_solve_test (no GPU)")
    a = ap.parse_args(argv)
    if a.convert:
        param = read_inputs(a.file)
        s = density_summary(param)
        print(f"min_density = {s['min_density']}\nmax_density = {s['max_density']}")
        if s["has_nan"]:
            print("nan detected")
        if s["has_inf"]:
            print("inf detected")
        if not s["has_nan"] and not s["has_inf"]:
            print("Good. No nan nor inf detected.")
        with open(a.text, "w") as f:
            f.write(human_readable(param, a.keep_density))
        return 0
    out, sec = replay(a.file, a.overrides, test=a.test)
    res = out.residuals[-1] if out.residuals else math.nan
    print(f"fem_solve: success={out.success} error_code={out.error_code} residual={res:.3e} time={sec:.3f}s")
    if out.info:
        print(out.info)
    if a.out:
        with open(a.out, "wb") as f:
            f.write(out.serialize())
    return 0 if out.success else 1


if __name__ == "__main__":
    sys.exit(main())
