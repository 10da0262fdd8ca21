"""Drop-in boundary: FEMInputs / FEMOutputs byte streams and the C entry points.

Mirrors the reference's only test, TC_TEST("fem_interface")
(topo_opt/interface_test.cpp:40-269): fixture fields, the fake solver
u_p = rho + p, the .tcb round trip and the 64-bit state-pointer split.  Byte
streams are compared with SHA-256 digests recorded from the reference's own
serializer (tests/golden/make_wire_golden.py), and, when the reference
sources are mounted, directly against oracle/_ref/wire_ref.  Runs on the CPU:
fem_solve_test and the error paths of fem_solve touch no GPU.
"""
import hashlib
import json
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from spgrid_topo_opt_amd import fem_interface as FI
from spgrid_topo_opt_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = json.load(open(os.path.join(HERE, "golden", "wire_golden.json")))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "wire_ref")
HAVE_REF = os.path.exists("/root/reference/topo_opt/fem_interface.h")


def interface_fixture(scale=1) -> FI.FEMInputs:
    """The interface test's FEMInputs (interface_test.cpp:45-166), resolution / scale."""
    res = [128 // scale, 256 // scale, 512 // scale]
    p = FI.FEMInputs(resolution=res, dx=float(np.float32(0.01)), use_density_only=True)
    p.forces_coord = np.array([[2 * i + 1, i * i + 14, i // 3 + 1] for i in range(17)], np.int32)
    p.forces_value = np.array([[math.sqrt(i), float(i * i), i * i * i / 10000.0] for i in range(17)])
    p.dirichlet_coord = np.array([[95, 6, 3], [94, 11, 16]], np.int32)
    p.dirichlet_axis = np.array([2, 1], np.int32)
    p.dirichlet_value = np.array([0.1234, 0.5432])
    p.krylov = FI.KrylovSolverParameters(1e-5, 10000, 1000, True)
    p.caller_method = "topology_optimization"
    p.extra_ints = {"multigrid_num_pre_smoothing": 2, "multigrid_num_post_smoothing": 3}
    p.extra_strings = {"log_file_path": "/tmp/fem_solver.log"}
    p.extra_doubles = {"smoother_jacobi_damping": float(np.float32(0.66667))}
    p.global_mu, p.global_lambda = 2.0, 1.0
    p.pre_and_post_smoothing_iter = 0
    p.density_resolution = list(res)
    centers = [r // 2 for r in res]
    depth = 6

    def node_inside(c):
        return any(centers[i] - depth < c[i] < centers[i] + depth for i in range(3))

    def cell_inside(c):
        return any(centers[i] - depth < c[i] < centers[i] + depth - 1 for i in range(3))

    bases, data = [], []
    for i in range(0, res[0], 4):
        for j in range(0, res[1], 4):
            for k in range(0, res[2], 4):
                # a block is kept if any of its 64 entries lies in the cross
                if not any(
                        node_inside((i + a, j + b, k + c)) for a in range(4) for b in range(4) for c in range(4)):
                    continue
                d = np.zeros(64)
                for ii in range(4):
                    for jj in range(4):
                        for kk in range(4):
                            if cell_inside((i + ii, j + jj, k + kk)):
                                d[ii * 16 + jj * 4 + kk] = i * jj + k - j + kk * ii
                bases.append((i, j, k))
                data.append(d)
    p.density_base = np.array(bases, np.int32)
    p.density = np.array(data)
    iface = FI.FEMInterface(p)
    iface.preserve_output(len(bases))
    return p


@pytest.fixture(scope="module")
def fx8():
    return interface_fixture(8)


def test_fixture_bytes_match_reference_serializer(fx8):
    data = fx8.serialize()
    g = GOLD["scale8"]
    assert len(data) == g["inputs_bytes"]
    assert hashlib.sha256(data).hexdigest() == g["inputs_sha256"]
    assert len(fx8.density_base) == g["blocks"]


def test_fixture_bytes_match_reference_serializer_scale4():
    data = interface_fixture(4).serialize()
    assert hashlib.sha256(data).hexdigest() == GOLD["scale4"]["inputs_sha256"]


def test_reading_back_known_answers(fx8):
    # interface_test.cpp:199-261
    p = FI.FEMInputs.deserialize(fx8.serialize())
    assert p.version_major == 3 and p.version_minor == 1
    assert p.resolution == [16, 32, 64]
    assert p.dx == float(np.float32(0.01)) and p.use_density_only
    assert len(p.forces_coord) == 17
    for i in range(17):
        assert list(p.forces_coord[i]) == [2 * i + 1, i * i + 14, i // 3 + 1]
        assert p.forces_value[i][0] == math.sqrt(1.0 * i)
        assert p.forces_value[i][1] == i * i and p.forces_value[i][2] == i * i * i / 10000.0
    assert list(p.dirichlet_coord[0]) == [95, 6, 3] and p.dirichlet_axis[0] == 2 and p.dirichlet_value[0] == 0.1234
    assert list(p.dirichlet_coord[1]) == [94, 11, 16] and p.dirichlet_axis[1] == 1 and p.dirichlet_value[1] == 0.5432
    assert (p.krylov.tolerance, p.krylov.max_iterations, p.krylov.restart_iterations, p.krylov.print_residuals) == (
        1e-5, 10000, 1000, True)
    assert p.caller_method == "topology_optimization"
    assert p.extra_ints == {"multigrid_num_pre_smoothing": 2, "multigrid_num_post_smoothing": 3}
    assert p.extra_strings["log_file_path"] == "/tmp/fem_solver.log"
    assert p.extra_doubles["smoother_jacobi_damping"] == float(np.float32(0.66667))
    assert p.global_mu == 2.0 and p.global_lambda == 1.0
    assert p.serialize() == fx8.serialize()


def test_state_pointer_split():
    # interface_test.cpp:264-267
    p = FI.FEMInputs()
    assert p.get_solver_state_ptr() is None
    p.set_solver_state_ptr(0x1234567890543200)
    assert p.get_solver_state_ptr() == 0x1234567890543200
    p.set_solver_state_ptr(0xFFFFFFFF80000001)
    assert p.get_solver_state_ptr() == 0xFFFFFFFF80000001


def test_fake_solver_through_c_abi(fx8):
    # fem_solve_test (interface_test.cpp:7-34) through the product's C ABI
    out_bytes = _lib.fem_solve_bytes(fx8.serialize(), test=True)
    assert hashlib.sha256(out_bytes).hexdigest() == GOLD["scale8"]["fake_outputs_sha256"]
    out = FI.FEMOutputs.deserialize(out_bytes)
    assert out.displacements.shape == (len(fx8.density_base), 64, 4)
    assert (out.displacement_base == fx8.density_base).all()
    for p in range(3):
        np.testing.assert_array_equal(out.displacements[:, :, p], fx8.density + p)


@pytest.mark.skipif(not HAVE_REF, reason="reference sources not mounted")
def test_reference_reader_accepts_product_streams(fx8, tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    fin = tmp_path / "in.tcb"
    fin.write_bytes(fx8.serialize())
    # reference re-encodes our FEMInputs bit-identically
    fre = tmp_path / "re.tcb"
    subprocess.run([REF_BIN, "reencode", str(fin), str(fre)], check=True)
    assert fre.read_bytes() == fin.read_bytes()
    # reference parses the FEMOutputs written by the product's C ABI
    fout = tmp_path / "out.bin"
    fout.write_bytes(_lib.fem_solve_bytes(fx8.serialize(), test=True))
    r = subprocess.run([REF_BIN, "outputs", str(fout)], check=True, capture_output=True, text=True).stdout
    assert "blocks=512" in r and "success=1" in r
    assert f"head={len(fout.read_bytes())}" in r


def test_fem_solve_reports_errors_without_throwing():
    # a stream with a wrong major version -> success=0, version error
    p = interface_fixture(8)
    p.version_major = 2
    out = FI.FEMOutputs.deserialize(_lib.fem_solve_bytes(p.serialize()))
    assert not out.success and out.error_code == _lib.SPG_E_VERSION
    # mu/lambda mode is unsupported (SPGrid_CPU_Solver.cpp:131-135)
    data = bytearray(interface_fixture(8).serialize())
    # flip use_density_only: it sits after the 3 maps, dx and preserved_buffer_size
    q = FI.FEMInputs.deserialize(bytes(data))
    w = FI._W()
    FI._w_base(w, q)
    off = len(b"".join(w.parts)) + 8 + 8
    data[off] = 0
    out = FI.FEMOutputs.deserialize(_lib.fem_solve_bytes(bytes(data) + b"\0" * 12 + (0).to_bytes(8, "little")))
    assert not out.success
    assert "Unsupported mode" in out.info or "truncated" in out.info or out.error_code != 0


def test_fem_solve_without_device_fails_cleanly():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present; the no-device path is exercised only on CPU hosts")
    from spgrid_topo_opt_amd import workloads as W
    p = W.box(8, 4, 4)
    out = FI.FEMOutputs.deserialize(_lib.fem_solve_bytes(FI.FEMInputs.from_problem(p).serialize()))
    assert not out.success and out.error_code == _lib.SPG_E_HIP
    assert "device" in out.info


def test_index_range_is_checked_before_any_device_call(monkeypatch):
    """Level vectors are indexed with 32-bit ints on the device: a level
    beyond that range is refused on the host (SPG_E_INVALID), before the
    device is touched.  SPG_INDEX_LIMIT lowers the limit to reach the check
    with a small problem."""
    from spgrid_topo_opt_amd import workloads as W
    p = W.box(16, 8, 8)
    monkeypatch.setenv("SPG_INDEX_LIMIT", str(192 * 10))
    out = FI.FEMOutputs.deserialize(_lib.fem_solve_bytes(FI.FEMInputs.from_problem(p).serialize()))
    assert not out.success and out.error_code == _lib.SPG_E_INVALID
    assert "32-bit" in out.info and "at most 10 blocks" in out.info
    with pytest.raises(_lib.SpgError, match="32-bit"):
        _lib.Solver(p)


def test_negative_counts_are_refused_before_any_device_call():
    """spg_create and spg_set_forces check their counts on the host: a
    negative block, Dirichlet or force count returns SPG_E_INVALID"""
    import ctypes as C
    from spgrid_topo_opt_amd import workloads as W
    p = W.box(8, 4, 4)
    cfg = _lib.make_config(p)
    L = _lib.lib()
    bb = np.ascontiguousarray(p.block_base, np.int32)
    de = np.ascontiguousarray(p.density, np.float64)
    h = C.c_void_p()
    ip, dp = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    for nb, nd in ((-1, 0), (bb.shape[0], -3)):
        rc = L.spg_create(C.byref(cfg), nb, bb.ctypes.data_as(ip), de.ctypes.data_as(dp), nd, None, None, None,
                          C.byref(h))
        assert rc == _lib.SPG_E_INVALID and "negative" in L.spg_last_error().decode()
        assert not h.value


@pytest.mark.parametrize("shift", [-8, -4, 4 * 60, 4 * 61])
def test_padding_violation_names_the_oracles_page(shift):
    """a block set too close to the grid edge is refused on the host
    (Multigrid_Hierarchy.h:145-157) with the page the oracle names"""
    import re
    from oracle import oracle as O
    from spgrid_topo_opt_amd import workloads as W
    p = W.box(8, 4, 4)
    r = W.Problem(**{**p.__dict__, "block_base": p.block_base + shift})
    with pytest.raises(RuntimeError, match="padding") as eo:
        O.Oracle(r)
    with pytest.raises(_lib.SpgError, match="padding") as ep:
        _lib.Solver(r)
    page = re.compile(r"page origin \((-?\d+),(-?\d+),(-?\d+)\)")
    assert page.search(str(eo.value)).groups() == page.search(str(ep.value)).groups()


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "spgrid_topo_opt_amd.h")).read()
    import re
    declared = set(re.findall(r"^\s*(?:const char \*|void|int)\s*(\w+)\(", hdr, re.M))
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)
    L = _lib.lib()
    for s in declared:
        assert hasattr(L, s), s
