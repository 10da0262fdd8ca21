"""The dump-replay tool (reference tasks fem_solve / convert_fem_solve,
topo_opt/misc.cpp:14-99): reading raw and zipped streams, the reference's
overrides, the density report, and (GPU) that a replayed dump gives the same
solution as solving the in-memory inputs.

The zip container of the reference's `.tcb.zip` dumps comes from the taichi
runtime, which is not part of the reference tree: reading a zip archive whose
first member is the stream is this tool's choice (parity unpinned for the
container; the stream inside is the pinned FEMInputs format)."""
import os
import zipfile

import numpy as np
import pytest

from spgrid_topo_opt_amd import fem_interface as FI
from spgrid_topo_opt_amd import replay as R
from spgrid_topo_opt_amd import workloads as W

from .test_wire_format import interface_fixture


@pytest.fixture()
def dumps(tmp_path):
    p = interface_fixture(8)
    raw = p.serialize()
    f1 = tmp_path / "00001.tcb"
    f1.write_bytes(raw)
    f2 = tmp_path / "00001.tcb.zip"
    with zipfile.ZipFile(f2, "w", compression=zipfile.ZIP_DEFLATED) as z:
        z.writestr("00001.tcb", raw)
    return p, str(f1), str(f2)


def test_raw_and_zip_streams_read_back(dumps):
    p, f1, f2 = dumps
    assert R.read_stream(f1) == p.serialize()
    assert R.read_stream(f2) == p.serialize()
    q = R.read_inputs(f2)
    np.testing.assert_array_equal(q.density, p.density)
    np.testing.assert_array_equal(q.density_base, p.density_base)


def test_overrides_follow_reference_task(dumps):
    p, f1, _ = dumps
    q = R.read_inputs(f1)
    R.apply_overrides(q, ["cg_max_iterations=7", "minimum_stiffness=0.25", "not_a_field_assignment"])
    assert q.krylov.max_iterations == 7
    assert q.extra_doubles["minimum_stiffness"] == 0.25
    with pytest.raises(ValueError, match="Unknown field"):
        R.apply_overrides(q, ["smoothing=3"])


def test_replay_with_fake_solver_matches_direct_call(dumps, tmp_path):
    p, f1, f2 = dumps
    direct = FI.FEMInterface(R.read_inputs(f1))
    direct.preserve_output(len(direct.param.density_base))
    want = direct.run(test=True)
    for f in (f1, f2):
        out, sec = R.replay(f, test=True)
        assert out.success and sec >= 0
        np.testing.assert_array_equal(out.displacements, want.displacements)
    # the CLI writes the serialized outputs
    o = tmp_path / "out.bin"
    assert R.main([f2, "--test", "--out", str(o)]) == 0
    got = FI.FEMOutputs.deserialize(o.read_bytes())
    np.testing.assert_array_equal(got.displacements, want.displacements)


def test_convert_reports_density_range_and_bad_values(dumps, tmp_path, capsys):
    p, f1, _ = dumps
    s = R.density_summary(p)
    nz = p.density[p.density != 0]
    assert s["min_density"] == min(1.0, nz.min()) and s["max_density"] == max(0.0, nz.max())
    assert not s["has_nan"] and not s["has_inf"]
    q = R.read_inputs(f1)
    q.density = q.density.copy()
    q.density[0, 5] = np.nan
    q.density[1, 7] = np.inf
    s = R.density_summary(q)
    assert s["has_nan"] and s["has_inf"]
    txt = tmp_path / "hr.txt"
    assert R.main([f1, "--convert", "--text", str(txt)]) == 0
    assert "Good. No nan nor inf detected." in capsys.readouterr().out
    body = txt.read_text()
    assert "forces: 17" in body and "dirichlet_nodes: 2" in body


@pytest.mark.gpu
def test_replayed_dump_solves_like_the_live_inputs(tmp_path):
    prob = W.box(16, 8, 8, seed=3, mg_levels=3)
    p = FI.FEMInputs.from_problem(prob, tolerance=1e-8, max_iterations=300,
                                  extra_ints={"defect_correction_iter": 10})
    live = FI.FEMInterface(p)
    live.preserve_output(len(p.density_base))
    want = live.run()
    assert want.success
    f = tmp_path / "00002.tcb.zip"
    with zipfile.ZipFile(f, "w") as z:
        z.writestr("00002.tcb", p.serialize())
    got, _ = R.replay(str(f))
    assert got.success
    np.testing.assert_array_equal(got.displacements, want.displacements)
    # an override reaches the solver: a stiffer void changes the solution
    stiff, _ = R.replay(str(f), ["minimum_stiffness=0.5"])
    assert stiff.success and not np.array_equal(stiff.displacements, want.displacements)
    # cg_max_iterations sets krylov.max_iterations, which the solver ignores
    # as the reference's does (SPGrid_CPU_Solver.cpp:326 runs 500)
    same, _ = R.replay(str(f), ["cg_max_iterations=1"])
    np.testing.assert_array_equal(same.displacements, want.displacements)
