"""The output path on the CPU: the oracle's Q-criterion and time average against
the reference build and closed forms, and the files of 3dalm_amd/output.py
(snapshot format, snapshot.json / partition.json layout, probe records,
ap_info.log, the output schedule) against the reference's own output folder for
the same run (tests/golden/output_small.npz)."""
import importlib
import math

import numpy as np
import pytest

from helpers import interior_mask, rand_field, small_mesh
from outputs import compare_output, golden_files, oracle_output

cases = importlib.import_module("3dalm_amd.cases")
output = importlib.import_module("3dalm_amd.output")
po = importlib.import_module("pyoracle")


@pytest.fixture(scope="module")
def K():
    return po.Kernels("oracle")


def test_q_closed_forms(K):
    """Central differences are exact for linear fields on any grid: solid-body
    rotation U = (-w y, w x, 0) has q = w^2, pure strain (a x, -a y, 0) q = -a^2,
    the shear (a y, 0, 0) q = 0."""
    m = small_mesh(9, 8, 7, 1.4)
    g = po.make_grid(m)
    X, Y, Z = np.meshgrid(m.x, m.y, m.z, indexing="ij")
    X, Y, Z = X.reshape(-1), Y.reshape(-1), Z.reshape(-1)
    inner = interior_mask(m)
    w, a = 1.7, 0.6
    for U, want in ((np.stack([-w * Y, w * X, 0 * Z], 1), w * w),
                    (np.stack([a * X, -a * Y, 0 * Z], 1), -a * a),
                    (np.stack([a * Y, 0 * Y, 0 * Z], 1), 0.0)):
        q = np.full(m.L, 7.0)
        K.calc_q(g, np.ascontiguousarray(U), q)
        assert np.allclose(q[inner], want, rtol=1e-12, atol=1e-12)
        assert np.all(q[~inner] == 7.0)  # only interior cells are written


def test_time_average_closed_forms(K):
    """With weights 1/n the running average is the arithmetic mean of the fields;
    the average of a constant field stays that constant (to rounding)."""
    m = small_mesh(5, 4, 6)
    g = po.make_grid(m)
    fields = [rand_field(m, s, 3) for s in range(5)]
    Ut = rand_field(m, 99, 3)  # overwritten by the first (weight 1) update
    for n, U in enumerate(fields, 1):
        K.time_average(g, Ut, U, 1.0 / n)
    assert np.allclose(Ut, np.mean(fields, axis=0), rtol=1e-13, atol=1e-14)
    c = np.full((m.L, 3), 0.3)
    Ut = c.copy()
    for n in range(2, 9):
        K.time_average(g, Ut, c, 1.0 / n)
    assert np.allclose(Ut, c, rtol=4e-16 * 8, atol=0)


@pytest.mark.skipif(not po.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("grid", [(7, 6, 5, 1.3), (5, 8, 66, 1.6)])
def test_q_and_average_match_reference(K, grid):
    R = po.Kernels("ref")
    m = small_mesh(*grid)
    g = po.make_grid(m)
    U = rand_field(m, 3, 3, 0.4, 1.0)
    qa, qb = np.zeros(m.L), np.zeros(m.L)
    K.calc_q(g, U, qa)
    R.calc_q(g, U, qb)
    assert np.array_equal(qa, qb)
    ta, tb = np.zeros((m.L, 3)), np.zeros((m.L, 3))
    for n in range(1, 4):
        Un = rand_field(m, 10 + n, 3)
        K.time_average(g, ta, Un, 1.0 / n)
        R.time_average(g, tb, Un, 1.0 / n)
        assert np.array_equal(ta, tb)


def test_json_numbers():
    cases_ = {0.0: "0.0", 0.01: "0.01", 0.015: "0.015", 2.0: "2.0", 1e-4: "0.0001", 1e-5: "1e-05",
              123.456: "123.456", 1e15: "1e+15", 1e14: "100000000000000.0", 0.1 + 0.2: "0.30000000000000004",
              -2.5e-7: "-2.5e-07", 5e-324: "5e-324"}
    for v, s in cases_.items():
        assert output.json_number(v) == s, (v, output.json_number(v), s)
        if math.isfinite(v):
            assert float(output.json_number(v)) == v
    assert output.json_number(3) == "3"
    assert output.json_dump([]) == "[]" and output.json_dump(None) == "null"


def test_binary_round_trip(tmp_path):
    m = small_mesh(4, 3, 5)
    U, p = rand_field(m, 1, 3), rand_field(m, 2, 1)
    path = str(tmp_path / output.rank_binary_filename("inst", 0, 12))
    assert path.endswith("inst_00000_0000000012")
    output.write_binary(path, m.size, m.gc, [("U", 3, U), ("p", 1, p)], m.x, m.y, m.z)
    s = output.read_binary(path)
    assert s.size == tuple(m.size) and s.gc == m.gc and s.dims == {"U": 3, "p": 1}
    assert np.array_equal(s.var["U"], U) and np.array_equal(s.var["p"], p)
    assert np.array_equal(s.x, m.x) and np.array_equal(s.z, m.z)


def test_reference_files_read_back():
    """The reference's snapshot files parse with this reader: header, mesh and
    variable set as src/solver.h:835-846 defines them."""
    files = golden_files()
    case = cases.case_output_small()
    m = case.mesh
    assert sorted(files) == ["ap_info.log", "inst_00000_0000000000", "inst_00000_0000000002",
                             "inst_00000_0000000004", "partition.json", "probe0.csv", "probe1.csv",
                             "snapshot.json", "tavg_00000_0000000002", "tavg_00000_0000000004"]
    import os
    import tempfile
    with tempfile.TemporaryDirectory() as t:
        for name in ("inst_00000_0000000004", "tavg_00000_0000000004"):
            path = os.path.join(t, name)
            open(path, "wb").write(files[name])
            s = output.read_binary(path)
            assert s.size == tuple(m.size) and s.gc == m.gc
            assert np.array_equal(s.x, m.x) and np.array_equal(s.y, m.y) and np.array_equal(s.z, m.z)
            want = output.INST_VARS if name.startswith("inst") else output.TAVG_VARS
            assert [(n, s.dims[n]) for n in s.var] == list(want)


def test_oracle_output_matches_reference(tmp_path):
    """The oracle's run of the same case, written through 3dalm_amd/output.py,
    reproduces the reference's output folder: every text file byte for byte and
    every snapshot variable bit for bit (both sides use glibc's libm)."""
    case = cases.case_output_small()
    oracle_output(case, str(tmp_path))
    worst = compare_output(str(tmp_path), golden_files(), rtol=0.0, atol=0.0)
    assert set(worst) == {"inst.U", "inst.p", "inst.div", "inst.q", "inst.f", "tavg.U"}
