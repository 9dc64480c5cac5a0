"""The C++ case driver (3dalm_amd/lib/dalm_solver): reads the reference's case files
and prints the reference's per-step line.  The GPU test runs a small case and
compares every printed line with the CPU oracle; the CPU tests cover argument and
case-file errors that end before any device call."""
import importlib
import json
import os
import subprocess

import numpy as np
import pytest

cases = importlib.import_module("3dalm_amd.cases")
build = importlib.import_module("3dalm_amd.build")
po = importlib.import_module("pyoracle")
hip = importlib.import_module("3dalm_amd.hip")
from outputs import check_report, expected_report  # noqa: E402
output = importlib.import_module("3dalm_amd.output")


def run(args, timeout=300):
    return subprocess.run([build.DRIVER, *args], capture_output=True, text=True, timeout=timeout)


def test_bad_turbulence_grid(tmp_path):
    """A grid spacing of zero is refused while the case is read, before any
    device call."""
    case = cases.case_tgg_small(steps=1)
    case.tgg = dict(case.tgg, mesh=0.0)
    cases.write_case_dir(case, str(tmp_path))
    r = run([str(tmp_path)])
    assert r.returncode == 1 and "turbulence grid: need mesh > 0" in r.stderr, r.stderr


def test_missing_case_file(tmp_path):
    r = run([str(tmp_path / "nowhere")])
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_usage():
    r = run([])
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("builder", [lambda: cases.case_c1(12, 10, 10, steps=4),
                                     lambda: cases.case_c2(24, 12, 12, steps=3, ap_per_blade=6),
                                     lambda: cases.case_tgg_small(steps=3)])
def test_driver_matches_oracle(tmp_path, builder):
    case = builder()
    print(f"[driver case {case.name} {case.mesh.interior}]", flush=True)
    cases.write_case_dir(case, str(tmp_path))
    r = run([str(tmp_path), "--perturb", str(case.noise * case.uin[0]), "--seed", str(case.seed), "--power"])
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.splitlines() if l and l.split()[0].isdigit()]
    power = [float(l.split()[-1]) for l in r.stdout.splitlines() if l.startswith("# wt")]
    o = po.OracleSolver(case)
    check_report(r.stdout, expected_report(case, str(tmp_path), o.stats()["max_diag"], hip.Context.device_count()))
    assert len(lines) == case.steps
    for n, l in enumerate(lines):
        s = o.step()
        assert int(l[0]) == s["step"]
        # printed with %e (7 significant digits).  SOR is bit-exact; BiCGStab's
        # Krylov path depends on the summation order of its dot products, so its
        # iteration count may move by a few and its final residual is only bounded
        # by the tolerance (the solution agrees to the tolerance)
        if case.method == "SOR":
            assert int(l[2]) == s["it"]
            checks = ((1, None, 1e-6), (3, "err", 1e-6), (4, "avg_div", 1e-6), (5, "max_cfl", 1e-6))
        else:
            assert abs(int(l[2]) - s["it"]) <= max(4, s["it"] // 5), (n, l[2], s["it"])
            assert float(l[3]) <= case.tol
            checks = ((1, None, 1e-6), (5, "max_cfl", 1e-6))
        for q, key, tol in checks:
            ref = s["step"] * case.dt if key is None else s[key]
            assert abs(float(l[q]) - ref) <= tol * abs(ref) + 1e-300, (n, key, l[q], ref)
        if case.tgg is not None and n == 0:  # the solid fraction written at start-up
            sol = output.read_binary(str(tmp_path / "output" / "solid_00000_0000000000"))
            assert list(sol.var) == ["solid"] and np.array_equal(sol.var["solid"], o.field("solid"))
        if power:
            rp = 1e-10 if case.method == "SOR" else 1e-7
            assert abs(power[n] - case.sites[0].rot_speed * s["torque"][0]) <= rp * abs(power[n])
