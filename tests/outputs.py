"""Checks of the output path shared by the CPU and GPU tests: the reference's own
output files for cases.case_output_small() (tests/golden/output_small.npz, written
by tests/golden/make_golden.py from the reference build), an oracle run that
writes the same files through 3dalm_amd/output.py, and a file-by-file comparison."""
import importlib
import os

import numpy as np

cases = importlib.import_module("3dalm_amd.cases")
output = importlib.import_module("3dalm_amd.output")
po = importlib.import_module("pyoracle")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "output_small.npz")


def golden_files():
    """file name -> bytes of the reference's output folder."""
    d = np.load(GOLDEN)
    case = cases.case_output_small()
    want = np.array([*case.mesh.interior, case.steps, case.dt, case.output_start, case.output_interval,
                     case.tavg_start])
    assert np.array_equal(d["case"], want), "output_small.npz was written for another case"
    return {k[5:].replace("__", "."): d[k].tobytes() for k in d.files if k.startswith("file_")}


def oracle_output(case, out_dir):
    """The oracle's whole steps with the output path of main_loop_once
    (src/solver.h:1193-1250) restated over its fields: q and the time average
    from the oracle's calc_q / time average, files through 3dalm_amd/output.py."""
    os.makedirs(out_dir, exist_ok=True)
    K = po.Kernels("oracle")
    o = po.OracleSolver(case)
    m = case.mesh
    size, gc = tuple(m.size), m.gc
    g = po.make_grid(m)
    q = np.zeros(m.L)
    Ut = np.zeros((m.L, 3))
    with open(os.path.join(out_dir, "partition.json"), "w") as fh:
        fh.write(output.partition_json(size, gc))
    if case.turbine is not None and case.sites:
        output.write_ap_info(os.path.join(out_dir, "ap_info.log"), o.ap, o.cd, o.cl, o.atk)
    probes = []
    for n, pos in enumerate(case.probes):
        cell = output.probe_cell(m.x, m.y, m.z, gc, pos)
        if cell is not None:
            probes.append(output.Probe(os.path.join(out_dir, f"probe{n}.csv"), cell,
                                       (m.x[cell[0]], m.y[cell[1]], m.z[cell[2]])))
    sched = output.OutputSchedule(case.dt, case.output_start, case.output_interval, case.tavg_start)
    snaps = []
    for step in range(case.steps + 1):
        if step:
            o.step()
        t = case.dt * step
        w = sched.tavg_weight(t)
        if w is not None:
            K.time_average(g, Ut, o.U, w)
        if sched.probes_due(t):
            for p in probes:
                i, j, k = p.cell
                p.add(t, o.U[(i * size[1] + j) * size[2] + k])
        if sched.snapshot_due(t):
            K.calc_q(g, o.U, q)
            output.write_binary(os.path.join(out_dir, output.rank_binary_filename("inst", 0, step)), size, gc,
                                [("U", 3, o.U), ("p", 1, o.p), ("div", 1, o.div), ("q", 1, q), ("f", 3, o.f)],
                                m.x, m.y, m.z)
            e = {"step": step, "time": t}
            if t >= sched.tavg_start:
                output.write_binary(os.path.join(out_dir, output.rank_binary_filename("tavg", 0, step)), size, gc,
                                    [("U", 3, Ut)], m.x, m.y, m.z)
                e["tavg"] = "yes"
            snaps.append(e)
    with open(os.path.join(out_dir, "snapshot.json"), "w") as fh:
        fh.write(output.json_dump(snaps))
    for p in probes:
        p.flush()


def compare_output(got_dir, expected, rtol, atol, exact_text=True, div_atol=0.0):
    """Every file of `expected` (name -> bytes) must exist in got_dir and agree:
    text files byte for byte (probe files, with exact_text False, value by value
    within 1e-5 relative), snapshot files with identical headers and mesh and
    every variable within rtol (relative to the variable's largest magnitude) +
    atol.  The divergence is what is left after the projection cancels O(1) face
    fluxes, so its bar is absolute: div_atol + atol.  Returns the largest
    relative difference per snapshot variable."""
    got_names = set(os.listdir(got_dir))
    assert set(expected) <= got_names, sorted(set(expected) - got_names)
    worst = {}
    for name, want in sorted(expected.items()):
        got = open(os.path.join(got_dir, name), "rb").read()
        if name.startswith(("inst_", "tavg_")):
            path_w = os.path.join(got_dir, "_expected_" + name)
            with open(path_w, "wb") as fh:
                fh.write(want)
            a, b = output.read_binary(os.path.join(got_dir, name)), output.read_binary(path_w)
            os.remove(path_w)
            assert (a.size, a.gc, a.dims) == (b.size, b.gc, b.dims), name
            assert list(a.var) == list(b.var), name
            for c in ("x", "y", "z"):
                assert np.array_equal(getattr(a, c), getattr(b, c)), (name, c)
            for v in b.var:
                s = np.max(np.abs(b.var[v]))
                d = np.max(np.abs(a.var[v] - b.var[v]))
                rel = d / s if s > 0 else d
                worst[f"{name[:4]}.{v}"] = max(worst.get(f"{name[:4]}.{v}", 0.0), rel)
                bar = div_atol + atol if v == "div" else rtol * s + atol
                assert d <= bar, (name, v, d, s)
        elif name.endswith(".csv") and not exact_text:
            ga, wa = got.decode().splitlines(), want.decode().splitlines()
            assert ga[:2] == wa[:2] and len(ga) == len(wa), name
            for lg, lw in zip(ga[2:], wa[2:]):
                vg = np.array([float(v) for v in lg.split(",")])
                vw = np.array([float(v) for v in lw.split(",")])
                assert np.allclose(vg, vw, rtol=1e-5, atol=1e-7), (name, lg, lw)
        else:
            assert got == want, name
    return worst


def expected_report(case, case_dir, max_diag, ngpu):
    """The start-up report that Solver::initialize prints on one rank
    (src/solver.h:855-1010) for a case written by cases.write_case_dir, as lines.
    The two %e lines (initial div(U), initial max cfl) end in "<value>"."""
    m = case.mesh
    n = tuple(m.size)
    out_dir = os.path.join(case_dir, "output")
    max_step = int(case.dt * (case.steps + 0.5) / case.dt)
    L = ["SETUP INFO", f"\tcase directory = {case_dir}", "DEVICE INFO", f"\tnumber of GPUs = {ngpu}",
         "CHARACTERISTIC SCALE INFO", "\tU = 1.000000", "\tL = 1.000000", "\tt = 1.000000", "\tp = 1.000000",
         "MESH INFO", f"\tpath = {os.path.join(case_dir, 'mesh.json')}", "\tglobal size = (%d %d %d)" % n,
         f"\tguide cell = {m.gc}", "RUNTIME INFO", "\tdt = %f" % case.dt, f"\tmax step = {max_step}",
         "OUTPUT INFO", f"\tdirectory = {out_dir}", "\toutput start time = %f" % case.output_start,
         "\toutput interval time = %f" % case.output_interval, "\ttime avg start time = %f" % case.tavg_start,
         "\toutput vars = U p div q f", "\ttime avg vars = U", "TURBULENCE GENERATING GRID INFO"]
    if case.tgg is not None:
        t = case.tgg
        L += ["\tthickness = %f" % t["thick"], "\tspacing = %f" % t["mesh"],
              "\tplacement = (%f %f %f)" % tuple(t["placement"])]
    else:
        L += ["\tno turbulence generating grid"]
    L += ["CFD INFO", "\tRe = %f" % case.Re, "\tCs = %f" % case.Cs, "\tinitial div(U) = <value>",
          "\tinitial max cfl = <value>"]
    if case.floor_start is not None:
        L += ["\tretracted floor boundary applied", "\tfloor boundary start from x = %f" % case.floor_start]
    L += ["EQ INFO", f"\tmax iteration = {case.max_it}", "\ttolerance = %f" % case.tol, "\tmax A diag = %f" % max_diag,
          f"\tmethod = {'BiCG' if case.method == 'BiCG' else 'SOR'}"]
    if case.method == "BiCG":
        L += ["PRECONDITIONER INFO", f"\tmethod = {case.pc}", f"\tmax iteration = {case.pc_it}"]
        if case.pc == "SOR":
            L += ["\trelaxation rate = %f" % case.pc_omega]
    else:
        L += ["\trelaxation rate = %f" % case.omega]
    if case.turbine is not None and case.sites:
        L += ["ALM INFO", f"\tpoints per blade = {case.ap_per_blade}", "\tprojection width = %f" % case.width,
              "WINDTURBINE INFO"]
        wt = cases.turbine_records(case.turbine, case.sites)
        names = {0: "roll", 1: "pitch", 2: "yaw"}
        for w, row in enumerate(wt):
            L += [f"\twt {w}", f"\t\tblades = {case.turbine.blades}", "\t\tbase = (%f %f %f)" % tuple(row[0:3]),
                  "\t\trotation speed = %f" % row[13], "\t\trotation center = (%f %f %f)" % tuple(row[3:6]),
                  f"\t\tangle type = {names.get(int(row[6]), 'none')}",
                  "\t\tangle formula = %f sin(%f t + %f) + %f" % tuple(row[9:13])]
        L += ["ACTUATOR POINT INFO", f"\toutput to {os.path.join(out_dir, 'ap_info.log')}"]
    else:
        L += ["WINDTURBINE INFO", "\tno windturbine"]
    L += ["PROC INFO 0/1", "\tgsize = (%d %d %d)" % n, "\tsize = (%d %d %d)" % n, "\toffset = (0 0 0)",
          "\tGPU id = 0", "\tprobes = "]
    for pos in case.probes:
        c = output.probe_cell(m.x, m.y, m.z, m.gc, pos)
        if c is not None:
            L.append("\t\tposition = (%f %f %f), local index = (%d %d %d)" % (m.x[c[0]], m.y[c[1]], m.z[c[2]], *c))
    return L


def check_report(stdout, want):
    """stdout's start-up report equals `want` line for line (values of the
    "<value>" lines: any %e number)."""
    lines = stdout.splitlines()
    first = lines.index("SETUP INFO")
    got = lines[first:first + len(want)]
    for g, w in zip(got, want):
        if w.endswith("<value>"):
            head = w[:-len("<value>")]
            assert g.startswith(head), (g, w)
            float(g[len(head):])
        else:
            assert g == w, (g, w)
    assert len(got) == len(want)
    assert lines[0].startswith("START ") and lines[-2].startswith("END "), (lines[0], lines[-2:])
