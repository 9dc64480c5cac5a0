"""The reference's output path on the host: snapshot files, their index files and
probe records, in the reference's formats, so that its post-processing tools
(``reconstructor``, ``2vtk``, ``script/read_bin.py``) read them unchanged.

Formats (the reference's, kept alike so that its tools read these files):

* snapshot file ``<prefix>_<rank:05d>_<step:010d>`` (src/io.h:11-62, 305-330):
  int64 ``size[3]``, ``gc``, ``var_count``, ``var_dim[var_count]``, then per variable
  an int64 length and its name bytes; then the cell centres ``x[size0]``,
  ``y[size1]``, ``z[size2]`` and every variable over the padded box, vector
  variables interleaved ([L][dim]), all little-endian FP64.  ``inst`` files hold
  U, p, div, q, f; ``tavg`` files the time-averaged U (src/solver.h:835-846);
* ``snapshot.json``: the list of written steps ``{"step", "tavg": "yes" (when a
  tavg file was written), "time"}``, and ``partition.json``: global size, gc and
  the per-rank size/offset, both as nlohmann::json writes them with setw(2)
  (keys sorted, two-space indent, no final newline);
* ``probe<m>.csv``: ``# x y z`` of the probe's cell centre, ``t,u,v,w``, then one
  line per recorded step, values as float printed by a default C++ ostream
  (6 significant digits, src/solver.h:406-465).

The schedule (src/solver.h:1193-1250): after every step, including step 0, the
time average is advanced once t >= tavg start (weight 1/count), probes record
once t >= output start, and a snapshot is written when the next output time
lies within half a step of t.
"""
from __future__ import annotations

import json
import math
import os
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

INST_VARS = (("U", 3), ("p", 1), ("div", 1), ("q", 1), ("f", 3))
TAVG_VARS = (("U", 3),)
PROBE_FLUSH = 10000  # records buffered per probe before an append (src/solver.h:411)


def rank_binary_filename(prefix: str, rank: int, step: int) -> str:
    """src/io.h:326-328"""
    return f"{prefix}_{rank:05d}_{step:010d}"


def write_binary(path: str, size: Sequence[int], gc: int, variables: Sequence[tuple], x, y, z) -> None:
    """variables: (name, dim, array of L*dim doubles) in file order."""
    with open(path, "wb") as fh:
        fh.write(struct.pack("<3q", *size))
        fh.write(struct.pack("<qq", gc, len(variables)))
        fh.write(struct.pack(f"<{len(variables)}q", *[d for _, d, _ in variables]))
        for name, _, _ in variables:
            b = name.encode()
            fh.write(struct.pack("<q", len(b)))
            fh.write(b)
        for c in (x, y, z):
            fh.write(np.ascontiguousarray(c, dtype="<f8").tobytes())
        L = int(size[0]) * int(size[1]) * int(size[2])
        for name, dim, a in variables:
            a = np.ascontiguousarray(a, dtype="<f8").reshape(-1)
            if a.size != L * dim:
                raise ValueError(f"{name}: {a.size} values, expected {L * dim}")
            fh.write(a.tobytes())


@dataclass
class Snapshot:
    size: tuple
    gc: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    var: Dict[str, np.ndarray] = field(default_factory=dict)
    dims: Dict[str, int] = field(default_factory=dict)


def read_binary(path: str) -> Snapshot:
    with open(path, "rb") as fh:
        data = fh.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from(fmt, data, off)
        off += struct.calcsize(fmt)
        return v

    size = take("<3q")
    gc, nv = take("<qq")
    dims = take(f"<{nv}q")
    names = []
    for _ in range(nv):
        (n,) = take("<q")
        names.append(data[off:off + n].decode())
        off += n
    arrs = []
    for cnt in (size[0], size[1], size[2], *[size[0] * size[1] * size[2] * d for d in dims]):
        arrs.append(np.frombuffer(data, dtype="<f8", count=cnt, offset=off).copy())
        off += 8 * cnt
    if off != len(data):
        raise ValueError(f"{path}: {len(data) - off} trailing bytes")
    s = Snapshot(tuple(size), gc, arrs[0], arrs[1], arrs[2])
    for name, d, a in zip(names, dims, arrs[3:]):
        s.var[name] = a.reshape(-1, d) if d > 1 else a
        s.dims[name] = d
    return s


# ----------------------------------------------------------------------------- JSON
def json_number(v) -> str:
    """A number as nlohmann::json dumps it: integers as such, doubles in the
    shortest form that reads back exactly, fixed notation while the decimal point
    falls within -3..15 digits of the first digit (".0" on integral values),
    otherwise d.ddde[+-]XX."""
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    v = float(v)
    if math.isnan(v) or math.isinf(v):
        return "null"
    if v == 0:
        return "-0.0" if math.copysign(1.0, v) < 0 else "0.0"
    sign, a = ("-" if v < 0 else ""), abs(v)
    for p in range(1, 18):  # fewest significant digits that read back exactly
        t = f"{a:.{p - 1}e}"
        if float(t) == a:
            break
    mant, e = t.split("e")
    digits = mant.replace(".", "").rstrip("0") or "0"
    k, n = len(digits), int(e) + 1  # value = 0.digits x 10^n
    if k <= n <= 15:
        s = digits + "0" * (n - k) + ".0"
    elif 0 < n <= 15:
        s = digits[:n] + "." + digits[n:]
    elif -4 < n <= 0:
        s = "0." + "0" * (-n) + digits
    else:
        x = n - 1
        s = digits[0] + ("." + digits[1:] if k > 1 else "") + "e" + ("-" if x < 0 else "+") + f"{abs(x):02d}"
    return sign + s


def json_dump(obj, indent: int = 2, level: int = 0) -> str:
    """nlohmann::json's dump(indent) layout: object keys sorted, every array and
    object element on its own line, "key": value, empty containers as [] / {}."""
    pad, inner = " " * (indent * level), " " * (indent * (level + 1))
    if obj is None:
        return "null"
    if isinstance(obj, str):
        return json.dumps(obj)
    if isinstance(obj, dict):
        if not obj:
            return "{}"
        items = [f"{inner}{json.dumps(k)}: {json_dump(obj[k], indent, level + 1)}" for k in sorted(obj)]
        return "{\n" + ",\n".join(items) + "\n" + pad + "}"
    if isinstance(obj, (list, tuple)):
        if not obj:
            return "[]"
        items = [inner + json_dump(v, indent, level + 1) for v in obj]
        return "[\n" + ",\n".join(items) + "\n" + pad + "]"
    return json_number(obj)


def partition_json(size: Sequence[int], gc: int) -> str:
    """partition.json of a one-rank run (src/solver.h:982-999, src/solver.h:27-50)."""
    size = [int(s) for s in size]
    return json_dump({"size": size, "gc": int(gc),
                      "partition": [{"size": size, "offset": [0, 0, 0], "rank": 0}]})


def write_ap_info(path: str, ap, cd, cl, atk) -> None:
    """ap_info.log (src/solver.h:950-977): per actuator point r, dr, chord, twist and
    its Cd / Cl rows, then the attack angles, every value as C's %lf."""
    ap, cd, cl = np.asarray(ap), np.asarray(cd).reshape(len(ap), -1), np.asarray(cl).reshape(len(ap), -1)
    out = [f"ap count = {len(ap)}, atk acount = {len(atk)}\n"]
    for i, a in enumerate(ap):
        out.append(f"ap {i}\n\tr = {a[3]:f}\n\tdr = {a[7]:f}\n\tchord = {a[4]:f}\n\ttwist = {a[5]:f}\n\tcd = \n")
        out.extend(f"\t\t{v:f}\n" for v in cd[i])
        out.append("\tcl = \n")
        out.extend(f"\t\t{v:f}\n" for v in cl[i])
    out.append("atk\n")
    out.extend(f"\t{v:f}\n" for v in atk)
    with open(path, "w") as fh:
        fh.write("".join(out))


# ----------------------------------------------------------------------------- probes
def cxx_general(v: float) -> str:
    """A value as a default-formatted C++ ostream prints it (%g, 6 digits)."""
    return "%g" % v


def find_nearest_index(c: np.ndarray, v: float) -> int:
    """src/util.h:92-106 over c[0:len]: the closer of the two bracketing centres
    (ties to the upper one), clamped to the ends."""
    n = len(c)
    if v < c[0]:
        return 0
    if v >= c[n - 1]:
        return n - 1
    for i in range(n - 1):
        if c[i] <= v < c[i + 1]:
            return i if v - c[i] < c[i + 1] - v else i + 1
    return -1


def probe_cell(x, y, z, gc: int, pos) -> Optional[tuple]:
    """Nearest cell to pos among the interior centres and the first ghost layer on
    each side; None unless that cell is interior (src/solver.h:656-700)."""
    size = (len(x), len(y), len(z))
    idx = []
    for c, v, n in zip((x, y, z), pos, size):
        idx.append(find_nearest_index(np.asarray(c[gc - 1:gc - 1 + n - 2 * gc + 2]), v) + gc - 1)
    if all(gc <= i < n - gc for i, n in zip(idx, size)):
        return tuple(idx)
    return None


class Probe:
    """One probe file: header at creation, records appended in blocks of
    PROBE_FLUSH and at close (src/solver.h:406-465)."""

    def __init__(self, path: str, cell: tuple, xyz: tuple):
        self.path, self.cell = path, cell
        self.recs: List[tuple] = []
        with open(path, "w") as fh:
            fh.write("# " + " ".join(cxx_general(float(c)) for c in xyz) + "\n")
            fh.write("t,u,v,w\n")

    def add(self, t: float, u) -> None:
        f32 = np.float32
        self.recs.append((float(f32(t)), *[float(f32(c)) for c in u]))
        if len(self.recs) == PROBE_FLUSH:
            self.flush()

    def flush(self) -> None:
        if self.recs:
            with open(self.path, "a") as fh:
                for r in self.recs:
                    fh.write(",".join(cxx_general(v) for v in r) + "\n")
            self.recs = []


# ----------------------------------------------------------------------------- schedule
class OutputSchedule:
    """Runtime's output bookkeeping (src/solver.h:176-205, 1193-1250)."""

    def __init__(self, dt: float, start: float, interval: float, tavg_start: float):
        self.dt, self.start, self.interval, self.tavg_start = dt, start, interval, tavg_start
        self.next_output = start
        self.tavg_count = 0

    def tavg_weight(self, t: float) -> Optional[float]:
        if t >= self.tavg_start:
            self.tavg_count += 1
            return 1.0 / self.tavg_count
        return None

    def probes_due(self, t: float) -> bool:
        return t >= self.start

    def snapshot_due(self, t: float) -> bool:
        if t - 0.5 * self.dt < self.next_output <= t + 0.5 * self.dt:
            self.next_output += self.interval
            return True
        return False


class OutputWriter:
    """The output side of Solver (src/solver.h:569-598, 656-700, 835-846,
    1193-1250) on a dalm context: call after_step(step) after the start-up pass
    (step 0) and after every step, close() at the end."""

    def __init__(self, ctx, case, out_dir: str, alm_tables=None):
        """alm_tables: (ap, cd, cl, atk) of build_ap_props, written to ap_info.log."""
        self.ctx, self.case, self.dir = ctx, case, out_dir
        os.makedirs(out_dir, exist_ok=True)
        m = case.mesh
        self.size, self.gc = (len(m.x), len(m.y), len(m.z)), m.gc
        self.sched = OutputSchedule(case.dt, case.output_start, case.output_interval, case.tavg_start)
        self.snapshots: list = []
        with open(os.path.join(out_dir, "partition.json"), "w") as fh:
            fh.write(partition_json(self.size, self.gc))
        if alm_tables is not None:
            write_ap_info(os.path.join(out_dir, "ap_info.log"), *alm_tables)
        self.probes: List[Probe] = []
        for n, pos in enumerate(case.probes):
            cell = probe_cell(m.x, m.y, m.z, m.gc, pos)
            if cell is not None:
                xyz = (m.x[cell[0]], m.y[cell[1]], m.z[cell[2]])
                self.probes.append(Probe(os.path.join(out_dir, f"probe{n}.csv"), cell, xyz))
        n1, n2 = self.size[1], self.size[2]
        self.cells = np.array([(i * n1 + j) * n2 + k for i, j, k in (p.cell for p in self.probes)],
                              dtype=np.int64)

    def after_step(self, step: int) -> None:
        t = self.case.dt * step
        w = self.sched.tavg_weight(t)
        if w is not None:
            self.ctx.time_average(w)
        if self.sched.probes_due(t) and self.probes:
            U = self.ctx.sample_U(self.cells)
            for p, u in zip(self.probes, U):
                p.add(t, u)
        if self.sched.snapshot_due(t):
            self.ctx.calc_q()
            self.ctx.calc_divergence()
            m = self.case.mesh
            var = [(n, d, self.ctx.download(n)) for n, d in INST_VARS]
            write_binary(os.path.join(self.dir, rank_binary_filename("inst", 0, step)), self.size, self.gc, var,
                         m.x, m.y, m.z)
            entry = {"step": int(step), "time": t}
            if t >= self.sched.tavg_start:
                write_binary(os.path.join(self.dir, rank_binary_filename("tavg", 0, step)), self.size, self.gc,
                             [("U", 3, self.ctx.download("Utavg"))], m.x, m.y, m.z)
                entry["tavg"] = "yes"
            self.snapshots.append(entry)
            self.write_index()

    def write_index(self) -> None:
        with open(os.path.join(self.dir, "snapshot.json"), "w") as fh:
            fh.write(json_dump(self.snapshots) if self.snapshots else "null")

    def close(self) -> None:
        self.write_index()
        for p in self.probes:
            p.flush()
