"""Synthetic case builders for the 3dalm hot path (host-side, numpy only).

They produce the inputs the reference reads from its case directory -- ``mesh.json``
(``src/io.h:98-164``), ``windturbine.json`` and the ``windturbine array`` entry of
``setup.json`` (``src/alm.h:35-206``, ``src/solver.h:702-708``) -- for the workloads
this project measures, plus the seeded initial perturbation the reference lacks.

Choices that the reference leaves open are fixed here once, so that every consumer
(the HIP library, the C++ driver, the CPU oracle and the reference build used for
golden vectors) sees bit-identical inputs:

* mesh nodes: x uniform; y, z two-sided tanh,
  ``s_j = L0 + (L1-L0) * (1 + tanh(beta*(2j/n-1))/tanh(beta)) / 2``;
  cell centres are node midpoints, ghost centres follow the extrapolation
  ``x_g = 3 x_{g+1} - 3 x_{g+2} + x_{g+3}`` of ``script/convertmesh.py:28-32``;
* widths: the reference rule ``src/io.h:115-125`` (central, one-sided 2nd order at
  the ends);
* noise: counter-based splitmix64 of (seed, component, interior index), top 53 bits
  mapped to [-1, 1);
* Fourier-mode turbulence (the at-scale workloads): a sum of seeded random
  divergence-free modes, ``fourier_mode_velocity``;
* polar: every 1 degree on [-180, 180], ``Cl = clip(2 pi alpha, +-1.2)``,
  ``Cd = 0.01 + 0.05 alpha^2`` (alpha in radians).
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

GC = 2


# ----------------------------------------------------------------------------- mesh
def uniform_nodes(l0: float, l1: float, n: int) -> np.ndarray:
    return l0 + (l1 - l0) * (np.arange(n + 1, dtype=np.float64) / n)


def tanh_nodes(l0: float, l1: float, n: int, beta: float) -> np.ndarray:
    s = np.arange(n + 1, dtype=np.float64)
    return l0 + (l1 - l0) * 0.5 * (1.0 + np.tanh(beta * (2.0 * s / n - 1.0)) / math.tanh(beta))


def centres_with_ghosts(nodes: np.ndarray, gc: int = GC) -> np.ndarray:
    """Interior centres at node midpoints, ghosts by cubic extrapolation
    (script/convertmesh.py:15-34)."""
    nint = len(nodes) - 1
    c = np.zeros(nint + 2 * gc)
    c[gc:gc + nint] = 0.5 * (nodes[:-1] + nodes[1:])
    for i in range(gc - 1, -1, -1):
        c[i] = 3 * c[i + 1] - 3 * c[i + 2] + c[i + 3]
    for i in range(gc + nint, nint + 2 * gc):
        c[i] = 3 * c[i - 1] - 3 * c[i - 2] + c[i - 3]
    return c


def widths(c: np.ndarray) -> np.ndarray:
    """Cell widths from centres, src/io.h:115-125."""
    d = np.empty_like(c)
    d[1:-1] = 0.5 * (c[2:] - c[:-2])
    d[0] = -1.5 * c[0] + 2.0 * c[1] - 0.5 * c[2]
    d[-1] = 1.5 * c[-1] - 2.0 * c[-2] + 0.5 * c[-3]
    return d


@dataclass
class Mesh:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    gc: int = GC

    def __post_init__(self):
        self.x = np.ascontiguousarray(self.x, dtype=np.float64)
        self.y = np.ascontiguousarray(self.y, dtype=np.float64)
        self.z = np.ascontiguousarray(self.z, dtype=np.float64)
        self.dx, self.dy, self.dz = widths(self.x), widths(self.y), widths(self.z)

    @property
    def size(self):
        return (len(self.x), len(self.y), len(self.z))

    @property
    def interior(self):
        return tuple(n - 2 * self.gc for n in self.size)

    @property
    def L(self) -> int:
        a, b, c = self.size
        return a * b * c

    @property
    def N(self) -> int:
        a, b, c = self.interior
        return a * b * c

    def to_json(self) -> str:
        """mesh.json text (src/io.h:105-151); repr keeps every double exact."""
        def arr(v):
            return "[" + ", ".join(repr(float(t)) for t in v) + "]"
        return ('{"guide cell": %d, "x": %s, "y": %s, "z": %s}'
                % (self.gc, arr(self.x), arr(self.y), arr(self.z)))


def make_mesh(nx, ny, nz, box, beta_y=None, beta_z=None, gc=GC) -> Mesh:
    """box = ((x0,x1),(y0,y1),(z0,z1)); beta None -> uniform."""
    (x0, x1), (y0, y1), (z0, z1) = box
    xn = uniform_nodes(x0, x1, nx)
    yn = tanh_nodes(y0, y1, ny, beta_y) if beta_y else uniform_nodes(y0, y1, ny)
    zn = tanh_nodes(z0, z1, nz, beta_z) if beta_z else uniform_nodes(z0, z1, nz)
    return Mesh(centres_with_ghosts(xn, gc), centres_with_ghosts(yn, gc),
                centres_with_ghosts(zn, gc), gc)


# ----------------------------------------------------------------------------- noise
_M64 = (1 << 64) - 1


def splitmix64(z: np.ndarray) -> np.ndarray:
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_M64)
    return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """count values in [-1, 1), counter-based (seed, stream, index)."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0xD1342543DE82EF95)
               + np.uint64(stream) * np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
        z = splitmix64(np.arange(count, dtype=np.uint64) ^ key)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -52) - 1.0


def initial_velocity(mesh: Mesh, uin, amp: float, seed: int) -> np.ndarray:
    """[L][3] velocity: Uin everywhere, interior perturbed by amp*xi per component
    (xi uniform in [-1,1), interior cells in (i,j,k) order)."""
    U = np.empty((mesh.L, 3))
    U[:] = np.asarray(uin, dtype=np.float64)
    if amp:
        g = mesh.gc
        nx, ny, nz = mesh.size
        idx = np.arange(mesh.L).reshape(nx, ny, nz)[g:nx - g, g:ny - g, g:nz - g].ravel()
        for m in range(3):
            U[idx, m] = U[idx, m] + amp * uniform_pm1(seed, len(idx), stream=m)
    return U


def fourier_mode_velocity(mesh: Mesh, uin, intensity: float, seed: int, modes: int = 200) -> np.ndarray:
    """[L][3] velocity of fourier_modes (Uin plus the modes on the interior)."""
    return fourier_modes(mesh, uin, intensity, seed, modes)[0]


def fourier_modes(mesh: Mesh, uin, intensity: float, seed: int, modes: int = 200):
    """([L][3] velocity, [modes][7] mode table): Uin everywhere, the interior plus a
    seeded synthetic isotropic turbulence of `modes` random Fourier modes,

        u'(x) = sum_n a_n sigma_n cos(k_n . x + psi_n),   sigma_n . k_n = 0,

    so every mode is divergence-free.  Directions of k_n are uniform on the sphere,
    |k_n| log-uniform between 2 pi / (smallest box extent) and 2 pi / (8 x the widest
    interior cell), so every mode spans at least 8 cells; sigma_n is the unit vector
    along k_n x w_n for a random unit w_n; a_n ~ |k_n|^(-1/3) (a -5/3 spectrum sampled
    log-uniformly); psi_n uniform in [0, 2 pi).  The field is scaled so that its
    per-component RMS over the interior is intensity * |Uin_x|.  Each random quantity
    comes from its own counter-based stream of the seed (streams 16-23, apart from
    the uniform noise's 0-2).  The sum is separable: per component, one complex
    (nx ny x modes) @ (modes x nz) product.  The table's rows are
    (kx, ky, kz, psi, ax, ay, az) with a = the scaled amplitude times sigma, so that
    u'(x) = sum_n a_n cos(k_n . x + psi_n) everywhere (the turbulent inflow
    evaluates it upstream of the box)."""
    U = np.empty((mesh.L, 3))
    U[:] = np.asarray(uin, dtype=np.float64)
    if not intensity or not modes:
        return U, np.zeros((0, 7))
    g = mesh.gc
    nx, ny, nz = mesh.size
    x, y, z = mesh.x[g:nx - g], mesh.y[g:ny - g], mesh.z[g:nz - g]
    r = [uniform_pm1(seed, modes, stream=16 + q) for q in range(8)]
    ct = r[0]
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = math.pi * (r[1] + 1.0)
    d = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    ext = min(x[-1] - x[0], y[-1] - y[0], z[-1] - z[0])
    wmax = max(mesh.dx[g:nx - g].max(), mesh.dy[g:ny - g].max(), mesh.dz[g:nz - g].max())
    kmin, kmax = 2 * math.pi / ext, 2 * math.pi / (8 * wmax)
    kmag = kmin * (kmax / kmin) ** (0.5 * (r[2] + 1.0))
    k = d * kmag[:, None]
    w = np.stack(r[3:6], axis=1)
    sig = np.cross(d, w)
    nrm = np.linalg.norm(sig, axis=1)
    bad = nrm < 1e-3  # w (almost) along k: use any vector normal to k instead
    alt = np.where(np.abs(d[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    sig[bad] = np.cross(d[bad], alt[bad])
    sig /= np.linalg.norm(sig, axis=1)[:, None]
    amp = kmag ** (-1.0 / 3.0)
    psi = math.pi * (r[6] + 1.0)
    Ey = np.exp(1j * np.outer(k[:, 1], y))   # [modes][ny]
    Ez = np.exp(1j * np.outer(k[:, 2], z))   # [modes][nz]
    Ex = np.exp(1j * (np.outer(x, k[:, 0]) + psi[None, :]))  # [nx][modes]
    idx = np.arange(mesh.L).reshape(nx, ny, nz)[g:nx - g, g:ny - g, g:nz - g].ravel()
    up = np.empty((len(idx), 3))
    ni, nj, nk = len(x), len(y), len(z)
    for c in range(3):
        T = (Ex * (amp * sig[:, c])[None, :])[:, None, :] * Ey.T[None, :, :]  # [nx][ny][modes]
        up[:, c] = (T.reshape(ni * nj, modes) @ Ez).real.ravel()
        del T
    rms = math.sqrt(float(np.mean(up * up)))
    scale = intensity * abs(float(uin[0])) / rms
    U[idx] += up * scale
    table = np.concatenate([k, psi[:, None], (amp * scale)[:, None] * sig], axis=1)
    return U, np.ascontiguousarray(table)


# ----------------------------------------------------------------------------- turbines
def polar_table(step_deg: float = 1.0):
    atk = np.arange(-180.0, 180.0 + 0.5 * step_deg, step_deg)
    a = np.deg2rad(atk)
    cl = np.clip(2 * np.pi * a, -1.2, 1.2)
    cd = 0.01 + 0.05 * a * a
    return atk, cl, cd


@dataclass
class TurbineModel:
    """windturbine.json (src/alm.h:39-63, 151-152)."""
    R: float = 0.5
    hub_R: float = 0.0
    blades: int = 3
    tower: float = 4.0
    overhang: float = 0.0
    stations_r: Optional[np.ndarray] = None      # written under "r/R" (absolute r, Q5)
    stations_chord: Optional[np.ndarray] = None  # written under "chord/R" (absolute)
    stations_twist: Optional[np.ndarray] = None  # degrees
    atk: Optional[np.ndarray] = None
    cl: Optional[np.ndarray] = None              # [stations][atk]
    cd: Optional[np.ndarray] = None

    def to_json(self) -> str:
        af = []
        for s in range(len(self.stations_r)):
            af.append({"r/R": float(self.stations_r[s]), "chord/R": float(self.stations_chord[s]),
                       "twist[deg]": float(self.stations_twist[s]),
                       "Cl": [float(v) for v in self.cl[s]], "Cd": [float(v) for v in self.cd[s]]})
        return json.dumps({"R": self.R, "hub R": self.hub_R, "number of blades": self.blades,
                           "tower": self.tower, "overhang": self.overhang,
                           "blade": {"airfoil": af, "attack angle": [float(a) for a in self.atk]}})


def default_turbine(R=0.5, tower=4.0, n_stations=16, chord=0.05, twist_root=13.0, twist_tip=1.0):
    atk, cl, cd = polar_table()
    r = R * np.arange(n_stations) / (n_stations - 1)
    return TurbineModel(R=R, hub_R=0.0, blades=3, tower=tower, overhang=0.0,
                        stations_r=r, stations_chord=np.full(n_stations, chord),
                        stations_twist=twist_root + (twist_tip - twist_root) * r / R,
                        atk=atk, cl=np.tile(cl, (n_stations, 1)), cd=np.tile(cd, (n_stations, 1)))


@dataclass
class TurbineSite:
    """One entry of setup.json "windturbine array" (src/solver.h:702-708)."""
    base: tuple
    rot_speed: float
    angle: tuple = (0.0, 0.0, 0.0)  # roll, pitch, yaw: number (deg) or dict {amp,T,phase,offset}
    pitch: float = 0.0              # collective blade pitch, degrees (not in the reference)
    angle_type: Optional[int] = None  # force the angle type (0 roll, 1 pitch, 2 yaw) even at angle 0

    def to_dict(self):
        return {"base": list(self.base), "angle": list(self.angle), "rotation speed": self.rot_speed}


def sites_json(sites: List[TurbineSite]) -> str:
    return json.dumps([s.to_dict() for s in sites])


# ----------------------------------------------------------------------------- cases
@dataclass
class Case:
    name: str
    mesh: Mesh
    uin: tuple = (1.0, 0.0, 0.0)
    Re: float = 1e4
    Cs: float = 0.1
    dt: float = 0.005
    method: str = "SOR"          # "SOR" | "BiCG"
    pc: str = "Jacobi"           # "Jacobi" | "SOR"
    omega: float = 1.2
    pc_omega: float = 1.0
    pc_it: int = 3
    max_it: int = 1000
    tol: float = 1e-6
    noise: float = 0.05
    seed: int = 1
    # > 0: the initial perturbation is `modes` random Fourier modes of RMS intensity
    # `noise` (fourier_mode_velocity) instead of uniform per-cell noise
    modes: int = 0
    steps: int = 100
    turbine: Optional[TurbineModel] = None
    sites: List[TurbineSite] = field(default_factory=list)
    ap_per_blade: int = 64
    width: float = 0.125
    # output path (setup.json "output" and "probe", src/solver.h:569-598, 656-700):
    # times of the first snapshot, between snapshots and of the first averaged step,
    # and probe positions; the defaults switch it off
    output_start: float = 1e9
    output_interval: float = 1e9
    tavg_start: float = 1e9
    probes: List[tuple] = field(default_factory=list)
    # turbulence-generating grid (setup.json "turbulence grid", src/solver.h:598-612):
    # {"bar": width, "mesh": spacing, "thick": x-thickness of a layer, "placement": [x, y, z]}
    tgg: Optional[dict] = None
    # retracted floor (setup.json "floor start", src/solver.h:620-635): slip floor for x < floor_start
    floor_start: Optional[float] = None
    # turbulent inflow (not in the reference, whose x- inflow is the constant Uin): with
    # modes > 0 the x- ghost cells and face flux take the initial field's Fourier modes
    # as frozen turbulence convected at Uin_x, Uin + u'(x - Uin_x t, y, z)
    inflow: bool = False
    _modes_cache: Optional[tuple] = field(default=None, init=False, repr=False, compare=False)

    def _fourier(self):
        if self._modes_cache is None:
            self._modes_cache = fourier_modes(self.mesh, self.uin, self.noise, self.seed, self.modes)
        return self._modes_cache

    def U0(self) -> np.ndarray:
        if self.modes:
            return self._fourier()[0].copy()
        return initial_velocity(self.mesh, self.uin, self.noise * abs(self.uin[0]), self.seed)

    def inflow_modes(self) -> Optional[np.ndarray]:
        """[n][7] mode table of the turbulent inflow (None: the constant Uin)."""
        if not (self.inflow and self.modes):
            return None
        return self._fourier()[1]

    def inflow_speed(self) -> float:
        return float(self.uin[0])


def case_c1(nx=64, ny=48, nz=48, steps=100) -> Case:
    """Plain FP64, no turbine: 64x48x48 on 10x4x4, tanh beta=1.2 in y,z, SOR 1.2 to 1e-6."""
    m = make_mesh(nx, ny, nz, ((0.0, 10.0), (-2.0, 2.0), (0.0, 4.0)), 1.2, 1.2)
    return Case("c1", m, method="SOR", omega=1.2, tol=1e-6, max_it=1000, steps=steps)


def case_c2(nx=256, ny=128, nz=128, steps=1000, ap_per_blade=64) -> Case:
    """One 3-blade rotor (D=1 at x=3), TSR 6, eps = 2 dx, BiCGStab(Jacobi 3) to 1e-8."""
    m = make_mesh(nx, ny, nz, ((0.0, 16.0), (-4.0, 4.0), (0.0, 8.0)), 1.5, 1.5)
    dx = 16.0 / nx
    t = default_turbine(R=0.5, tower=4.0)
    return Case("c2", m, method="BiCG", pc="Jacobi", pc_it=3, tol=1e-8, max_it=1000,
                steps=steps, turbine=t, sites=[TurbineSite((3.0, 0.0, 0.0), 12.0)],
                ap_per_blade=ap_per_blade, width=2 * dx)


def case_c3(nx=512, ny=256, nz=256, steps=500, ap_per_blade=64, tol=4e-9) -> Case:
    """The precision form at scale: 512x256x256 on a 16x8x8 box, two rotors (D = 1)
    5 D apart at x = 3 and 8, TSR 6, eps = 2 dx, a seeded isotropic turbulence of 200
    random Fourier modes at 10 % intensity as the initial field and, frozen and
    convected at Uin, as the inflow; BiCGStab(Jacobi 3) to 4e-9, the tolerance that
    keeps avg div <= 1e-6 at this size (dt * max_diag = 230), at most 6000
    iterations (the first steps after the impulsive start need 1000-3000)."""
    m = make_mesh(nx, ny, nz, ((0.0, 16.0), (-4.0, 4.0), (0.0, 8.0)), 1.5, 1.5)
    dx = 16.0 / nx
    t = default_turbine(R=0.5, tower=4.0)
    sites = [TurbineSite((3.0, 0.0, 0.0), 12.0), TurbineSite((8.0, 0.0, 0.0), 12.0)]
    return Case("c3", m, method="BiCG", pc="Jacobi", pc_it=3, tol=tol, max_it=6000, steps=steps, noise=0.10,
                modes=200, inflow=True, turbine=t, sites=sites, ap_per_blade=ap_per_blade, width=2 * dx)


def case_c4(nx=192, ny=96, nz=96, steps=200, ap_per_blade=64, yaw_deg=(5.0, 0.0, -5.0),
            pitch_deg=(0.0, 0.0, 0.0)) -> Case:
    """Differentiation form: three rotors (D=1) in tandem at x = 3, 8, 13 on a 24x8x8
    box, yaw set explicitly on each (angle type yaw even at 0), collective pitch,
    TSR 6, eps = 2 dx, BiCGStab(Jacobi 3) to 1e-8."""
    m = make_mesh(nx, ny, nz, ((0.0, 24.0), (-4.0, 4.0), (0.0, 8.0)), 1.5, 1.5)
    dx = 24.0 / nx
    t = default_turbine(R=0.5, tower=4.0)
    sites = [TurbineSite((x, 0.0, 0.0), 12.0, angle=(0.0, 0.0, yw), pitch=pt, angle_type=2)
             for x, yw, pt in zip((3.0, 8.0, 13.0), yaw_deg, pitch_deg)]
    return Case("c4", m, method="BiCG", pc="Jacobi", pc_it=3, tol=1e-8, max_it=1000, steps=steps,
                turbine=t, sites=sites, ap_per_blade=ap_per_blade, width=2 * dx)


def case_c5(nx=768, ny=384, nz=384, steps=2000, ap_per_blade=96) -> Case:
    """The largest configuration: 768x384x384 on a 32x16x8 box, a 3x3 array of
    rotors (D = 1) at 6 D along x and 4 D along y from (4, -4, 0), TSR 6,
    eps = 2 dx, a turbulence of 200 random Fourier modes at 8 % intensity as the
    initial field and the inflow, BiCGStab(Jacobi 3) to 1e-8, at most 6000
    iterations."""
    m = make_mesh(nx, ny, nz, ((0.0, 32.0), (-8.0, 8.0), (0.0, 8.0)), 1.5, 1.5)
    dx = 32.0 / nx
    t = default_turbine(R=0.5, tower=4.0)
    sites = [TurbineSite((4.0 + 6.0 * a, -4.0 + 4.0 * b, 0.0), 12.0) for a in range(3) for b in range(3)]
    return Case("c5", m, method="BiCG", pc="Jacobi", pc_it=3, tol=1e-8, max_it=6000, steps=steps, noise=0.08,
                modes=200, inflow=True,
                turbine=t, sites=sites, ap_per_blade=ap_per_blade, width=2 * dx)


def case_adjoint_small(nx=20, ny=12, nz=12, steps=4, ap_per_blade=6) -> Case:
    """A few-thousand-cell rotor case for finite-difference checks of the adjoint:
    one rotor (R = 0.5) at x = 1 on a 4x3x3 box, yaw 10 deg, pitch 2 deg,
    BiCGStab(SOR 2) converged to 1e-13."""
    m = make_mesh(nx, ny, nz, ((0.0, 4.0), (-1.5, 1.5), (0.0, 3.0)), 1.3, 1.3)
    t = default_turbine(R=0.5, tower=1.5)
    site = TurbineSite((1.0, 0.0, 0.0), 12.0, angle=(0.0, 0.0, 10.0), pitch=2.0, angle_type=2)
    return Case("adj", m, method="BiCG", pc="SOR", pc_it=2, pc_omega=1.2, tol=1e-13, max_it=5000,
                steps=steps, turbine=t, sites=[site], ap_per_blade=ap_per_blade, width=4.0 / nx * 2)


def case_output_small(steps=4) -> Case:
    """A 10x6x6 rotor case with every output switched on: snapshots at t = 0, 2 dt,
    4 dt, ..., the time average from t = dt, three probes (one outside the interior,
    so inactive).  SOR to 1e-8 keeps the device and CPU paths identical."""
    m = make_mesh(10, 6, 6, ((0.0, 4.0), (-1.5, 1.5), (0.0, 3.0)), 1.3, 1.3)
    t = default_turbine(R=0.5, tower=1.5)
    dt = 0.005
    return Case("out", m, method="SOR", omega=1.2, tol=1e-8, max_it=2000, steps=steps, dt=dt, turbine=t,
                sites=[TurbineSite((1.2, 0.0, 0.0), 12.0)], ap_per_blade=6, width=0.8,
                output_start=0.0, output_interval=2 * dt, tavg_start=dt,
                probes=[(1.7, 0.1, 1.4), (3.1, -0.6, 0.5), (9.0, 0.0, 1.0)])


def case_tgg_small(nx=32, ny=12, nz=12, steps=3) -> Case:
    """A grid of bars across the inflow (the reference's tgg cases, scaled): box
    8 x 3 x 3, bars 0.18 wide every 0.6, layers 0.15 thick ending at x = 2,
    floor retracted up to x = 1.5 (the option those cases use), SOR 1.5 to 1e-6."""
    m = make_mesh(nx, ny, nz, ((0.0, 8.0), (-1.5, 1.5), (0.0, 3.0)), 1.2, 1.2)
    c = Case("tgg", m, method="SOR", omega=1.5, tol=1e-6, max_it=2000, steps=steps,
             tgg={"bar": 0.18, "mesh": 0.6, "thick": 0.15, "placement": [2.0, 0.3, 1.2]}, floor_start=1.5)
    return c


def turbine_records(turbine: TurbineModel, sites: List[TurbineSite]) -> np.ndarray:
    """Turbine rows (include/dalm_hip.h layout), src/alm.h:144-206: rotation centre
    (-overhang, 0, tower); the angle type is the first non-zero or object entry of
    (roll, pitch, yaw), with formula (0, 0, 0, rad(a)) for a number and
    (rad(amp), 2 pi / T, rad(phase), rad(offset)) for an object.  All-zero angles
    give type "none" (the reference leaves that case unset) unless the site names
    its type explicitly."""
    wt = np.zeros((len(sites), 16))
    for w, s in enumerate(sites):
        wt[w, 0:3] = s.base
        wt[w, 3:6] = (-turbine.overhang, 0.0, turbine.tower)
        wt[w, 6] = 3
        for j, a in enumerate(s.angle):
            if s.angle_type is not None and j != s.angle_type:
                continue
            if isinstance(a, dict) or a != 0 or s.angle_type is not None:
                wt[w, 6] = j
                if isinstance(a, dict):
                    wt[w, 9:13] = (a["amp"] * math.pi / 180, 2 * math.pi / a["T"],
                                   a["phase"] * math.pi / 180, a["offset"] * math.pi / 180)
                else:
                    wt[w, 9:13] = (0.0, 0.0, 0.0, a * math.pi / 180)
                break
        wt[w, 13] = s.rot_speed
    return wt


def write_case_dir(case: Case, path: str, time_steps: Optional[int] = None) -> None:
    """Write setup.json / mesh.json / windturbine.json for a Case in the reference's
    schemas (src/solver.h:564-708, src/io.h:105-151, src/alm.h:39-63)."""
    import os
    os.makedirs(path, exist_ok=True)
    steps = case.steps if time_steps is None else time_steps
    eq = {"tolerance": case.tol, "max iteration": case.max_it,
          "method": "BiCG" if case.method == "BiCG" else "SOR"}
    if case.method == "BiCG":
        eq["preconditioner"] = {"method": case.pc, "max iteration": case.pc_it,
                                "relaxation rate": case.pc_omega}
    else:
        eq["relaxation rate"] = case.omega
    setup = {"runtime": {"dt": case.dt, "time": case.dt * (steps + 0.5)},
             "output": {"output start time": case.output_start, "output interval time": case.output_interval,
                        "time avg start time": case.tavg_start},
             "inflow": {"value": list(case.uin)}, "cfd": {"Re": case.Re, "Cs": case.Cs}, "eq": eq}
    if case.probes:
        setup["probe"] = [list(p) for p in case.probes]
    if case.floor_start is not None:
        setup["floor start"] = case.floor_start
    if case.tgg is not None:
        setup["turbulence grid"] = {"bar": case.tgg["bar"], "mesh": case.tgg["mesh"], "thick": case.tgg["thick"],
                                    "placement": list(case.tgg["placement"])}
    if case.turbine is not None and case.sites:
        setup["windturbine array"] = [s.to_dict() for s in case.sites]
        setup["alm"] = {"blade points": case.ap_per_blade, "projection width": case.width}
        with open(os.path.join(path, "windturbine.json"), "w") as fh:
            fh.write(case.turbine.to_json())
    with open(os.path.join(path, "setup.json"), "w") as fh:
        fh.write(json.dumps(setup, indent=1))
    with open(os.path.join(path, "mesh.json"), "w") as fh:
        fh.write(case.mesh.to_json())
