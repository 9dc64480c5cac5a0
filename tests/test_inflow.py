"""The turbulent inflow of the oracle (or_inflow_eval / or_apply_inflow_*), which the
reference does not have (its x- inflow is the constant Uin, src/bc.h:16-30 and
243-257), pinned by what the construction fixes:

  * every mode of the table is divergence-free (a_n . k_n = 0) and the table
    reproduces the initial field that fourier_modes builds by a separable complex
    product -- two independent evaluations of the same modal sum;
  * the shift property of frozen turbulence: the inflow at shift s = speed t on the
    ghost plane x_g equals the initial field at x_g - s; with s = -(a multiple of
    the uniform dx) that is an interior plane of the initial field, for the cells
    and (from the face position) for the face velocity;
  * whole steps: the constant inflow is untouched when no modes are set (bit for
    bit), the inflow values land on the x- ghosts and face exactly, and a case with
    the inflow stays divergence-free to the solve's tolerance."""
import ctypes as C
import importlib

import numpy as np

cases = importlib.import_module("3dalm_amd.cases")
po = importlib.import_module("pyoracle")


def _case(modes=24, inflow=True):
    m = cases.make_mesh(24, 12, 10, ((0.0, 6.0), (-1.5, 1.5), (0.0, 3.0)), 1.3, 1.3)
    c = cases.Case("inflow", m, method="BiCG", pc="Jacobi", pc_it=3, tol=1e-12, max_it=4000, steps=3,
                   noise=0.1, modes=modes, inflow=inflow)
    return c


def _eval(case, shift):
    L, _ = po.lib("oracle")
    m = case.mesh
    g = m.gc
    ny, nz = m.size[1] - 2 * g, m.size[2] - 2 * g
    cells, face = np.zeros(g * ny * nz * 3), np.zeros(ny * nz)
    modes = np.ascontiguousarray(case.inflow_modes())
    uin = np.array(case.uin, dtype=np.float64)
    grid = po.make_grid(m)
    L.or_inflow_eval(C.byref(grid), po.ptr(modes), len(modes), po.ptr(uin), shift, po.ptr(cells), po.ptr(face))
    return cells.reshape(g, ny, nz, 3), face.reshape(ny, nz)


def test_modes_divergence_free_and_reproduce_initial_field():
    case = _case()
    T = case.inflow_modes()
    k, a = T[:, 0:3], T[:, 4:7]
    assert T.shape == (24, 7)
    assert np.max(np.abs(np.sum(k * a, axis=1))) <= 1e-14 * np.max(np.abs(k)) * np.max(np.abs(a))
    # the cosine sum of the table at every interior cell = the complex product's field
    m = case.mesh
    g = m.gc
    nx, ny, nz = m.size
    X, Y, Z = np.meshgrid(m.x[g:nx - g], m.y[g:ny - g], m.z[g:nz - g], indexing="ij")
    arg = (X[..., None] * k[:, 0] + Y[..., None] * k[:, 1] + Z[..., None] * k[:, 2] + T[:, 3])
    up = np.cos(arg) @ a
    U = case.U0().reshape(nx, ny, nz, 3)[g:nx - g, g:ny - g, g:nz - g]
    ref = U - np.asarray(case.uin)
    assert np.max(np.abs(up - ref)) <= 1e-12 * np.max(np.abs(ref))
    # the per-component RMS is the stated intensity
    assert abs(np.sqrt(np.mean(ref * ref)) - 0.1) <= 1e-12


def test_shift_property_against_initial_field():
    """Uniform x: x_i = x_0 + i dx.  The ghost plane i (< gc) at shift -(n dx)
    samples the initial field at x_{i+n}; the face x_f = (x_{g-1} + x_g)/2 at
    shift -(n dx + dx/2) samples the x-velocity at cell centre x_{g+n}."""
    case = _case()
    m = case.mesh
    g = m.gc
    nx, ny, nz = m.size
    dx = m.x[1] - m.x[0]
    U = case.U0().reshape(nx, ny, nz, 3)
    for n in (g + 1, g + 7, nx - 2 * g - 1):
        cells, _ = _eval(case, -n * dx)
        for i in range(g):
            want = U[i + n, g:ny - g, g:nz - g]
            assert np.max(np.abs(cells[i] - want)) <= 1e-12, (n, i)
        _, face = _eval(case, -(n - g) * dx - 0.5 * dx)
        want = U[n, g:ny - g, g:nz - g, 0]
        assert np.max(np.abs(face - want)) <= 1e-12, n


def test_single_mode_closed_form():
    """One mode, k = (2, 0, 0), a = (0, 0.3, 0), psi = 0.4 (divergence-free): the
    ghost values are Uin + (0, 0.3 cos(2 (x - s) + 0.4), 0), the face uses x_f."""
    case = _case()
    case._modes_cache = (case.U0(), np.array([[2.0, 0.0, 0.0, 0.4, 0.0, 0.3, 0.0]]))
    m = case.mesh
    g = m.gc
    s = 0.37
    cells, face = _eval(case, s)
    for i in range(g):
        assert np.all(cells[i, :, :, 0] == 1.0) and np.all(cells[i, :, :, 2] == 0.0)
        assert np.allclose(cells[i, :, :, 1], 0.3 * np.cos(2 * (m.x[i] - s) + 0.4), rtol=0, atol=1e-15)
    assert np.all(face == 1.0)  # a_x = 0: the face keeps Uin_x


def test_whole_steps_with_inflow():
    """Modes of zero amplitude give the constant-inflow step bit for bit (Uin + 0 is
    Uin); with the real modes the x- ghosts and face carry the inflow of the step's
    time and the projection identity holds."""
    plain = _case(modes=24, inflow=False)
    zero = _case(modes=24, inflow=True)
    U0, T = zero._fourier()
    T = T.copy()
    T[:, 4:7] = 0.0
    zero._modes_cache = (U0, T)
    o1, o2 = po.OracleSolver(plain), po.OracleSolver(zero)
    for _ in range(2):
        o1.step(), o2.step()
    for name in ("U", "p", "JU"):
        assert np.array_equal(o1.field(name), o2.field(name))
    case = _case()
    o = po.OracleSolver(case)
    st = None
    for _ in range(2):
        st = o.step()
    assert st["err"] <= case.tol
    m = case.mesh
    g = m.gc
    nx, ny, nz = m.size
    t = case.dt * st["step"]
    cells, face = _eval(case, case.uin[0] * t)
    U = o.field("U").reshape(nx, ny, nz, 3)
    JU = o.field("JU").reshape(nx, ny, nz, 3)
    assert np.array_equal(U[:g, g:ny - g, g:nz - g], cells)
    ayz = m.dy[g:ny - g, None] * m.dz[None, g:nz - g]
    assert np.array_equal(JU[g - 1, g:ny - g, g:nz - g, 0], ayz * face)
    # avg div = dt * max_diag * err (the projection identity) with the inflow flux;
    # at err ~ 1e-12 the divergence's own rounding (~1e-17) limits it to ~1e-4
    assert abs(st["avg_div"] - case.dt * st["max_diag"] * st["err"]) <= 1e-3 * st["avg_div"]
    # the turbulent inflow changes the flow: the first interior plane differs from
    # the constant-inflow run's
    o3 = po.OracleSolver(plain)
    for _ in range(2):
        o3.step()
    assert not np.allclose(o3.field("U").reshape(nx, ny, nz, 3)[g], U[g])
