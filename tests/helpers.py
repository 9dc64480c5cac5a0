"""Shared helpers for the tests: seeded fields on a grid and comparisons."""
import importlib

import numpy as np

cases = importlib.import_module("3dalm_amd.cases")


def small_mesh(nx=10, ny=9, nz=7, beta=1.3):
    return cases.make_mesh(nx, ny, nz, ((0.0, 2.0), (-1.0, 1.0), (0.0, 1.5)), beta, beta)


def rand_field(mesh, seed, comps=3, scale=1.0, offset=0.0):
    rng = np.random.default_rng(seed)
    a = offset + scale * rng.uniform(-1, 1, size=(mesh.L, comps))
    return np.ascontiguousarray(a if comps > 1 else a[:, 0])


def interior_mask(mesh, grow=0):
    nx, ny, nz = mesh.size
    g = mesh.gc - grow
    m = np.zeros((nx, ny, nz), dtype=bool)
    m[g:nx - g, g:ny - g, g:nz - g] = True
    return m.reshape(-1)


def face_mask(mesh):
    """cells [gc-1, n-gc) in every direction (the face-flux box)."""
    nx, ny, nz = mesh.size
    g = mesh.gc
    m = np.zeros((nx, ny, nz), dtype=bool)
    m[g - 1:nx - g, g - 1:ny - g, g - 1:nz - g] = True
    return m.reshape(-1)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def max_rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)
