"""The C-ABI library loads without a GPU and exports every entry point that
include/dalm_hip.h declares; host-only helpers agree with the oracle."""
import ctypes
import importlib

import numpy as np

hip = importlib.import_module("3dalm_amd.hip")
cases = importlib.import_module("3dalm_amd.cases")
po = importlib.import_module("pyoracle")


def test_exports_every_declared_symbol():
    L = hip.lib()
    names = hip.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_helpers_match_oracle():
    """dalm_mesh_widths and dalm_build_ap_props are host code: no GPU needed."""
    m = cases.make_mesh(9, 7, 6, ((0.0, 2.0), (-1.0, 1.0), (0.0, 1.0)), 1.2, 1.4)
    out = np.zeros_like(m.y)
    assert hip.lib().dalm_mesh_widths(hip.ptr(m.y), len(m.y), hip.ptr(out)) == 0
    assert np.array_equal(out, m.dy)
    t = cases.default_turbine()
    a1 = hip.build_ap_props(t, 2, 9)
    a2 = po.ap_props_oracle(t, 2, 9)
    for x, y in zip(a1, a2):
        assert np.array_equal(x, y)


def test_bad_arguments_rejected_without_gpu():
    L = hip.lib()
    h = ctypes.c_void_p()
    size = (ctypes.c_int64 * 3)(8, 8, 8)
    x = np.zeros(8)
    # gc < 2 is refused before any device work (src/solver.h:615 asserts gc >= 2)
    assert L.dalm_create(ctypes.byref(h), size, 1, hip.ptr(x), hip.ptr(x), hip.ptr(x), 0) == hip.BAD_ARG
    # fewer than 3 interior cells along an axis, a padded box past 32-bit cell
    # indices (also where the product of the sizes would overflow 64 bits), an
    # unknown precision: refused before any device work
    for bad in ((6, 8, 8), (8, 8, 6), (2048, 2048, 512), (1 << 40, 1 << 40, 1 << 40), (8, 1 << 62, 8)):
        size = (ctypes.c_int64 * 3)(*bad)
        assert L.dalm_create(ctypes.byref(h), size, 2, hip.ptr(x), hip.ptr(x), hip.ptr(x), 0) == hip.BAD_ARG, bad
    size = (ctypes.c_int64 * 3)(8, 8, 8)
    assert L.dalm_create(ctypes.byref(h), size, 2, hip.ptr(x), hip.ptr(x), hip.ptr(x), 7) == hip.BAD_ARG
    assert L.dalm_create(ctypes.byref(h), size, 1 << 62, hip.ptr(x), hip.ptr(x), hip.ptr(x), 0) == hip.BAD_ARG
    assert L.dalm_mesh_widths(hip.ptr(x), 2, hip.ptr(x)) == hip.BAD_ARG
