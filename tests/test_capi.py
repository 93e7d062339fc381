"""The C ABI library loads and exports every symbol include/hyperdeal_amd.h
declares (no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hyperdeal_amd.h")
LIB = os.path.join(ROOT, "hyperdeal_amd", "libhyperdeal_amd.so")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("hd_matrix_free_create", "hd_advection_create",
                     "hd_advection_apply", "hd_lsrk_stage", "hd_lsrk_step",
                     "hd_last_error"):
        assert required in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_lists_the_same_symbols():
    import hyperdeal_amd
    assert sorted(hyperdeal_amd.EXPORTED) == declared_functions()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_host_only_entry_points():
    """Entry points that never touch the device: version, RK tables,
    argument validation and error reporting."""
    lib = ctypes.CDLL(LIB)
    lib.hd_last_error.restype = ctypes.c_char_p
    ma, mi = ctypes.c_int(), ctypes.c_int()
    assert lib.hd_version(ctypes.byref(ma), ctypes.byref(mi)) == 0
    a = (ctypes.c_double * 8)()
    b = (ctypes.c_double * 9)()
    assert lib.hd_lsrk_coefficients(b"rk45", a, b) == 5
    import oracle as O
    oa, ob = O.lsrk_coefficients("rk45")
    assert list(a[:4]) == list(oa) and list(b[:5]) == list(ob)
    for name, n in (("rk33", 3), ("rk47", 7), ("rk59", 9)):
        assert lib.hd_lsrk_coefficients(name.encode(), a, b) == n
    assert lib.hd_lsrk_coefficients(b"rk99", a, b) == -1
    assert b"unknown RK scheme" in lib.hd_last_error()
    # invalid mesh description is rejected before any device call
    import hyperdeal_amd as H
    with pytest.raises(H.HyperdealError, match="dim_x and dim_v"):
        H.MatrixFree(4, 1, 3, [2] * 5, [0] * 5, [1] * 5)
    with pytest.raises(H.HyperdealError, match="n_points"):
        H.MatrixFree(1, 1, 3, [2, 2], [0, 0], [1, 1], n_points=2)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_matrix_free_sizes_without_device():
    import hyperdeal_amd as H
    mf = H.MatrixFree(3, 3, 3, [8] * 6, [0] * 6, [1] * 6)
    assert mf.n_dofs() == 8 ** 6 * 4 ** 6 == 1073741824
    assert mf.dofs_per_cell() == 4096
    assert mf.n_cells_total() == 262144
    dt = H.compute_critical_time_step(mf, [1, 0.5, 0.25, 0.125, 0.1, 0.1])
    assert dt == pytest.approx(1.0 / 8.0)
