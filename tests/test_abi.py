"""The C ABI: every symbol include/kkt_ldl.h declares is exported by
libkktldl.so, argument checking returns negative codes with a message, and
no call needs a GPU before the first factorization."""

import ctypes
import os
import re

import numpy as np
import pytest

from amigo_amd import ldl as L

from .conftest import ROOT, has_gpu


def _header_functions():
    text = open(os.path.join(ROOT, "include", "kkt_ldl.h")).read()
    return sorted(set(re.findall(r"\b(kkt_ldl_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = L.load_library()
    decl = _header_functions()
    assert len(decl) >= 14
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(L.EXPORTS) == decl


def test_every_assembly_symbol_is_exported():
    from amigo_amd import assembly as A
    lib = L.load_library()
    text = open(os.path.join(ROOT, "include", "kkt_asm.h")).read()
    decl = sorted(set(re.findall(r"\b(kkt_asm_[a-z_]+)\s*\(", text)))
    assert len(decl) == 5
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(A.EXPORTS) == decl


def test_default_options_match_amigo_solver():
    lib = L.load_library()
    o = L.Options()
    lib.kkt_ldl_default_options(ctypes.byref(o))
    assert o.ustab == 0.01 and o.pivot_tol == 1e-14 and o.delay_growth == 2.0
    assert o.order == L.OrderingType.AMD and o.solver_type == 0
    assert o.refine_tol == 1e-10


def _create(n, rowp, cols, **kw):
    lib = L.load_library()
    o = L.Options()
    lib.kkt_ldl_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    h = ctypes.c_void_p()
    rowp = np.ascontiguousarray(rowp, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    st = lib.kkt_ldl_create(n, L._ip(rowp), L._ip(cols), ctypes.byref(o), ctypes.byref(h))
    return lib, st, h


@pytest.mark.parametrize("rowp,cols,msg", [
    ([0, 2, 3], [1, 0, 1], "sorted"),
    ([0, 1, 2], [0, 5], "range"),
    ([0, 2, 1], [0, 1], "monotone"),
])
def test_create_rejects_bad_patterns(rowp, cols, msg):
    lib, st, h = _create(2, rowp, cols)
    assert st < 0
    assert msg in lib.kkt_ldl_last_error().decode()


def test_given_order_requires_a_permutation():
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1], order=3)
    assert st < 0 and "perm" in lib.kkt_ldl_last_error().decode()
    bad = np.array([0, 0], np.int32)
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1], order=3,
                         perm=bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert st < 0 and "permutation" in lib.kkt_ldl_last_error().decode()


def test_calls_before_factor_are_errors():
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1])
    assert st == 0
    a, b = ctypes.c_int(), ctypes.c_int()
    assert lib.kkt_ldl_inertia(h, ctypes.byref(a), ctypes.byref(b), None) < 0
    x = np.zeros(2)
    assert lib.kkt_ldl_solve(h, 1, x.ctypes.data_as(ctypes.c_void_p), 2, 0, None) < 0
    lib.kkt_ldl_destroy(h)


@pytest.mark.skipif(has_gpu(), reason="checks the no-GPU error path")
def test_factor_without_gpu_fails_loudly():
    """No CPU fallback: without a device the factorization reports an error."""
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1])
    vals = np.array([2.0, 1.0, 1.0, 2.0])
    st = lib.kkt_ldl_factor(h, vals.ctypes.data_as(ctypes.c_void_p), None, 0, None)
    assert st < 0
    assert "HIP" in lib.kkt_ldl_last_error().decode()
    lib.kkt_ldl_destroy(h)
