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
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1], order=L.OrderingType.GIVEN)
    assert st < 0 and "perm" in lib.kkt_ldl_last_error().decode()
    bad = np.array([0, 0], np.int32)
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1], order=L.OrderingType.GIVEN,
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


class _RefEnum:
    """Stand-in for a value of the reference's pybind OrderingType: an int
    with a ``name`` (amigo/amigo.cpp:324-334)."""

    def __init__(self, name, value):
        self.name, self.value = name, value

    def __int__(self):
        return self.value

    def __str__(self):
        return f"OrderingType.{self.name}"


def test_ordering_codes_follow_the_reference_enum():
    """include/ordering_utils.h:21: NESTED_DISSECTION=0, AMD=1, MULTI_COLOR=2,
    NATURAL=3.  Each reference value selects its own ordering."""
    assert (L.OrderingType.NESTED_DISSECTION, L.OrderingType.AMD, L.OrderingType.MULTI_COLOR,
            L.OrderingType.NATURAL) == (0, 1, 2, 3)
    from amigo_amd import workloads as W
    P = W.poisson2d(N=12)
    mat = L.CSRMat(P.n, P.n, P.rowp, P.cols, P.vals)
    perms = {}
    for name, value in (("NESTED_DISSECTION", 0), ("AMD", 1), ("NATURAL", 3)):
        for order in (value, _RefEnum(name, value), name, f"OrderingType.{name}"):
            perms.setdefault(name, []).append(L.SparseLDL(mat, order=order).get_perm()[0])
    for name, ps in perms.items():
        assert all(np.array_equal(ps[0], q) for q in ps), name
    assert np.array_equal(perms["NATURAL"][0], np.arange(P.n))
    assert not np.array_equal(perms["AMD"][0], perms["NATURAL"][0])
    assert not np.array_equal(perms["NESTED_DISSECTION"][0], perms["AMD"][0])
    assert sorted(perms["NESTED_DISSECTION"][0]) == list(range(P.n))
    assert L.SparseLDL(mat, order=_RefEnum("DEFAULT", 1)).get_perm()[0].tolist() == \
        perms["AMD"][0].tolist()
    for bad in (2, _RefEnum("MULTI_COLOR", 2), 5, -1, "SPECTRAL"):
        with pytest.raises(ValueError):
            L.SparseLDL(mat, order=bad)


def test_abi_rejects_unknown_orderings():
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1], order=L.OrderingType.MULTI_COLOR)
    assert st < 0 and "MULTI_COLOR" in lib.kkt_ldl_last_error().decode()
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1], order=7)
    assert st < 0 and "unknown ordering" in lib.kkt_ldl_last_error().decode()


def test_refactor_and_pivot_threshold_arguments():
    lib, st, h = _create(2, [0, 2, 4], [0, 1, 0, 1])
    assert st == 0
    d = np.ones(2)
    # no values loaded yet: the refactorization has nothing to reuse
    assert lib.kkt_ldl_refactor(h, d.ctypes.data_as(ctypes.c_void_p), 0, None) == -3
    assert "before factor" in lib.kkt_ldl_last_error().decode()
    assert lib.kkt_ldl_set_pivot_threshold(h, 0.1, -1.0) == 0
    assert lib.kkt_ldl_set_pivot_threshold(h, float("nan"), -1.0) < 0
    lib.kkt_ldl_destroy(h)


def test_refinement_policy(tmp_path):
    """The refinement's host decisions (amigo_amd/csrc/refine_policy.hpp),
    compiled on their own: a NaN or infinite residual never reads as
    converged, a step that raises the residual is taken back."""
    import subprocess
    drv = tmp_path / "drv.cpp"
    drv.write_text(r'''
#include <cstdio>
#include <cmath>
#include "refine_policy.hpp"
using namespace kktldl;
int main() {
  const double nan = std::nan(""), inf = INFINITY;
  double b[3] = {4.0, 1.0, 0.0}, r[3];
  r[0] = 4e-20; r[1] = 1e-22; r[2] = 0.0;  printf("%.3e\n", refine_worst(b, r, 3));
  r[1] = nan;                               printf("%.3e\n", refine_worst(b, r, 3));
  r[1] = 1e-22; r[0] = nan;                 printf("%.3e\n", refine_worst(b, r, 3));
  r[0] = inf;                               printf("%.3e\n", refine_worst(b, r, 3));
  r[0] = 4e-20; r[2] = 9.0;                 printf("%.3e\n", refine_worst(b, r, 3));
  printf("%d %d %d %d %d %d %d\n",
         refine_next(0, 1e-12, inf, 1e-10, 10),   // converged
         refine_next(0, 1e-6, inf, 1e-10, 10),    // go on
         refine_next(0, inf, inf, 1e-10, 10),     // non-finite plain solve: stop
         refine_next(2, 1e-8, 1e-9, 1e-10, 10),   // residual grew: undo
         refine_next(2, 1e-9, 1e-9, 1e-10, 10),   // no progress: undo
         refine_next(1, inf, 1e-6, 1e-10, 10),    // NaN after a step: undo
         refine_next(10, 1e-9, 1e-8, 1e-10, 10)); // step limit
  return 0;
}
''')
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "amigo_amd", "csrc"),
                    str(drv), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert float(out[0]) == pytest.approx(1e-10)
    assert out[1] == "inf" and out[2] == "inf" and out[3] == "inf"
    assert float(out[4]) == pytest.approx(3.0)
    assert out[5].split() == ["0", "1", "0", "2", "2", "2", "0"]
