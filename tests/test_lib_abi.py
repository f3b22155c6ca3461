"""The C-ABI library: loads, exports every entry point include/pyqed_amd.h declares, reports errors
through pyqed_last_error, and builds tables on the host.  No GPU compute calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "pyqed_amd.h")).read()
    return sorted(set(re.findall(r"\b(pyqed_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    from pyqed_amd import _lib
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) <= set(_lib.SYMBOLS) | {"pyqed_heom_tables_size", "pyqed_heom_tables_build"}


def test_version_and_error_reporting():
    from pyqed_amd import _lib
    L = _lib.lib()
    assert L.pyqed_version() >= 100
    rc = L.pyqed_heom_tables_build(-1, 2, None, None, None)
    assert rc == 1
    assert b"bad K or L" in L.pyqed_last_error()
    n = C.c_int64()
    assert L.pyqed_heom_tables_size(40, 20, C.byref(n)) == 1  # C(60, 20) > 2^31
    assert b"exceeds" in L.pyqed_last_error()
    with pytest.raises(_lib.PyqedError):
        _lib.check(L.pyqed_heom_tables_build(300, 1, None, None, None))


def test_tables_size_matches_binomial():
    from math import comb
    from pyqed_amd import _lib
    L = _lib.lib()
    n = C.c_int64()
    for K, Lv in [(2, 5), (14, 4), (48, 3), (32, 4), (0, 3), (5, 0)]:
        assert L.pyqed_heom_tables_size(K, Lv, C.byref(n)) == 0
        assert n.value == comb(K + Lv, Lv)


def test_tables_against_oracle_random_shapes():
    from hypothesis import given, settings, strategies as st
    from oracle import heom as O
    from pyqed_amd._lib import hierarchy_tables

    @settings(max_examples=25, deadline=None)
    @given(st.integers(1, 7), st.integers(0, 5))
    def check(K, L):
        keys, plus, minus = hierarchy_tables(K, L)
        okeys = O.gen_keys(K, L)
        op, om = O.neighbour_tables(okeys, L)
        assert (keys == okeys).all() and (plus == op).all() and (minus == om).all()

    check()
