"""The C-ABI library: built, loadable, and exporting exactly what include/qca.h
declares (no GPU needed; no compute calls)."""
import ctypes
import os
import re

from conftest import ROOT

import qca_native as qn

HEADER = os.path.join(ROOT, "include", "qca.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:int|const char\*)\s+(qca_\w+)\s*\(", text, flags=re.M))


def test_library_is_built_and_loads():
    assert os.path.exists(qn.LIB_PATH)
    assert qn.lib().qca_version() >= 10000


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(qn.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    assert declared() == set(qn.SIGNATURES)


def test_argument_errors_are_reported_without_a_device():
    lib = qn.lib()
    assert lib.qca_engine_step(None, 1) == -2
    assert b"null engine" in lib.qca_last_error()
    assert lib.qca_heff_apply(3, 0, None, None, None, None, None, 1, 1, 1, 1, None) == -2
    assert b"precision" in lib.qca_last_error()
    assert lib.qca_exact_expm(None, 10, 1, 1, 2, 0.1, 0.0, None, None) == -2
    failed = ctypes.c_int(7)
    assert lib.qca_batch_step(None, 2, 1, 4, ctypes.byref(failed)) == -2
    assert b"null engine list" in lib.qca_last_error()
    hs = (ctypes.c_void_p * 1)(None)
    assert lib.qca_batch_step(hs, 1, 1, 0, None) == -2
    assert b"nthreads" in lib.qca_last_error()
    assert lib.qca_batch_step(hs, 1, 1, 2, None) == -2
    assert b"null engine in batch" in lib.qca_last_error()
    assert lib.qca_batch_step(None, 0, 5, 1, ctypes.byref(failed)) == 0
    assert failed.value == -1
    assert lib.qca_engine_set_fused_krylov_limit(None, 1024) == -2
    assert b"null engine" in lib.qca_last_error()
