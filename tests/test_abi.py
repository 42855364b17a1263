"""CPU checks of the C ABI: the library loads, exports every function that
include/eolc.h declares, and reports errors instead of aborting."""
import ctypes as C
import os
import re

import eolc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "eolc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(eolc_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(eolc.EXPORTED_SYMBOLS)


def test_library_exports_all_symbols():
    L = C.CDLL(eolc.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_null_context_is_an_error_not_an_abort():
    L = eolc.lib()
    assert L.eolc_set_options(None, 0) == eolc.ERR_INVALID
    assert L.eolc_last_error(None) == b"null context"
    rc = L.eolc_create(0, None)
    assert rc == eolc.ERR_INVALID


def test_create_reports_status():
    h = C.c_void_p()
    rc = eolc.lib().eolc_create(0, C.byref(h))
    assert rc in (eolc.OK, eolc.ERR_NO_DEVICE)
    if rc == eolc.OK:
        eolc.lib().eolc_destroy(h)
