"""The C-ABI library loads on any host and exports every entry point that
include/idx2g.h declares (no compute calls here)."""
import ctypes
import os
import re

from conftest import ROOT


def test_library_exports_header_symbols():
    from idx2_amd import api
    header = open(os.path.join(ROOT, "include", "idx2g.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(idx2g_\w+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(api.LIB_PATH)
    for sym in sorted(declared):
        assert hasattr(lib, sym), sym
    assert declared == set(api.EXPORTS)


def test_error_message_without_device():
    from idx2_amd import api
    msg = api.lib().idx2g_last_error_message()
    assert isinstance(msg, bytes)


def test_synth_params_deterministic():
    from idx2_amd import synth
    a = synth.field_params(3)
    b = synth.field_params(3)
    assert (a == b).all() and a.shape == (192,)
    # integer wavevectors in [-8, 8], not all zero
    for j in range(8):
        k = a[160 + 4 * j:160 + 4 * j + 3]
        assert all(float(v).is_integer() and -8 <= v <= 8 for v in k) and any(k)
