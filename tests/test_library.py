"""The C-ABI library loads on any host and exports every entry point that
include/idx2g.h declares (no compute calls here)."""
import ctypes
import os
import re

from conftest import ROOT


def test_library_exports_header_symbols():
    from idx2_amd import api
    header = open(os.path.join(ROOT, "include", "idx2g.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(idx2g_\w+)\s*\(", header, re.M))
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


def test_brick_extents_up_to_128_validate_like_the_reference():
    """Generic block words hold block origins in 8 bits per axis and the
    generic kernels' tables 17^3 blocks per subband, so brick extents up to
    128 are taken (also anisotropic ones such as 64x16x16, whose subbands
    have few blocks but origins past 31); an extent of 256 is refused with
    NotSupportedInVersion after the reference's own checks pass, and
    parameters the reference rejects keep the reference's code."""
    from idx2_amd import api
    code = {name: c for c, name in api.ERRORS.items()}
    assert api.validate_params(dims=(128, 128, 128), dtype="float32", brick=(32, 32, 32), nlevels=2) == 0
    for brick in [(64, 16, 16), (16, 64, 16), (16, 16, 64), (64, 64, 64), (128, 8, 8), (128, 128, 128)]:
        assert api.validate_params(dims=(256, 256, 256), dtype="float32", brick=brick, nlevels=1) == 0, brick
    assert api.validate_params(dims=(512, 64, 64), dtype="float32", brick=(256, 16, 16), nlevels=1) == \
        code["NotSupportedInVersion"]
    assert api.validate_params(dims=(128, 128, 128), dtype="float32", brick=(24, 16, 16), nlevels=1) == \
        code["BrickSizeNotPowerOfTwo"]
