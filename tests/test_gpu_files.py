"""Malformed dataset files are rejected on open with an error code, never
read out of range (the reader checks every size and address it takes from
a file against the file; a stream that passes the checks is still decoded
inside a padded buffer).  The corrupted datasets are opened only, never
decoded.  Reference-written files from tests/golden/ are the starting
point."""
import os
import shutil
import struct
import tempfile

import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from idx2_amd import api
    c = api.Context(0)
    yield c
    c.close()


def _copy_case(wd, case="c2d_f64"):
    dst = os.path.join(wd, "G")  # the dataset name in the metadata is G
    shutil.copytree(os.path.join(GOLDEN, case, "G"), dst)
    return os.path.join(dst, "f.idx2"), os.path.join(dst, "f", "L00.bin")


def _open_fails(ctx, idx2_path):
    from idx2_amd import api
    with pytest.raises(api.Idx2Error) as e:
        api.Dataset.open(ctx, idx2_path)
    return e.value.code


def test_intact_copy_opens(ctx):
    from idx2_amd import api
    with tempfile.TemporaryDirectory() as wd:
        meta, _ = _copy_case(wd)
        ds = api.Dataset.open(ctx, meta)
        ds.free()


@pytest.mark.parametrize("how", ["truncated", "empty", "huge_exponent_section", "huge_chunk_count",
                                 "huge_address_table"])
def test_corrupt_level_file_is_rejected(ctx, how):
    print(f"case {how}", flush=True)
    with tempfile.TemporaryDirectory() as wd:
        meta, lvl = _copy_case(wd)
        b = bytearray(open(lvl, "rb").read())
        if how == "truncated":
            b = b[: len(b) // 2]
        elif how == "empty":
            b = bytearray()
        elif how == "huge_exponent_section":   # trailer: exponent part size
            b[-4:] = struct.pack("<i", 0x7FFFFFF0)
        elif how == "huge_chunk_count":        # trailer: number of exponent chunks
            b[-8:-4] = struct.pack("<i", 0x7FFFFFF0)
        else:                                  # trailer: exponent address table size
            b[-12:-8] = struct.pack("<i", 0x7FFFFFF0)
        open(lvl, "wb").write(bytes(b))
        assert _open_fails(ctx, meta) != 0
