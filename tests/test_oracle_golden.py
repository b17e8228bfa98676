"""The oracle against the reference's own outputs (golden fixtures written by
the reference idx2App, tests/golden/make_fixtures.py)."""
import glob
import hashlib
import os
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, golden_cases
from idx2_amd import synth
import idx2file

CASES = golden_cases()
IDS = [c[0] for c in CASES]


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def _files(root):
    return sorted(p[len(root):] for p in glob.glob(os.path.join(root, "**", "*.bin"), recursive=True))


@pytest.mark.parametrize("name,spec", CASES, ids=IDS)
def test_input_generator_stable(name, spec):
    data = synth.make_field(spec["dims"], spec["dtype"], seed=spec["seed"])
    assert _sha(data) == spec["input_sha256"]


def _encode_with(O, spec, out):
    data = synth.make_field(spec["dims"], spec["dtype"], seed=spec["seed"])
    rc = O.encode(out, "G", "f", data, brick=spec["brick"], nlevels=spec["nlevels"], tolerance=spec["tolerance"],
                  bricks_per_chunk=spec["bpc"], chunks_per_file=spec["cpf"], files_per_dir=spec["fpd"])
    assert rc == 0


@pytest.mark.parametrize("name,spec", CASES, ids=IDS)
def test_oracle_encode_matches_reference(oracle_lib, name, spec):
    """Same chunks, addresses and exponent streams as the reference; with the
    reference's own zstd (oracle/_ref build) the files are byte-identical."""
    O = oracle_lib
    gold = os.path.join(GOLDEN, name, "G", "f")
    with tempfile.TemporaryDirectory() as wd:
        O.use_variant("system")
        try:
            _encode_with(O, spec, wd)
        finally:
            O.use_variant("system")
        mine = os.path.join(wd, "G", "f")
        assert _files(mine) == _files(gold)
        for rel in _files(gold):
            a = open(gold + rel, "rb").read()
            b = open(mine + rel, "rb").read()
            assert idx2file.same_content(a, b), rel
        meta_g = open(os.path.join(GOLDEN, name, "G", "f.idx2")).read()
        meta_m = open(os.path.join(wd, "G", "f.idx2")).read()
        head = meta_g.split("---------------- Statistics")[0]
        assert meta_m.split("---------------- Statistics")[0] == head


@pytest.mark.parametrize("name,spec", CASES, ids=IDS)
def test_oracle_encode_byte_identical_with_reference_zstd(oracle_lib, name, spec):
    O = oracle_lib
    if not os.path.exists(O.LIB_Z145_PATH):
        pytest.skip("reference zstd build (oracle/_ref) not present")
    gold = os.path.join(GOLDEN, name, "G", "f")
    with tempfile.TemporaryDirectory() as wd:
        O.use_variant("z145")
        try:
            _encode_with(O, spec, wd)
        finally:
            O.use_variant("system")
        mine = os.path.join(wd, "G", "f")
        for rel in _files(gold):
            assert open(gold + rel, "rb").read() == open(mine + rel, "rb").read(), rel
        assert open(os.path.join(GOLDEN, name, "G", "f.idx2")).read() == open(os.path.join(wd, "G", "f.idx2")).read()


@pytest.mark.parametrize("name,spec", CASES, ids=IDS)
def test_oracle_decode_matches_reference(oracle_lib, name, spec):
    O = oracle_lib
    idx = os.path.join(GOLDEN, name, "G", "f.idx2")
    for q in spec["queries"]:
        kw = {k: tuple(v) if isinstance(v, list) else v for k, v in q.items()
              if k in ("first", "last", "downsampling", "tolerance")}
        rc, out = O.decode(idx, os.path.join(GOLDEN, name), **kw)
        assert rc == 0
        assert out.size == q["size"]
        assert _sha(out) == q["sha256"], q
