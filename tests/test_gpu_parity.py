"""HIP path parity, called through the C ABI on the GPU.

* encode: files byte-identical with the oracle (same system zstd) and
  section-identical with the reference's golden files (bit-plane chunks,
  addresses, decompressed exponent streams);
* decode: reference-written golden files decode bit-identically to the
  reference's own decode (sha256 of every recorded query);
* round trips and ROI/crop consistency at larger sizes.
"""
import glob
import hashlib
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, golden_cases
from idx2_amd import synth
import idx2file

pytestmark = pytest.mark.gpu
CASES = golden_cases()
IDS = [c[0] for c in CASES]


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def _files(root):
    return sorted(p[len(root):] for p in glob.glob(os.path.join(root, "**", "*.bin"), recursive=True))


@pytest.fixture(scope="module")
def ctx():
    from idx2_amd import api
    c = api.Context(0)
    yield c
    c.close()


def _kw(spec):
    return dict(brick=spec["brick"], nlevels=spec["nlevels"], tolerance=spec["tolerance"],
                bricks_per_chunk=spec["bpc"], chunks_per_file=spec["cpf"], files_per_dir=spec["fpd"])


@pytest.mark.parametrize("name,spec", CASES, ids=IDS)
def test_encode_matches_reference_and_oracle(ctx, oracle_lib, name, spec):
    from idx2_amd import api
    print(f"case {name}", flush=True)
    data = synth.make_field(spec["dims"], spec["dtype"], seed=spec["seed"])
    gold = os.path.join(GOLDEN, name, "G", "f")
    with tempfile.TemporaryDirectory() as wd:
        ds = api.Dataset.encode(ctx, data, name="G", field="f", **_kw(spec))
        ds.write(os.path.join(wd, "gpu"))
        ds.free()
        mine = os.path.join(wd, "gpu", "G", "f")
        assert _files(mine) == _files(gold)
        for rel in _files(gold):
            assert idx2file.same_content(open(gold + rel, "rb").read(), open(mine + rel, "rb").read()), rel
        oracle_lib.use_variant("system")
        assert oracle_lib.encode(os.path.join(wd, "ora"), "G", "f", data, **_kw(spec)) == 0
        ora = os.path.join(wd, "ora", "G", "f")
        for rel in _files(ora):
            assert open(ora + rel, "rb").read() == open(mine + rel, "rb").read(), rel
        assert open(os.path.join(wd, "ora", "G", "f.idx2")).read() == open(os.path.join(wd, "gpu", "G", "f.idx2")).read()


@pytest.mark.parametrize("name,spec", CASES, ids=IDS)
def test_decode_reference_files(ctx, name, spec):
    from idx2_amd import api
    print(f"case {name}", flush=True)
    ds = api.Dataset.open(ctx, os.path.join(GOLDEN, name, "G", "f.idx2"))
    for q in spec["queries"]:
        kw = {k: tuple(v) if isinstance(v, list) else v for k, v in q.items()
              if k in ("first", "last", "downsampling", "tolerance")}
        out = ds.decode(**kw)
        assert out.size == q["size"]
        assert _sha(out) == q["sha256"], q
    ds.free()


def _roundtrip_case(ctx, oracle_lib, dims, dtype, nlevels, tol, brick, queries, seed=11, data=None):
    from idx2_amd import api
    if data is None:
        data = synth.make_field(dims, dtype, seed=seed)
    with tempfile.TemporaryDirectory() as wd:
        ds = api.Dataset.encode(ctx, data, name="R", field="f", brick=brick, nlevels=nlevels, tolerance=tol)
        ds.write(wd)
        oracle_lib.use_variant("system")
        assert oracle_lib.encode(os.path.join(wd, "ora"), "R", "f", data, brick=brick, nlevels=nlevels,
                                 tolerance=tol) == 0
        for rel in _files(os.path.join(wd, "ora", "R", "f")):
            assert open(os.path.join(wd, "ora", "R", "f") + rel, "rb").read() == \
                open(os.path.join(wd, "R", "f") + rel, "rb").read(), rel
        for q in queries:
            g = ds.decode(**q)
            rc, o = oracle_lib.decode(os.path.join(wd, "R", "f.idx2"), wd, **q)
            assert rc == 0 and g.shape == o.shape
            assert np.array_equal(g.view(np.uint8), o.view(np.uint8)), q
        ds.free()
    return data


def test_larger_3d_f32_three_levels(ctx, oracle_lib):
    print("case 256x256x160 f32 3 levels", flush=True)
    _roundtrip_case(ctx, oracle_lib, (256, 256, 160), "float32", 3, 1e-4, (32, 32, 32),
                    [dict(), dict(downsampling=(1, 1, 1)), dict(tolerance=1e-2),
                     dict(first=(17, 40, 3), last=(200, 131, 150), downsampling=(0, 1, 0))])


SHAPES = [((256, 160, 130), "float32", 2, 1e-4, (64, 64, 64)),
          ((200, 150, 100), "float64", 2, 1e-5, (64, 32, 16)),
          ((300, 40, 24), "float32", 1, 1e-3, (128, 8, 8)),
          ((260, 200, 1), "float32", 2, 1e-4, (128, 64, 1)),
          ((260, 260, 130), "float32", 1, 1e-3, (128, 128, 128)),
          ((150, 70, 50), "float64", 2, 1e-5, (32, 16, 8)),
          ((200, 90, 1), "float32", 2, 1e-4, (32, 8, 1)),
          ((96, 100, 40), "float32", 2, 1e-4, (16, 32, 8))]


@pytest.mark.parametrize("dims,dtype,nlevels,tol,brick", SHAPES, ids=[f"{b[0]}x{b[1]}x{b[2]}" for *_, b in SHAPES])
def test_large_and_anisotropic_bricks_match_oracle(ctx, oracle_lib, dims, dtype, nlevels, tol, brick):
    """Bricks above 32 samples per axis (the generic kernels with 64-bit
    block words, per-subband tables in the workgroup's scratch slot when they
    outgrow LDS) and anisotropic bricks, whose axes shrink at different rates
    so that the reference lifts the clamped boundary lines of the coarse
    extrapolation levels more than once: files byte-identical with the
    oracle, queries bit-identical."""
    print(f"case {dims} {dtype} brick {brick} {nlevels} levels", flush=True)
    z = 0 if dims[2] == 1 else 1
    _roundtrip_case(ctx, oracle_lib, dims, dtype, nlevels, tol, brick,
                    [dict(), dict(downsampling=(1, 1, z)), dict(tolerance=1e-2),
                     dict(first=(5, 3, 0 if z == 0 else 2), last=(dims[0] - 9, dims[1] - 4, max(0, dims[2] - 7)),
                          downsampling=(0, 1, 0))])


def test_extreme_magnitudes_match_oracle(ctx, oracle_lib):
    """float64 bricks scaled near the bottom of the exponent range: at 2^-1000
    the dequantisation power k = (62 - ndims) - emax exceeds 1023, so
    Dequantize's scale 1 / 2^k (Function.h:146) divides by inf and is 0, and
    the quantisation scale 2^k is inf (Function.h:123-127); at 2^-960 the scale
    is a small normal power.  Tolerance 1e-310 (subnormal) keeps these blocks'
    planes in the stream.  Files byte-identical with the oracle, decodes
    bit-identical (signs of zeros included)."""
    print("case 96^3 f64 extreme magnitudes", flush=True)
    dims = (96, 96, 96)
    data = synth.make_field(dims, "float64", seed=5)
    data[:, :, :32] *= 2.0 ** -1000
    data[:, :, 32:64] *= 2.0 ** -960
    data[:, :, 64:] *= 2.0 ** -1000
    _roundtrip_case(ctx, oracle_lib, dims, "float64", 2, 1e-310, (32, 32, 32),
                    [dict(), dict(downsampling=(1, 1, 1)), dict(tolerance=1e-300)], data=data)


def test_subband_key_range_limit(ctx):
    """A level-1 brick whose subband spans 2^-1000 .. 2^1000 needs more bit-plane
    keys than a brick-subband's tables hold (128): the encoder refuses it with
    NotSupportedInVersion and a message, and the context stays usable."""
    from idx2_amd import api
    print("case 96^3 f64 key range beyond the tables", flush=True)
    dims = (96, 96, 96)
    data = synth.make_field(dims, "float64", seed=5)
    data[:, :, :32] *= 2.0 ** -1000
    data[:, :, 64:] *= 2.0 ** 1000
    with pytest.raises(api.Idx2Error) as ei:
        api.Dataset.encode(ctx, data, name="R", field="f", brick=(32, 32, 32), nlevels=2, tolerance=1e-310)
    assert ei.value.code == 28 and "bit-plane keys" in str(ei.value)
    ds = api.Dataset.encode(ctx, synth.make_field((64, 64, 64), "float32", seed=1), name="R", field="f",
                            brick=(32, 32, 32), nlevels=1, tolerance=1e-3)
    assert ds.decode().shape == (64, 64, 64)
    ds.free()


def test_2d_f64_slice(ctx, oracle_lib):
    print("case 1024x768x1 f64 3 levels", flush=True)
    _roundtrip_case(ctx, oracle_lib, (1024, 768, 1), "float64", 3, 1e-4, (32, 32, 1),
                    [dict(), dict(downsampling=(2, 2, 0)), dict(first=(100, 37, 0), last=(900, 700, 0))])


EDGE = [
    # slivers: partial bricks of 1 and 8 samples, float64, tight tolerance
    ((33, 40, 65), "float64", 1, 1e-6, (32, 32, 32)),
    # tolerance = machine epsilon of float32 (idx2App's default,
    # Applications/idx2App.cpp:151): every plane down to the last
    ((96, 64, 64), "float32", 1, 2.0 ** -23, (32, 32, 32)),
    # 2D, coarse tolerance
    ((64, 64, 1), "float32", 1, 1e-2, (32, 32, 1)),
    # 2D, full and partial 32x32 bricks on two levels, float32
    ((160, 100, 1), "float32", 2, 1e-5, (32, 32, 1)),
    # partial bricks on two levels
    ((130, 66, 34), "float32", 2, 1e-5, (32, 32, 32)),
    # 16^3 bricks (generic kernels, bricks kept in LDS)
    ((48, 48, 48), "float64", 2, 1e-4, (16, 16, 16)),
]


@pytest.mark.parametrize("dims,dtype,nlevels,tol,brick", EDGE, ids=[f"{d[0]}x{d[1]}x{d[2]}_{t}_{n}l" for d, t, n, _, _ in EDGE])
def test_edge_cases_match_oracle(ctx, oracle_lib, dims, dtype, nlevels, tol, brick):
    print(f"edge case {dims} {dtype} {nlevels} levels tol {tol} brick {brick}", flush=True)
    last = tuple(d - 1 for d in dims)
    corner = tuple(max(0, d - 3) for d in dims)
    flat = dims[2] == 1
    queries = [dict(), dict(first=corner, last=last), dict(first=last, last=last),
               dict(downsampling=(1, 2, 0)), dict(tolerance=1e-1),
               dict(first=(1, 2, 0 if flat else 3), last=(dims[0] - 2, dims[1] - 1, 0 if flat else dims[2] - 2),
                    downsampling=(0, 1, 0 if flat else 1), tolerance=1e-3)]
    _roundtrip_case(ctx, oracle_lib, dims, dtype, nlevels, tol, brick, queries, seed=29)


def test_empty_query_is_an_error(ctx):
    """A query box that is empty after clamping to the volume is rejected, not decoded."""
    from idx2_amd import api
    data = synth.make_field((64, 64, 64), "float32", seed=4)
    ds = api.Dataset.encode(ctx, data, brick=(32, 32, 32), nlevels=1, tolerance=1e-3)
    with pytest.raises(api.Idx2Error):
        ds.decode(first=(10, 10, 10), last=(5, 5, 5))
    with pytest.raises(api.Idx2Error):
        ds.decode(first=(70, 0, 0), last=(80, 10, 10))
    ds.free()


def test_roi_is_crop_of_full_decode(ctx):
    from idx2_amd import api
    print("case roi crop", flush=True)
    data = synth.make_field((160, 128, 96), "float32", seed=2)
    ds = api.Dataset.encode(ctx, data, brick=(32, 32, 32), nlevels=2, tolerance=1e-4)
    full = ds.decode()
    f, l_ = (13, 7, 20), (150, 99, 60)
    roi = ds.decode(first=f, last=l_)
    assert np.array_equal(roi, full[f[2]:l_[2] + 1, f[1]:l_[1] + 1, f[0]:l_[0] + 1])
    ds.free()


def test_reference_decodes_gpu_files(ctx, oracle_lib):
    """The reference idx2App (oracle/_ref) reads files the GPU wrote."""
    from idx2_amd import api
    if not oracle_lib.ref_available():
        pytest.skip("reference build not present")
    print("case reference reads gpu files", flush=True)
    data = synth.make_field((96, 80, 70), "float64", seed=4)
    with tempfile.TemporaryDirectory() as wd:
        ds = api.Dataset.encode(ctx, data, name="X", field="f", brick=(32, 32, 32), nlevels=1, tolerance=1e-5)
        ds.write(wd)
        g = ds.decode(downsampling=(1, 0, 1), tolerance=1e-3)
        rc, path = oracle_lib.ref_decode(wd, os.path.join(wd, "X", "f.idx2"), downsampling=(1, 0, 1),
                                         tolerance=1e-3)
        assert rc == 0
        assert np.array_equal(np.fromfile(path, dtype=np.float64), g.reshape(-1))
        ds.free()


def test_invalid_parameters_return_reference_codes(ctx):
    from idx2_amd import api
    data = np.zeros((64, 64, 64), np.float32)
    cases = [(dict(brick=(24, 32, 32)), 21), (dict(brick=(128, 32, 32)), 22), (dict(nlevels=3), 23),
             (dict(bricks_per_chunk=3000), 34), (dict(chunks_per_file=3), 33), (dict(bricks_per_chunk=65536), 31)]
    for kw, code in cases:
        with pytest.raises(api.Idx2Error) as e:
            api.Dataset.encode(ctx, data, **kw)
        assert e.value.code == code, kw


def test_missing_file_matches_reference(ctx):
    """A dataset with one level-0 file deleted (tests/golden/missing_file.json,
    written by the reference idx2App): it opens, queries that need a brick of
    the missing file fail with FileNotFound as the reference's do, and the
    others decode to the reference's bytes."""
    import json
    from idx2_amd import api
    fx = json.load(open(os.path.join(GOLDEN, "missing_file.json")))
    with tempfile.TemporaryDirectory() as wd:
        shutil.copytree(os.path.join(GOLDEN, fx["source"], "G"), os.path.join(wd, "G"))
        os.remove(os.path.join(wd, "G", "f", fx["removed"]))
        ds = api.Dataset.open(ctx, os.path.join(wd, "G", "f.idx2"))
        for q in fx["queries"]:
            kw = {k: tuple(v) if isinstance(v, list) else v for k, v in q.items()
                  if k in ("first", "last", "downsampling", "tolerance")}
            print(f"  query {kw} -> {q['result']}", flush=True)
            if q["result"] == "ok":
                out = ds.decode(**kw)
                assert out.size == q["size"] and _sha(out) == q["sha256"], q
            else:
                with pytest.raises(api.Idx2Error) as e:
                    ds.decode(**kw)
                assert e.value.code == 37, q  # FileNotFound
        ds.free()


def test_device_synth_matches_numpy(ctx):
    import torch
    from idx2_amd import api  # noqa: F401
    for dims, dt in (((48, 40, 36), "float64"), ((70, 33, 21), "float32")):
        ref = synth.make_field(dims, dt, seed=9)
        buf = torch.empty(ref.size, dtype=torch.float64 if dt == "float64" else torch.float32, device="cuda")
        ctx.synth_field(buf.data_ptr(), dims, dt, 9)
        torch.cuda.synchronize()
        assert np.array_equal(buf.cpu().numpy().view(np.uint8), ref.reshape(-1).view(np.uint8))
