"""Sparse-brick output (the reference's HashMap mode, idx2App --brick_map)
through the C ABI on the GPU.

Fixtures (tests/golden/brick_map/brick_map.json, written by the reference
idx2App built from its sources, tests/golden/make_fixtures.py): for
full-volume decodes of the golden datasets and of a field that is zero on a
slab (so some level-0 bricks are not significant and are written from their
level-1 parents), the sha256 of the reference's "bricks" file
(idx2SparseBricks.cpp:207-243) and its per-level significant-brick counts
(PrintStatistics, idx2SparseBricks.cpp:49-65).
"""
import hashlib
import json
import os
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN

FX = json.load(open(os.path.join(GOLDEN, "brick_map", "brick_map.json")))


def _sparse_field():
    from idx2_amd import synth
    s = FX["sparse"]
    data = synth.make_field(tuple(s["dims"]), s["dtype"], seed=s["seed"])
    x0, x1 = s["zero_x"]
    data[:, :, x0:x1] = 0
    return data


def test_sparse_fixture_input_regenerates():
    data = _sparse_field()
    assert hashlib.sha256(np.ascontiguousarray(data).view(np.uint8).tobytes()).hexdigest() == FX["sparse"]["input_sha256"]


@pytest.fixture(scope="module")
def ctx():
    pytest.importorskip("torch")
    from idx2_amd import api
    c = api.Context(0)
    yield c
    c.close()


def _dataset(ctx, source):
    from idx2_amd import api
    d = os.path.join(GOLDEN, "brick_map" if source == "sparse" else source)
    return api.Dataset.open(ctx, os.path.join(d, "G", "f.idx2"), in_dir=d)


def _query(rec):
    return {k: tuple(rec[k]) if isinstance(rec[k], list) else rec[k] for k in ("downsampling", "tolerance") if k in rec}


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(len(FX["decodes"])),
                         ids=[f"{r['source']}-{'-'.join(f'{k}{v}' for k, v in _query(r).items()) or 'full'}"
                              for r in FX["decodes"]])
def test_brick_file_matches_reference(ctx, i):
    rec = FX["decodes"][i]
    print(f"brick map {rec['source']} {_query(rec)}", flush=True)
    ds = _dataset(ctx, rec["source"])
    pool = ds.decode_bricks(**_query(rec))
    try:
        for level, kept, total in rec["counts"]:
            assert pool.count(level) == (kept, total), level
        with tempfile.TemporaryDirectory() as wd:
            path = os.path.join(wd, "bricks")
            pool.write(path)
            blob = open(path, "rb").read()
        assert len(blob) == rec["size"]
        assert hashlib.sha256(blob).hexdigest() == rec["sha256"]
    finally:
        pool.free()
        ds.free()


@pytest.mark.gpu
def test_kept_bricks_match_the_grid_decode_and_device_lookup(ctx):
    """Every kept level-0 brick of a full-resolution decode holds the grid
    decode's samples on its brick extent (the extra extended samples are the
    brick's extrapolation); the device hash table returns the same addresses
    as the host lookup, and NULL for bricks that are not kept."""
    import torch
    ds = _dataset(ctx, "sparse")
    grid = ds.decode()
    pool = ds.decode_bricks()
    try:
        nz, ny, nx = grid.shape
        keys, want = [], []
        for bz in range((nz + 31) // 32):
            for by in range((ny + 31) // 32):
                for bx in range((nx + 31) // 32):
                    for level in (0, 1):
                        b3 = (bx, by, bz) if level == 0 else (bx // 2, by // 2, bz // 2)
                        found = pool.find(level, b3)
                        keys.append(pool.key(level, b3))
                        want.append(found[0] if found else 0)
                        if level != 0 or not found:
                            continue
                        assert found[1] == (33, 33, 33)
                        v = pool.samples(0, b3)
                        ez, ey, ex = min(32, nz - 32 * bz), min(32, ny - 32 * by), min(32, nx - 32 * bx)
                        crop = grid[32 * bz:32 * bz + ez, 32 * by:32 * by + ey, 32 * bx:32 * bx + ex]
                        assert np.array_equal(v[:ez, :ey, :ex].astype(np.float32), crop), (bx, by, bz)
        assert sum(1 for w in want if w == 0) > 0  # the zero slab's level-0 bricks are not kept
        k = torch.tensor(np.array(keys, dtype=np.uint64).view(np.int64), device="cuda")
        out = torch.zeros(len(keys), dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        pool.lookup(ctx, len(keys), k.data_ptr(), out.data_ptr())
        got = out.cpu().numpy().view(np.uint64).tolist()
        assert got == want
        # resolution: level 0 where the level-0 brick is kept, else its parent's level
        for bx in range((nx + 31) // 32):
            r = pool.resolution((bx, 0, 0))
            assert r == (0 if pool.find(0, (bx, 0, 0)) else 1)
    finally:
        pool.free()
        ds.free()
