"""Batched encode/decode (idx2g_encode_batch / idx2g_decode_batch) on the GPU:
a mix of fields of uneven shapes (3D and 2D, float32 and float64, full and
partial bricks, 1-3 levels) encoded in one pass gives the same chunk bytes as
encoding each field alone, and a batch of queries decodes to the same grids
as each query alone."""
import numpy as np
import pytest

from idx2_amd import synth

pytestmark = pytest.mark.gpu

MIX = [
    ((64, 96, 128), "float32", 2),
    ((128, 64, 64), "float64", 2),
    ((96, 96, 96), "float32", 1),
    ((100, 70, 45), "float64", 2),
    ((160, 128, 1), "float64", 3),
    ((256, 96, 1), "float32", 2),
    ((64, 64, 64), "float64", 1),
    ((192, 64, 96), "float32", 3),
]


def _fields():
    from idx2_amd import api
    out = []
    for i, (dims, dt, nl) in enumerate(MIX):
        brick = (32, 32, 1) if dims[2] == 1 else (32, 32, 32)
        nl = min(nl, api.max_levels(dims, brick))
        data = synth.make_field(dims, dt, seed=100 + i)
        out.append(dict(data=data, name=f"F{i}", field="f", brick=brick, nlevels=nl, tolerance=1e-4 if i % 2 else 1e-6))
    return out


@pytest.fixture(scope="module")
def batch():
    from idx2_amd import api
    ctx = api.Context(0)
    fields = _fields()
    dss = api.Dataset.encode_batch(ctx, fields)
    yield ctx, fields, dss
    for ds in dss:
        ds.free()
    ctx.close()


def test_encode_batch_equals_single(batch):
    from idx2_amd import api
    ctx, fields, dss = batch
    for f, ds in zip(fields, dss):
        alone = api.Dataset.encode(ctx, **f)
        try:
            assert np.array_equal(ds.chunk_bytes(), alone.chunk_bytes()), f["name"]
            assert ds.info() == alone.info(), f["name"]
            assert np.array_equal(ds.decode().view(np.uint8), alone.decode().view(np.uint8)), f["name"]
        finally:
            alone.free()


def test_decode_batch_equals_single(batch):
    import torch
    from idx2_amd import api
    ctx, fields, dss = batch
    rng = np.random.default_rng(3)
    items = []
    for rep in range(3):
        for f, ds in zip(fields, dss):
            nz, ny, nx = f["data"].shape
            dims = (nx, ny, nz)
            first, last = [], []
            for d in range(3):
                e = int(rng.integers(1, dims[d] + 1))
                s0 = int(rng.integers(0, dims[d] - e + 1))
                first.append(s0)
                last.append(s0 + e - 1)
            ds3 = [int(x) for x in rng.integers(0, 3, 3)]
            if nz == 1:
                ds3[2] = 0
            q = dict(first=tuple(first), last=tuple(last), downsampling=tuple(ds3),
                     tolerance=float(10 ** rng.uniform(-6, -2)))
            if rep == 0:
                q = dict(downsampling=(0, 0, 0), tolerance=0.0)
            items.append((ds, q))
    outs = []
    for ds, q in items:
        dims, nbytes = ds.output_dims(**q)
        outs.append(torch.zeros(max(nbytes, 1), dtype=torch.uint8, device="cuda"))
    api.decode_batch(ctx, [ds for ds, _ in items], [o.data_ptr() for o in outs], [q for _, q in items])
    torch.cuda.synchronize()
    for (ds, q), o in zip(items, outs):
        want = ds.decode(**q)
        got = o.cpu().numpy()[: want.nbytes]
        assert np.array_equal(got, want.reshape(-1).view(np.uint8)), q


def test_batch_of_one_and_repeat(batch):
    """A batch of one field is the single path; a repeated batch is deterministic."""
    from idx2_amd import api
    ctx, fields, dss = batch
    again = api.Dataset.encode_batch(ctx, fields[:3])
    try:
        for a, b in zip(again, dss[:3]):
            assert np.array_equal(a.chunk_bytes(), b.chunk_bytes())
    finally:
        for a in again:
            a.free()
