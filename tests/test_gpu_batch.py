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


def _files_of(ds, root):
    import os
    ds.write(root)
    out = {}
    for dirpath, _, names in os.walk(root):
        for nm in names:
            p = os.path.join(dirpath, nm)
            out[os.path.relpath(p, root)] = open(p, "rb").read()
    return out


def _same_batches(ctx, fields, a, b):
    import tempfile
    with tempfile.TemporaryDirectory() as wd:
        for i, (x, y) in enumerate(zip(a, b)):
            fx = _files_of(x, f"{wd}/a{i}")
            fy = _files_of(y, f"{wd}/b{i}")
            assert fx == fy, fields[i]["name"]
            assert np.array_equal(x.decode().view(np.uint8), y.decode().view(np.uint8)), fields[i]["name"]


def test_batch_graph_equals_stepwise(batch, monkeypatch):
    """The batch above ran as one graph launch; the same batch encoded
    stepwise (one launch group per level, host-read key ranges and chunk
    layout) writes the same files."""
    from idx2_amd import api
    ctx, fields, dss = batch
    monkeypatch.setenv("IDX2G_BATCH_GRAPH", "0")
    step = api.Dataset.encode_batch(ctx, fields)
    monkeypatch.delenv("IDX2G_BATCH_GRAPH")
    try:
        _same_batches(ctx, fields, dss, step)
    finally:
        for ds in step:
            ds.free()


def test_graph_falls_back_when_a_level_spans_too_many_keys(batch, monkeypatch):
    """A one-level float64 field with values near 1e-5 in its bricks of
    x < 32 and near 1e20 in those of x >= 32, at tolerance 1e-20, spans about
    140 bit-plane keys on its level, more than the graph form's rows hold: the
    batch is encoded again stepwise and matches the stepwise batch."""
    from idx2_amd import api
    ctx, _, _ = batch
    data = synth.make_field((64, 64, 64), "float64", seed=77)
    data[:, :, :32] *= 1e-5
    data[:, :, 32:] *= 1e20
    fields = [dict(data=data, name="W", field="f", brick=(32, 32, 32), nlevels=1, tolerance=1e-20),
              dict(data=synth.make_field((96, 80, 70), "float32", seed=78), name="N", field="f",
                   nlevels=api.max_levels((96, 80, 70), (32, 32, 32)), tolerance=1e-4)]
    graph = api.Dataset.encode_batch(ctx, fields)
    monkeypatch.setenv("IDX2G_BATCH_GRAPH", "0")
    step = api.Dataset.encode_batch(ctx, fields)
    monkeypatch.delenv("IDX2G_BATCH_GRAPH")
    try:
        _same_batches(ctx, fields, graph, step)
    finally:
        for ds in graph + step:
            ds.free()


def test_arena_regrowth_keeps_results(batch, monkeypatch):
    """Encodes that start from a 1 MiB arena overflow it, grow it and encode
    again (stepwise for one field, as a graph for the batch): the same files
    and decodes as without the overflow."""
    from idx2_amd import api
    ctx, fields, dss = batch
    monkeypatch.setenv("IDX2G_TEST_ARENA_BYTES", str(1 << 20))
    small = api.Dataset.encode_batch(ctx, fields[:3])
    alone = api.Dataset.encode(ctx, **fields[0])
    monkeypatch.delenv("IDX2G_TEST_ARENA_BYTES")
    try:
        _same_batches(ctx, fields[:3], small, dss[:3])
        _same_batches(ctx, fields[:1], [alone], dss[:1])
    finally:
        for ds in small + [alone]:
            ds.free()
