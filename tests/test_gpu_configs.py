"""The workload configurations at their own sizes, through the C ABI on the GPU.

* 384x384x256 float64, brick 32^3, 2 levels, tolerance 1e-3 (the CPU
  reference's configuration): every L*.bin and the .idx2 text byte-identical
  to the oracle's, and ROI / downsampling / tolerance queries decoding
  bit-identically.  This runs the on-chip float64 kernels below the coarsest
  level (k_encode_b32<1>, k_walk, k_decode_b32<1>) with parents passed
  through the dense level grid.
* 1024x1024x256 float32, 3 levels, tolerance 1e-4: a slab of the 1024^3
  workload (same brick, levels and tolerance, a quarter of its z extent; the
  single-threaded oracle encodes the whole 1024^3 grid in about two minutes,
  past a test's time budget), chunk files and full decode against the oracle.
* the 64-field batched mix: every field's files and decode identical batched
  and alone, and 8 fields (2D/3D x float32/float64) against the oracle.
* 2048x2048x1024 float32 encoded once: at every downsampling x tolerance of
  the progressive sweep, each of the 32 random boxes decodes to the crop of
  the full decode.
* the serving loop on its three full-size datasets with 200 requests of the
  serving request stream: each result of the concurrent mix equals the same
  query decoded alone.
"""
import glob
import os
import tempfile

import numpy as np
import pytest

from idx2_amd import workloads

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from idx2_amd import api
    c = api.Context(0)
    yield c
    c.close()


def _tdt(dt):
    return torch.float64 if dt == "float64" else torch.float32


def _field(ctx, dims, dt, seed):
    t = torch.empty(dims[0] * dims[1] * dims[2], dtype=_tdt(dt), device="cuda")
    ctx.synth_field(t.data_ptr(), dims, dt, seed)
    torch.cuda.synchronize()
    return t


def _files(root):
    return sorted(p[len(root):] for p in glob.glob(os.path.join(root, "**", "*.bin"), recursive=True))


def _same_files(a, b):
    fa, fb = _files(a), _files(b)
    assert fa == fb
    for rel in fa:
        assert open(a + rel, "rb").read() == open(b + rel, "rb").read(), rel


def _decode_dev(ds, dt, **q):
    dims, _ = ds.output_dims(**q)
    out = torch.zeros(dims[0] * dims[1] * dims[2], dtype=_tdt(dt), device="cuda")
    ds.decode(out_dev_ptr=out.data_ptr(), **q)
    torch.cuda.synchronize()
    return out.view(dims[2], dims[1], dims[0])


def _against_oracle(ctx, oracle_lib, wd, x, dims, dt, enc, queries):
    """Encode the device grid x, write it, encode its host copy with the
    oracle; files, .idx2 and every query's decode must be identical."""
    from idx2_amd import api
    host = x.cpu().numpy().reshape(dims[2], dims[1], dims[0])
    ds = api.Dataset.encode(ctx, dev_ptr=x.data_ptr(), dims=dims, dtype=dt, name="C", field="f", **enc)
    ds.write(os.path.join(wd, "gpu"))
    oracle_lib.use_variant("system")
    assert oracle_lib.encode(os.path.join(wd, "ora"), "C", "f", host, **enc) == 0
    _same_files(os.path.join(wd, "ora", "C", "f"), os.path.join(wd, "gpu", "C", "f"))
    assert open(os.path.join(wd, "ora", "C", "f.idx2")).read() == open(os.path.join(wd, "gpu", "C", "f.idx2")).read()
    for q in queries:
        print(f"  query {q}", flush=True)
        g = ds.decode(**q)
        rc, o = oracle_lib.decode(os.path.join(wd, "gpu", "C", "f.idx2"), os.path.join(wd, "gpu"), **q)
        assert rc == 0 and g.shape == o.shape, q
        assert np.array_equal(g.view(np.uint8), o.view(np.uint8)), q
    ds.free()


def test_cpu_reference_config_384x384x256_f64(ctx, oracle_lib):
    dims, dt = (384, 384, 256), "float64"
    print(f"case {dims} {dt} 2 levels tol 1e-3", flush=True)
    x = _field(ctx, dims, dt, seed=0)
    enc = dict(brick=(32, 32, 32), nlevels=2, tolerance=1e-3)
    queries = [dict(), dict(downsampling=(1, 1, 1)), dict(tolerance=1e-1),
               dict(first=(37, 5, 100), last=(300, 383, 255)),
               dict(first=(0, 64, 31), last=(383, 200, 97), downsampling=(0, 1, 1), tolerance=1e-2)]
    with tempfile.TemporaryDirectory() as wd:
        _against_oracle(ctx, oracle_lib, wd, x, dims, dt, enc, queries)


@pytest.mark.timeout(400)
def test_plain_workload_slab_1024x1024x256_f32(ctx, oracle_lib):
    dims, dt = (1024, 1024, 256), "float32"
    print(f"case {dims} {dt} 3 levels tol 1e-4", flush=True)
    x = _field(ctx, dims, dt, seed=0)
    enc = dict(brick=(32, 32, 32), nlevels=3, tolerance=1e-4)
    with tempfile.TemporaryDirectory() as wd:
        _against_oracle(ctx, oracle_lib, wd, x, dims, dt, enc, [dict()])


@pytest.mark.timeout(400)
def test_batched_mix_64_fields(ctx, oracle_lib):
    from idx2_amd import api
    specs = workloads.batched_specs(seed=0)
    print(f"case 64-field mix ({sum(1 for s in specs if s['dims'][2] == 1)} 2D)", flush=True)
    xs = [_field(ctx, s["dims"], s["dtype"], s["seed"]) for s in specs]
    enc = [dict(dev_ptr=x.data_ptr(), dims=s["dims"], dtype=s["dtype"], brick=s["brick"], nlevels=s["nlevels"],
                tolerance=s["tolerance"], name=s["name"], field=s["field"]) for s, x in zip(specs, xs)]
    batch = api.Dataset.encode_batch(ctx, enc)
    with tempfile.TemporaryDirectory() as wd:
        for i, (kw, db) in enumerate(zip(enc, batch)):
            alone = api.Dataset.encode(ctx, **kw)
            db.write(os.path.join(wd, "b%d" % i))
            alone.write(os.path.join(wd, "a%d" % i))
            alone.free()
            _same_files(os.path.join(wd, "a%d" % i, kw["name"], "f"), os.path.join(wd, "b%d" % i, kw["name"], "f"))
        outs = [torch.zeros_like(x) for x in xs]
        api.decode_batch(ctx, batch, [o.data_ptr() for o in outs])
        torch.cuda.synchronize()
        for i, (s, ds, o) in enumerate(zip(specs, batch, outs)):
            alone = _decode_dev(ds, s["dtype"]).reshape(-1)
            assert torch.equal(alone.view(torch.uint8), o.view(torch.uint8)), i
        # 8 fields against the oracle: the two smallest of each of 2D/3D x float32/float64
        picks = []
        for is2d in (True, False):
            for dt in ("float32", "float64"):
                cand = sorted((int(np.prod(s["dims"])), i) for i, s in enumerate(specs)
                              if (s["dims"][2] == 1) == is2d and s["dtype"] == dt)
                picks += [i for _, i in cand[:2]]
        assert len(picks) == 8
        oracle_lib.use_variant("system")
        for i in picks:
            s = specs[i]
            print(f"  oracle field {i} {s['dims']} {s['dtype']} {s['nlevels']} levels", flush=True)
            host = xs[i].cpu().numpy().reshape(s["dims"][2], s["dims"][1], s["dims"][0])
            assert oracle_lib.encode(os.path.join(wd, "o%d" % i), s["name"], "f", host, brick=s["brick"],
                                     nlevels=s["nlevels"], tolerance=s["tolerance"]) == 0
            _same_files(os.path.join(wd, "o%d" % i, s["name"], "f"), os.path.join(wd, "b%d" % i, s["name"], "f"))
            rc, o = oracle_lib.decode(os.path.join(wd, "b%d" % i, s["name"], "f.idx2"), os.path.join(wd, "b%d" % i))
            assert rc == 0
            assert np.array_equal(o.reshape(-1).view(np.uint8), outs[i].cpu().numpy().view(np.uint8)), i
    for ds in batch:
        ds.free()


def _crop(full, first, last, dsmp):
    """The samples of full (a decode at downsampling dsmp) that an ROI decode
    of [first, last] writes: its output grid starts at first rounded down to
    the stride and the comparison stops at `last` (a sample past it can lie
    in a brick the ROI traversal skips)."""
    st = [1 << v for v in dsmp]
    f2 = [first[d] // st[d] for d in range(3)]
    n = [(last[d] - f2[d] * st[d]) // st[d] + 1 for d in range(3)]
    return f2, n, full[f2[2]:f2[2] + n[2], f2[1]:f2[1] + n[1], f2[0]:f2[0] + n[0]]


@pytest.mark.timeout(400)
def test_progressive_sweep_rois_are_crops(ctx):
    from idx2_amd import api
    c = workloads.PROGRESSIVE
    dims = c["dims"]
    print(f"case progressive {dims}", flush=True)
    x = _field(ctx, dims, "float32", seed=0)
    ds = api.Dataset.encode(ctx, dev_ptr=x.data_ptr(), dims=dims, dtype="float32", name="p", field="f",
                            brick=c["brick"], nlevels=c["nlevels"], tolerance=c["tolerance"])
    del x
    boxes = workloads.progressive_boxes(0, dims)
    for dsmp in workloads.PROGRESSIVE_DOWNSAMPLING:
        for tol in workloads.PROGRESSIVE_TOLERANCES:
            print(f"  downsampling {dsmp} tolerance {tol}", flush=True)
            full = _decode_dev(ds, "float32", downsampling=dsmp, tolerance=tol)
            for first, last in boxes:
                r = _decode_dev(ds, "float32", first=first, last=last, downsampling=dsmp, tolerance=tol)
                _, n, want = _crop(full, first, last, dsmp)
                assert torch.equal(r[:n[2], :n[1], :n[0]], want), (dsmp, tol, first, last)
            del full
    ds.free()


@pytest.mark.timeout(400)
def test_serving_full_size_mix_equals_alone(ctx):
    from idx2_amd import api
    sets = []
    for i, (dims, dt) in enumerate(workloads.SERVE_SETS):
        print(f"case serving dataset {dims} {dt}", flush=True)
        x = _field(ctx, dims, dt, seed=100 + i)
        sets.append((api.Dataset.encode(ctx, dev_ptr=x.data_ptr(), dims=dims, dtype=dt, name="s%d" % i, field="f",
                                        **workloads.SERVE_ENCODE), dt))
        del x
    reqs = workloads.serve_requests(0, 200, 400.0)
    alone = []
    for k, q, _ in reqs:
        ds, dt = sets[k]
        alone.append(_decode_dev(ds, dt, **q).reshape(-1))
    print("  alone done", flush=True)
    server = api.Server(0, hbm_budget=200 << 30, slots=4)
    outs, ids = [], []
    for (k, q, _), want in zip(reqs, alone):
        ds, dt = sets[k]
        o = torch.empty_like(want)
        outs.append(o)
        ids.append(server.submit(ds, o.data_ptr(), out_owner=o, **q))
    for i in ids:
        server.wait(i)
    info = server.info()
    server.close()
    torch.cuda.synchronize()
    bad = [i for i, (o, w) in enumerate(zip(outs, alone)) if not torch.equal(o.view(torch.uint8), w.view(torch.uint8))]
    assert not bad, bad[:10]
    assert info["completed"] == len(reqs)
    assert info["peak_running"] >= 2
    for ds, _ in sets:
        ds.free()
