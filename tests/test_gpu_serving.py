"""Decode serving loop (idx2g_server) on the GPU: requests of different ROI /
downsampling / tolerance on resident datasets run concurrently under an HBM
budget, and every result is bit-identical to the same query decoded alone."""
import numpy as np
import pytest

from idx2_amd import synth

pytestmark = pytest.mark.gpu


def _queries(rng, dims, n):
    qs = []
    for _ in range(n):
        first, last = [], []
        for d in range(3):
            e = int(rng.integers(min(8, dims[d]), dims[d] + 1))
            f = int(rng.integers(0, dims[d] - e + 1))
            first.append(f)
            last.append(f + e - 1)
        ds = tuple(int(x) for x in rng.integers(0, 3, 3))
        tol = float(10 ** rng.uniform(-6, -2))
        qs.append(dict(first=tuple(first), last=tuple(last), downsampling=ds, tolerance=tol))
    return qs


@pytest.fixture(scope="module")
def resident():
    from idx2_amd import api
    ctx = api.Context(0)
    sets = []
    for dims, dt, nl in (((160, 128, 96), "float32", 2), ((96, 80, 64), "float64", 1), ((192, 160, 1), "float64", 3)):
        data = synth.make_field(dims, dt, seed=len(sets) + 21)
        brick = (32, 32, 1) if dims[2] == 1 else (32, 32, 32)
        ds = api.Dataset.encode(ctx, data, name="S", field="f", brick=brick, nlevels=nl, tolerance=1e-6)
        sets.append((ds, dims))
    yield ctx, sets
    for ds, _ in sets:
        ds.free()
    ctx.close()


def _run_mix(server, jobs):
    import torch
    outs, ids = [], []
    for ds, q, want in jobs:
        buf = torch.empty(want.size, dtype=torch.float64 if want.dtype == np.float64 else torch.float32, device="cuda")
        outs.append(buf)
        ids.append(server.submit(ds, buf.data_ptr(), **q))
    lats = [server.wait(i)[0] for i in ids]
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs], lats


def test_mix_equals_alone(resident):
    from idx2_amd import api
    ctx, sets = resident
    rng = np.random.default_rng(5)
    jobs = []
    for k in range(30):
        ds, dims = sets[k % len(sets)]
        q = _queries(rng, dims, 1)[0]
        if dims[2] == 1:
            q["first"] = (q["first"][0], q["first"][1], 0)
            q["last"] = (q["last"][0], q["last"][1], 0)
            q["downsampling"] = (q["downsampling"][0], q["downsampling"][1], 0)
        jobs.append((ds, q, ds.decode(**q)))
    server = api.Server(0, hbm_budget=4 << 30, slots=4)
    got, lats = _run_mix(server, jobs)
    info = server.info()
    server.close()
    for (ds, q, want), g in zip(jobs, got):
        assert np.array_equal(g.view(np.uint8), want.reshape(-1).view(np.uint8)), q
    assert info["completed"] == len(jobs)
    assert info["peak_running"] >= 2          # requests overlapped on the device
    assert info["peak_bytes"] <= 4 << 30
    assert all(x > 0 for x in lats)


def test_budget_serialises(resident):
    """A budget below every request's footprint admits one request at a time
    (a request larger than the budget runs alone), results unchanged."""
    from idx2_amd import api
    ctx, sets = resident
    ds, dims = sets[0]
    rng = np.random.default_rng(9)
    qs = _queries(rng, dims, 6)
    jobs = [(ds, q, ds.decode(**q)) for q in qs]
    server = api.Server(0, hbm_budget=1, slots=3)
    got, _ = _run_mix(server, jobs)
    info = server.info()
    server.close()
    assert info["peak_running"] == 1
    for (_, q, want), g in zip(jobs, got):
        assert np.array_equal(g.view(np.uint8), want.reshape(-1).view(np.uint8)), q


def test_bad_request_is_rejected_at_submit(resident):
    """A query that is empty after clamping is refused by submit with an
    error code; the server keeps serving the next requests."""
    import torch
    from idx2_amd import api
    ctx, sets = resident
    ds, dims = sets[0]
    server = api.Server(0, hbm_budget=1 << 30, slots=2)
    buf = torch.empty(16, dtype=torch.float32, device="cuda")
    with pytest.raises(api.Idx2Error):
        server.submit(ds, buf.data_ptr(), first=(10, 10, 10), last=(5, 5, 5))
    q = dict(first=(0, 0, 0), last=(7, 7, 7))
    want = ds.decode(**q)
    out = torch.empty(want.size, dtype=torch.float32, device="cuda")
    rid = server.submit(ds, out.data_ptr(), **q)
    server.wait(rid)
    server.close()
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint8), want.reshape(-1).view(np.uint8))
