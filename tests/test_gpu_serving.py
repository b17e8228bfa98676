"""Decode serving loop (idx2g_server) on the GPU: requests of different ROI /
downsampling / tolerance on resident datasets run concurrently under an HBM
budget, batched into shared launches, with decoded coarse levels cached;
every result is bit-identical to the same query decoded alone."""
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
    cnt = server.counters()
    server.close()
    for (ds, q, want), g in zip(jobs, got):
        assert np.array_equal(g.view(np.uint8), want.reshape(-1).view(np.uint8)), q
    assert info["completed"] == len(jobs)
    assert info["peak_bytes"] <= 4 << 30
    assert cnt["batched_requests"] == len(jobs)
    assert cnt["batches"] < len(jobs)        # requests shared launches
    assert all(x > 0 for x in lats)


def test_repeated_key_hits_the_level_cache(resident):
    """A coarse level decoded whole for one request is reused by later
    requests with the same tolerance and downsampling; their results equal
    the same queries decoded alone (without the cache)."""
    from idx2_amd import api
    ctx, sets = resident
    ds, dims = sets[0]  # 2 levels
    qs = [dict(first=(5, 9, 3), last=(100, 77, 50), tolerance=1e-4),
          dict(first=(0, 0, 0), last=(40, 40, 40), tolerance=1e-4),
          dict(first=(60, 30, 20), last=(159, 127, 95), tolerance=1e-4)]
    jobs = [(ds, q, ds.decode(**q)) for q in qs]
    server = api.Server(0, hbm_budget=4 << 30, slots=1)
    got0, _ = _run_mix(server, jobs[:1])   # decodes level 1 whole and keeps it
    got1, _ = _run_mix(server, jobs[1:])
    cnt = server.counters()
    server.close()
    for (_, q, want), g in zip(jobs, got0 + got1):
        assert np.array_equal(g.view(np.uint8), want.reshape(-1).view(np.uint8)), q
    assert cnt["cache_inserts"] >= 1
    assert cnt["cache_hits"] >= 2
    assert cnt["cache_bytes"] > 0


def test_small_requests_pass_a_blocked_head(resident):
    """While the oldest request does not fit next to the one admitted first,
    a later small one that fits is admitted ahead of it; all results are
    unchanged."""
    import torch
    from idx2_amd import api
    ctx, sets = resident
    ds, dims = sets[0]
    big = dict(tolerance=1e-6)
    small = dict(first=(0, 0, 0), last=(15, 15, 15), tolerance=1e-2)
    fb, fs = api.decode_footprint(ds, **big), api.decode_footprint(ds, **small)
    assert fs < fb
    budget = fb + fs  # the first big request and a small one fit, two big ones do not
    jobs = [(ds, big, ds.decode(**big)), (ds, big, ds.decode(**big)), (ds, small, ds.decode(**small))]
    # two slots: while the first request runs (milliseconds), the second
    # slot finds the second big one blocked at the head and the small one
    # fitting behind it (the three submissions take microseconds)
    server = api.Server(0, hbm_budget=budget, slots=2)
    outs, ids = [], []
    for d, q, want in jobs:
        buf = torch.empty(want.size, dtype=torch.float32, device="cuda")
        outs.append(buf)
        ids.append(server.submit(d, buf.data_ptr(), out_owner=buf, **q))
    for i in ids:
        server.wait(i)
    cnt = server.counters()
    server.close()
    torch.cuda.synchronize()
    for (_, q, want), g in zip(jobs, outs):
        assert np.array_equal(g.cpu().numpy().view(np.uint8), want.reshape(-1).view(np.uint8)), q
    assert cnt["bypasses"] >= 1


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
