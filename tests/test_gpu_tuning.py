"""Launch-shape tuning (idx2g_set_tuning / idx2g_autotune) changes only time:
chunk bytes and decoded grids are identical under every shape."""
import numpy as np
import pytest

from idx2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from idx2_amd import api
    c = api.Context(0)
    yield c
    api.set_tuning(0)
    c.close()


CASES = [((160, 128, 96), "float32", 2, (32, 32, 32)), ((320, 192, 1), "float64", 2, (32, 32, 1))]


def _run(ctx, dims, dt, nl, brick):
    from idx2_amd import api
    data = synth.make_field(dims, dt, seed=41)
    ds = api.Dataset.encode(ctx, data, brick=brick, nlevels=nl, tolerance=1e-5)
    out = (ds.chunk_bytes(), ds.decode(), ds.decode(first=(5, 7, 0), last=(150, 100, 0 if dims[2] == 1 else 90),
                                                   downsampling=(1, 0, 0)))
    ds.free()
    return out


@pytest.mark.parametrize("shape", [dict(enc_wg_per_cu=1, dec_wg_per_cu=1, walk_batch=7, walk_wide_max=1),
                                   dict(enc_wg_per_cu=2, dec_wg_per_cu=4, walk_batch=16, walk_wide_max=8),
                                   dict(enc_wg_per_cu=8, dec_wg_per_cu=8, walk_batch=4096, walk_wide_max=1 << 30)])
def test_shapes_do_not_change_results(ctx, shape):
    from idx2_amd import api
    for dims, dt, nl, brick in CASES:
        api.set_tuning(0)
        want = _run(ctx, dims, dt, nl, brick)
        api.set_tuning(0, **shape)
        got = _run(ctx, dims, dt, nl, brick)
        api.set_tuning(0)
        for w, g in zip(want, got):
            assert np.array_equal(w.view(np.uint8), g.view(np.uint8)), (dims, shape)


def test_autotune_picks_valid_shapes(ctx):
    from idx2_amd import api
    shapes, best, dflt = api.autotune(ctx)
    assert 0 <= shapes["enc_wg_per_cu"] <= 8 and 0 <= shapes["dec_wg_per_cu"] <= 8
    assert shapes["walk_batch"] in (0, 2048, 4096)
    assert shapes["walk_wide_max"] in (0, 1023, 8191, 1 << 30)
    assert 0 < best <= dflt
    assert api.get_tuning(0) == shapes
    api.set_tuning(0)


def test_release_device_cache_keeps_results(ctx):
    """idx2g_release_device_cache hands the kept free blocks back (a
    non-negative byte count; an invalid device is an error), and the calls
    after it allocate again and give the same bytes and grids."""
    from idx2_amd import api
    print("case release_device_cache", flush=True)
    dims, dt, nl, brick = CASES[0]
    a = _run(ctx, dims, dt, nl, brick)
    assert api.release_device_cache(0) > 0  # the encode/decode freed their work buffers
    assert api.release_device_cache(0) == 0
    with pytest.raises(api.Idx2Error):
        api.release_device_cache(1 << 20)
    b = _run(ctx, dims, dt, nl, brick)
    for w, g in zip(a, b):
        assert np.array_equal(w.view(np.uint8), g.view(np.uint8))
