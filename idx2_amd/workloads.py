"""Workload definitions shared by the benchmark and the GPU tests: the 64-field
batched mix, the progressive decode sweep and the serving loop's request
stream.  Only shapes, parameters and seeded random draws live here; the
field values come from the seeded generator (idx2_amd/synth.py and its device
twin idx2g_synth_field).
"""
import numpy as np

# 64 fields of uneven shapes: dims per axis from BATCH_DIMS, 2D and 3D 50/50,
# float32 and float64 50/50, 2-4 levels clamped to each field's maximum,
# tolerance 1e-4, field f generated with seed + f.
BATCH_DIMS = (64, 96, 128, 192, 256, 384, 512)


def batched_specs(seed, n_fields=64):
    from idx2_amd import api
    rng = np.random.default_rng(seed + 3000)
    out = []
    for f in range(n_fields):
        is2d = bool(rng.random() < 0.5)
        dims = [int(rng.choice(BATCH_DIMS)) for _ in range(3)]
        if is2d:
            dims[2] = 1
        dt = "float64" if rng.random() < 0.5 else "float32"
        brick = (32, 32, 1) if is2d else (32, 32, 32)
        nl = max(1, min(int(rng.integers(2, 5)), api.max_levels(dims, brick)))
        out.append(dict(dims=tuple(dims), dtype=dt, brick=brick, nlevels=nl, tolerance=1e-4,
                        name=f"F{f}", field="f", seed=seed + f))
    return out


# Progressive decode sweep: 2048x2048x1024 float32 encoded once, decoded at
# three downsamplings x four tolerances over the full volume and random boxes.
PROGRESSIVE = dict(dims=(2048, 2048, 1024), dtype="float32", nlevels=3, tolerance=1e-6, brick=(32, 32, 32))
PROGRESSIVE_DOWNSAMPLING = ((2, 2, 2), (1, 1, 1), (0, 0, 0))
PROGRESSIVE_TOLERANCES = (1e-2, 1e-3, 1e-4, 1e-6)


def progressive_boxes(seed, dims, n_roi=32, lo=64, hi=512):
    """n_roi seeded random boxes (first, last inclusive) of edge lo..hi."""
    rng = np.random.default_rng(seed + 11)
    boxes = []
    for _ in range(n_roi):
        first, last = [], []
        for d in range(3):
            e = int(rng.integers(lo, hi + 1))
            f = int(rng.integers(0, dims[d] - e + 1))
            first.append(f)
            last.append(f + e - 1)
        boxes.append((tuple(first), tuple(last)))
    return boxes


# Serving loop: resident datasets, Poisson arrivals at a fixed rate, random
# ROI (edge 32-1024, clipped), downsampling U{0,1,2} per axis, tolerance
# log-uniform in [1e-6, 1e-2], HBM admission budget 200 GB.
SERVE_SETS = [((512, 512, 512), "float64"), ((1024, 1024, 1024), "float32"), ((2048, 2048, 512), "float32")]
SERVE_ENCODE = dict(brick=(32, 32, 32), nlevels=3, tolerance=1e-6)


def serve_requests(seed, n_requests, rate, sets=SERVE_SETS):
    """[(dataset index, query dict, gap to the previous arrival in s)]."""
    rng = np.random.default_rng(seed + 7)
    reqs = []
    for _ in range(n_requests):
        k = int(rng.integers(0, len(sets)))
        dims = sets[k][0]
        first, last = [], []
        for d in range(3):
            e = min(int(rng.integers(32, 1025)), dims[d])
            f = int(rng.integers(0, dims[d] - e + 1))
            first.append(f)
            last.append(f + e - 1)
        q = dict(first=tuple(first), last=tuple(last), downsampling=tuple(int(x) for x in rng.integers(0, 3, 3)),
                 tolerance=float(10 ** rng.uniform(-6, -2)))
        reqs.append((k, q, float(rng.exponential(1.0 / rate))))
    return reqs
