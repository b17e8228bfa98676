"""Benchmark of the idx2 encode/decode path on one MI355X.

A step is one encode of the whole grid (input resident in HBM -> chunks
assembled in HBM) followed by one full-resolution decode (chunks in HBM ->
grid in HBM).  value = raw input bytes x steps / wall time of the timed
steps (GB/s).  Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "encode & decode GB/s (raw input) per GPU + % HBM roofline; p50/p99 decode latency"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # the single-GPU plain path of the workload list: 1024^3 float32, 3 levels, tol 1e-4
    "c2": dict(dims=(1024, 1024, 1024), dtype="float32", nlevels=3, tolerance=1e-4, brick=(32, 32, 32),
               workload="3D float32 1024^3 (4 GiB), brick 32^3, 3 levels, tolerance 1e-4, encode + full decode"),
    # the 2D float64 slice of the same workload: 8192^2, brick 32x32x1
    "c2b": dict(dims=(8192, 8192, 1), dtype="float64", nlevels=3, tolerance=1e-4, brick=(32, 32, 1),
                workload="2D float64 8192^2 (512 MiB), brick 32x32x1, 3 levels, tolerance 1e-4"),
    # the reference's CPU case: 384x384x256 float64, 2 levels, tol 1e-3
    "c1": dict(dims=(384, 384, 256), dtype="float64", nlevels=2, tolerance=1e-3, brick=(32, 32, 32),
               workload="3D float64 384x384x256, brick 32^3, 2 levels, tolerance 1e-3"),
}
# The CPU baseline runs the reference's own CPU configuration (the 384x384x256
# float64 case, 2 levels, tolerance 1e-3) whatever GPU workload is measured:
# one core for encode and serial decode, and the reference's thread-pool
# decoder on at most CPU_SHARE cores (a one-GPU job's share of the host).
CPU_SAMPLE = CONFIGS["c1"]
CPU_SHARE = 16


def log(msg):
    """Progress on stderr (stdout carries only the JSON result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(ctx, torch, seed):
    """The reference idx2App (built from the reference sources into
    oracle/_ref) on its CPU configuration: encode + serial full decode pinned
    to one core (the value), and the --parallel decode pinned to CPU_SHARE
    cores.  The raw input file is written before the clock starts; file
    output of the encode and decode is included (it is part of idx2App)."""
    from oracle import oracle as O  # checker/baseline only
    if not O.ref_available():
        return None
    s = CPU_SAMPLE
    nx, ny, nz = s["dims"]
    tdt = torch.float64 if s["dtype"] == "float64" else torch.float32
    buf = torch.empty(nx * ny * nz, dtype=tdt, device="cuda")
    ctx.synth_field(buf.data_ptr(), s["dims"], s["dtype"], seed)
    torch.cuda.synchronize()
    host = buf.cpu().numpy().reshape(nz, ny, nx)
    raw_bytes = host.nbytes
    avail = sorted(os.sched_getaffinity(0))
    one, share = avail[:1], avail[:min(CPU_SHARE, len(avail))]
    with tempfile.TemporaryDirectory() as wd:
        raw = O.ref_write_raw(wd, "B", "f", host)
        out_dir = os.path.join(wd, "ref")
        t0 = time.perf_counter()
        rc = O.ref_encode_raw(raw, out_dir, brick=s["brick"], nlevels=s["nlevels"], tolerance=s["tolerance"], cpus=one)
        t1 = time.perf_counter()
        if rc:
            return None
        meta = os.path.join(out_dir, "B", "f.idx2")
        rc, _ = O.ref_decode(wd, meta, cpus=one)
        t2 = time.perf_counter()
        rc2, _ = O.ref_decode(wd, meta, parallel=True, cpus=share)
        t3 = time.perf_counter()
        if rc or rc2:
            return None
    return {"value": round(raw_bytes / (t2 - t0) / 1e9, 6), "unit": "GB/s", "cores": 1, "kind": "reference",
            "sample": f"{s['workload']}: reference idx2App encode {t1 - t0:.2f}s + serial decode {t2 - t1:.2f}s "
                      f"on one core (file output included, raw input written before timing)",
            "parallel_decode": {"seconds": round(t3 - t2, 3), "gbps": round(raw_bytes / (t3 - t2) / 1e9, 6),
                                "cores": len(share),
                                "note": "idx2App --decode --parallel (thread pool of hardware_concurrency() threads) "
                                        "pinned to the listed number of cores"}}


# FP64 operations of the reference's transform of one full 32^3 brick,
# counted from its loops (FLiftCdf53 / ILiftCdf53, Wavelet.h:177-310): a
# line of D points has P = len(range(1, D - 2, 2)) interior predicts (3 ops:
# add, halve, subtract) and as many updates (4 ops: two quarter-and-add),
# plus for odd D one more predict (3) and update (4), for even D the
# extrapolated point 2B - A (2).  ExtrapolateCdf53 (Wavelet.cpp:67-77) lifts
# X, Y, Z of the 32-point data (X over 32x32 lines, Y over 33x32, Z over
# 33x33) and the 17^3, 9^3, 5^3, 3^3 subgrids forward, then the same
# subgrids and 33^3 inverse; the subband transform lifts 33^3 once more
# (Wavelet.cpp:80-138) and weighs all but subband 0 (17^3).
def fp64_ops_full_brick():
    def line(d):
        n = len(range(1, d - 2, 2))
        return 3 * n + 4 * n + (7 if d % 2 else 2)
    ext0 = 32 * 32 * line(32) + 33 * 32 * line(32) + 33 * 33 * line(32)
    sub = sum(3 * n * n * line(n) for n in (17, 9, 5, 3))
    full33 = 3 * 33 * 33 * line(33)
    weights = 33 ** 3 - 17 ** 3
    return ext0 + 2 * sub + full33 + full33 + weights


def profile_counter(kernel, key):
    """A per-launch figure (e.g. the rocprofv3 VALU instruction count) from the
    committed profile summary of the dominant kernel; None if absent."""
    name = kernel.split()[0].split("::")[-1]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{name}_1024cube_f32.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return int(line.split(":")[1])
    return None


COPY_MIB = 1024  # calibration copy of the HBM counters


def pmc_probe(torch, cfg, seed):
    """Child of measure_traffic, run under rocprofv3 --pmc: one plain device
    copy of COPY_MIB MiB (the counters' calibration) and one encode of the
    workload (its level-0 on-chip launch is the dominant kernel)."""
    from idx2_amd import api
    n = COPY_MIB << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    ctx = api.Context(0)
    nx, ny, nz = cfg["dims"]
    tdt = torch.float64 if cfg["dtype"] == "float64" else torch.float32
    inp = torch.empty(nx * ny * nz, dtype=tdt, device="cuda")
    ctx.synth_field(inp.data_ptr(), cfg["dims"], cfg["dtype"], seed)
    torch.cuda.synchronize()
    ds = api.Dataset.encode(ctx, dev_ptr=inp.data_ptr(), name="probe", field="f", dims=cfg["dims"],
                            dtype=cfg["dtype"], brick=cfg["brick"], nlevels=cfg["nlevels"],
                            tolerance=cfg["tolerance"])
    torch.cuda.synchronize()
    ds.free()
    ctx.close()


def measure_traffic(kernel, config):
    """HBM bytes of the dominant launch, measured in this run: two
    `rocprofv3 --pmc` passes (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch)
    over a child process (pmc_probe), each scaled by what the same counter
    reads for the child's plain copy of COPY_MIB MiB.  Returns a dict, or
    None when the profiler is not available or a pass fails."""
    import csv
    import shutil
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    short = kernel.split()[0].split("::")[-1]
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE", "SQ_INSTS_VALU"):
        with tempfile.TemporaryDirectory(dir="/tmp") as wd:
            cmd = [prof, "--pmc", counter, "-d", wd, "-o", "run", "-f", "csv", "--", sys.executable,
                   os.path.abspath(__file__), "--pmc-probe", "--config", config]
            try:
                r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=150)
            except subprocess.TimeoutExpired:
                log(f"traffic: rocprofv3 --pmc {counter} timed out")
                return None
            rows = []
            for root, _, files in os.walk(wd):
                for f in files:
                    if f.endswith("counter_collection.csv"):
                        rows += list(csv.DictReader(open(os.path.join(root, f))))
            if r.returncode != 0 and not rows:
                log(f"traffic: rocprofv3 --pmc {counter} failed (rc {r.returncode})")
                return None
            copy, launches = None, []
            for row in rows:
                if row.get("Counter_Name") != counter:
                    continue
                name, val = row.get("Kernel_Name", ""), float(row.get("Counter_Value", 0))
                if "copyBuffer" in name or "copy_kernel" in name or "CatArrayBatchedCopy" in name:
                    copy = val if copy is None else max(copy, val)
                elif short in name:
                    launches.append(val)
            if not launches or (copy is None and counter != "SQ_INSTS_VALU"):
                log(f"traffic: {counter}: no copy or no {short} dispatch in the profile")
                return None
            if counter == "SQ_INSTS_VALU":
                out[counter] = {"wave_instructions": int(max(launches))}
                continue
            factor = copy * 1024 / (COPY_MIB << 20)  # counter bytes per copied byte
            out[counter] = {"kib": max(launches), "copy_kib": copy, "factor": round(factor, 4),
                            "bytes": max(launches) * 1024 / factor}
    return {"traffic": int(out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]),
            "valu": out["SQ_INSTS_VALU"]["wave_instructions"],
            "counters": {k: out[k] for k in ("FETCH_SIZE", "WRITE_SIZE")}}


def _serve_run(torch, sets, reqs, rate, slots, budget, max_out):
    """Poisson arrivals of reqs at `rate`; the client holds at most max_out
    bytes of output buffers (an overloaded server makes it wait: stalls)."""
    import threading
    from idx2_amd import api
    server = api.Server(0, hbm_budget=budget, slots=slots)
    lat, qms, out_bytes = [], [], 0
    pending = []
    lock = threading.Condition()
    done_flag = [False]

    def waiter():
        nonlocal out_bytes
        while True:
            with lock:
                while not pending and not done_flag[0]:
                    lock.wait()
                if not pending:
                    return
                rid, buf, nb = pending.pop(0)
            l_, q_, _ = server.wait(rid)
            lat.append(l_)
            qms.append(q_)
            out_bytes += nb  # this request's output grid
            del buf

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    stalls = 0
    try:
        for i, (k, q, gap) in enumerate(reqs):
            if i and i % 2000 == 0:
                log(f"serving: {i} of {len(reqs)} requests submitted")
            ds, _, dt = sets[k]
            t_next += gap * 400.0 / rate  # the stream's gaps are drawn at 400/s
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            dims, nbytes = ds.output_dims(**q)
            while True:
                with lock:
                    held = sum(b.numel() * b.element_size() for _, b, _ in pending)
                if held + nbytes <= max_out or held == 0:
                    break
                stalls += 1
                time.sleep(0.001)
            released = False
            while True:
                try:
                    buf = torch.empty(max(1, nbytes // (8 if dt == "float64" else 4)),
                                      dtype=torch.float64 if dt == "float64" else torch.float32, device="cuda")
                    break
                except torch.OutOfMemoryError:
                    # device memory is short (admitted requests, cached levels, the
                    # library's kept blocks): wait for a pending result to free its
                    # buffer, or once nothing is pending hand the kept blocks back
                    stalls += 1
                    with lock:
                        busy = bool(pending)
                    if busy:
                        time.sleep(0.001)
                    elif released:
                        raise
                    else:
                        api.release_device_cache(0)
                        torch.cuda.empty_cache()
                        released = True
            rid = server.submit(ds, buf.data_ptr(), out_owner=buf, **q)
            with lock:
                pending.append((rid, buf, nbytes))
                lock.notify()
    finally:
        with lock:
            done_flag[0] = True
            lock.notify()
        th.join()
    wall = time.perf_counter() - t0
    info = server.info()
    cnt = server.counters()
    server.close()
    lat = np.array(lat)
    return {"requests": len(reqs), "rate_per_s": rate,
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_queue_ms": round(float(np.mean(qms)), 3), "served_per_s": round(len(reqs) / wall, 2),
            "output_gbps": round(out_bytes / wall / 1e9, 3),
            "peak_concurrent_batches": info["peak_running"], "peak_admitted_bytes": info["peak_bytes"],
            "batches": cnt["batches"], "mean_batch": round(cnt["batched_requests"] / max(1, cnt["batches"]), 2),
            "bypasses": cnt["bypasses"], "cache_hits": cnt["cache_hits"], "cache_misses": cnt["cache_misses"],
            "cache_bytes": cnt["cache_bytes"], "client_stalls": stalls}


def serve_bench(ctx, torch, n_requests, rate, seed, slots=4, budget=200 << 30, max_out=32 << 30,
                sweep=(800.0, 1600.0, 3200.0, 6400.0), sweep_requests=2000):
    """The serving loop on its three resident datasets: n_requests of the
    seeded request stream at `rate`, then a sweep of higher rates
    (sweep_requests each) up to saturation."""
    from idx2_amd import api, workloads
    sets = []
    for i, (dims, dt) in enumerate(workloads.SERVE_SETS):
        n = dims[0] * dims[1] * dims[2]
        buf = torch.empty(n, dtype=torch.float64 if dt == "float64" else torch.float32, device="cuda")
        ctx.synth_field(buf.data_ptr(), dims, dt, seed + 100 + i)
        ds = api.Dataset.encode(ctx, dev_ptr=buf.data_ptr(), dims=dims, dtype=dt, name="s%d" % i, field="f",
                                **workloads.SERVE_ENCODE)
        del buf
        sets.append((ds, dims, dt))
    torch.cuda.synchronize()
    reqs = workloads.serve_requests(seed, n_requests, rate)
    res = _serve_run(torch, sets, reqs, rate, slots, budget, max_out)
    res.update({"slots": slots, "hbm_budget_bytes": budget,
                "datasets": [list(d) + [t] for d, t in workloads.SERVE_SETS]})
    if sweep:
        res["sweep"] = []
        for r in sweep:
            log(f"serving sweep: {sweep_requests} requests at {r}/s")
            res["sweep"].append(_serve_run(torch, sets, workloads.serve_requests(seed + 1, sweep_requests, 400.0), r,
                                           slots, budget, max_out))
    for ds, _, _ in sets:
        ds.free()
    return res


def batched_fields(torch, ctx, seed, n_fields=64):
    """The 64-field mix (idx2_amd.workloads.batched_specs) generated on the device."""
    from idx2_amd import workloads
    fields, bufs = [], []
    for sp in workloads.batched_specs(seed, n_fields):
        dims, dt = sp["dims"], sp["dtype"]
        buf = torch.empty(dims[0] * dims[1] * dims[2], dtype=torch.float64 if dt == "float64" else torch.float32,
                          device="cuda")
        ctx.synth_field(buf.data_ptr(), dims, dt, sp["seed"])
        bufs.append(buf)
        fields.append(dict(dev_ptr=buf.data_ptr(), dims=dims, dtype=dt, brick=sp["brick"], nlevels=sp["nlevels"],
                           tolerance=sp["tolerance"], name=sp["name"], field=sp["field"]))
    return fields, bufs


def batched_bench(ctx, torch, seed, reps=3):
    """Encode + full decode of the 64-field mix: one call per field (the plain
    path) against one idx2g_encode_batch + one idx2g_decode_batch; the batch
    encode as one graph launch (the default) and stepwise
    (IDX2G_BATCH_GRAPH=0: one launch group per level, key ranges and chunk
    layout read back by the host)."""
    from idx2_amd import api
    fields, bufs = batched_fields(torch, ctx, seed)
    modes = ("per_field", "batched_stepwise", "batched")
    outs_by_mode = {m: [torch.empty_like(b) for b in bufs] for m in modes}
    raw = sum(b.numel() * b.element_size() for b in bufs)
    res = {}
    for mode in modes:
        outs = outs_by_mode[mode]
        enc, dec = [], []
        for rep in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if mode == "per_field":
                dss = [api.Dataset.encode(ctx, **f) for f in fields]
            elif mode == "batched_stepwise":
                os.environ["IDX2G_BATCH_GRAPH"] = "0"
                dss = api.Dataset.encode_batch(ctx, fields)
                del os.environ["IDX2G_BATCH_GRAPH"]
            else:
                dss = api.Dataset.encode_batch(ctx, fields)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if mode == "per_field":
                for ds, o in zip(dss, outs):
                    ds.decode(out_dev_ptr=o.data_ptr())
            else:
                api.decode_batch(ctx, dss, [o.data_ptr() for o in outs])
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if rep:  # first round warms the device memory cache
                enc.append(t1 - t0)
                dec.append(t2 - t1)
            if mode == "batched" and rep == reps:
                chunk = sum(ds.info()["chunk_bytes"] for ds in dss)
            for ds in dss:
                ds.free()
        e, d = float(np.mean(enc)), float(np.mean(dec))
        res[mode] = {"encode_ms": round(e * 1e3, 3), "decode_ms": round(d * 1e3, 3),
                     "gbps": round(raw / (e + d) / 1e9, 3)}
    res["speedup"] = round(res["batched"]["gbps"] / res["per_field"]["gbps"], 3)
    res["graph_encode_speedup"] = round(res["batched_stepwise"]["encode_ms"] / res["batched"]["encode_ms"], 3)
    # every batched decode must be bit-identical to the same field's per-field decode
    mismatched = sum(0 if torch.equal(a.view(torch.uint8), b.view(torch.uint8)) else 1
                     for m in ("batched_stepwise", "batched")
                     for a, b in zip(outs_by_mode["per_field"], outs_by_mode[m]))
    res.update({"fields": len(fields), "raw_bytes": int(raw), "chunk_bytes": int(chunk),
                "fields_2d": sum(1 for f in fields if f["dims"][2] == 1),
                "fields_f64": sum(1 for f in fields if f["dtype"] == "float64"),
                "levels": [f["nlevels"] for f in fields], "batch_vs_per_field_mismatches": mismatched})
    del bufs, outs_by_mode
    return res


def progressive_bench(ctx, torch, seed, n_roi=32):
    """Encode 2048x2048x1024 float32 once; decode at downsampling (2,2,2),
    (1,1,1), (0,0,0) x tolerance {1e-2, 1e-3, 1e-4, 1e-6}, full volume and
    n_roi seeded random boxes (edge 64-512)."""
    from idx2_amd import api, workloads
    c = workloads.PROGRESSIVE
    dims = c["dims"]
    buf = torch.empty(dims[0] * dims[1] * dims[2], dtype=torch.float32, device="cuda")
    ctx.synth_field(buf.data_ptr(), dims, "float32", seed)
    ds = api.Dataset.encode(ctx, dev_ptr=buf.data_ptr(), dims=dims, dtype="float32", name="p", field="f",
                            brick=c["brick"], nlevels=c["nlevels"], tolerance=c["tolerance"])
    enc_ms = ctx.stats()["total_ms"]
    del buf
    boxes = [None] + workloads.progressive_boxes(seed, dims, n_roi)
    rows = []
    out = torch.empty(dims[0] * dims[1] * dims[2], dtype=torch.float32, device="cuda")
    for dsmp in ((2, 2, 2), (1, 1, 1), (0, 0, 0)):
        for tol in (1e-2, 1e-3, 1e-4, 1e-6):
            for bx in boxes:
                q = dict(downsampling=dsmp, tolerance=tol)
                if bx is not None:
                    q.update(first=bx[0], last=bx[1])
                ds.decode(out_dev_ptr=out.data_ptr(), **q)
                st = ctx.stats()
                rows.append((dsmp, tol, bx is None, st["total_ms"], st["raw_bytes"], st["chunk_bytes"]))
    ds.free()
    summary = []
    for dsmp in ((2, 2, 2), (1, 1, 1), (0, 0, 0)):
        for tol in (1e-2, 1e-3, 1e-4, 1e-6):
            full = [r for r in rows if r[0] == dsmp and r[1] == tol and r[2]][0]
            rois = [r for r in rows if r[0] == dsmp and r[1] == tol and not r[2]]
            roi_ms = np.array([r[3] for r in rois])
            summary.append({"downsampling": list(dsmp), "tolerance": tol,
                            "full_ms": round(full[3], 3), "full_out_gbps": round(full[4] / full[3] / 1e6, 2),
                            "full_stream_bytes_read": int(full[5]),
                            "roi_p50_ms": round(float(np.percentile(roi_ms, 50)), 3),
                            "roi_out_gbps": round(sum(r[4] for r in rois) / roi_ms.sum() / 1e6, 2),
                            "roi_stream_bytes_read": int(sum(r[5] for r in rois))})
    return {"encode_ms": round(enc_ms, 2), "queries": len(rows), "sweep": summary}


def host_path_bench(ctx, torch, cfg, seed, reps=2):
    """The same encode + decode with the grid and the output in (pageable) host
    memory: the rate a caller that hands over host buffers sees (PCIe copies
    included)."""
    from idx2_amd import api
    nx, ny, nz = cfg["dims"]
    dt = cfg["dtype"]
    tdt = torch.float64 if dt == "float64" else torch.float32
    buf = torch.empty(nx * ny * nz, dtype=tdt, device="cuda")
    ctx.synth_field(buf.data_ptr(), cfg["dims"], dt, seed)
    host = buf.cpu().numpy().reshape(nz, ny, nx)
    del buf
    raw = host.nbytes
    kw = dict(brick=cfg["brick"], nlevels=cfg["nlevels"], tolerance=cfg["tolerance"])
    enc, dec = [], []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        ds = api.Dataset.encode(ctx, host, name="h", field="f", **kw)
        t1 = time.perf_counter()
        out = ds.decode()
        t2 = time.perf_counter()
        ds.free()
        enc.append(t1 - t0)
        dec.append(t2 - t1)
    e, d = float(np.mean(enc[1:])), float(np.mean(dec[1:]))
    ok = bool(np.abs(out.astype(np.float64) - host.astype(np.float64)).max() < 1.0)
    return {"encode_ms": round(e * 1e3, 2), "decode_ms": round(d * 1e3, 2), "gbps": round(raw / (e + d) / 1e9, 3),
            "note": "host input / host output, pageable memory, PCIe copies included", "decoded": ok}


def plain_bench(ctx, torch, cfg, steps, warmup, seed):
    """Encode + full decode of one grid resident in HBM, `steps` timed rounds."""
    from idx2_amd import api
    nx, ny, nz = cfg["dims"]
    dt = cfg["dtype"]
    tdt = torch.float64 if dt == "float64" else torch.float32
    inp = torch.empty(nx * ny * nz, dtype=tdt, device="cuda")
    ctx.synth_field(inp.data_ptr(), cfg["dims"], dt, seed)
    out = torch.empty_like(inp)
    raw = inp.numel() * inp.element_size()
    kw = dict(dims=cfg["dims"], dtype=dt, brick=cfg["brick"], nlevels=cfg["nlevels"], tolerance=cfg["tolerance"])

    def step():
        ds = api.Dataset.encode(ctx, dev_ptr=inp.data_ptr(), name="bench", field="f", **kw)
        es = ctx.stats()
        ds.decode(out_dev_ptr=out.data_ptr())
        dsx = ctx.stats()
        chunk_bytes = ds.info()["chunk_bytes"]
        ds.free()
        return es, dsx, chunk_bytes

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    enc, dec = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        es, dsx, chunk_bytes = step()
        enc.append(es)
        dec.append(dsx)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # error check of the last round trip (size-independent property: |x - x'| small)
    err = float((out - inp).abs().max().item())
    del inp, out
    enc_ms = np.array([e["total_ms"] for e in enc])
    dec_ms = np.array([d["total_ms"] for d in dec])
    # the level-0 brick launches: encode's brick kernel, decode's walk + on-chip pair
    brick = tuple(cfg["brick"])
    kind = "b32" if brick == (32, 32, 32) and nz > 1 else "2d" if brick == (32, 32, 1) and nz == 1 else None
    dom_e = float(np.mean([e["dominant_ms"] for e in enc]))
    dom_d = float(np.mean([d["dominant_ms"] for d in dec]))
    enc_name = f"idx2g::k_encode_{kind} (level 0)" if kind else "idx2g::k_encode_level (level 0)"
    dec_name = f"idx2g::k_walk + idx2g::k_decode_{kind} (level 0)" if kind else "idx2g::k_decode_level (level 0)"
    eb = float(np.mean([e["dominant_bytes"] for e in enc]))
    db = float(np.mean([d["dominant_bytes"] for d in dec]))
    return {"value": raw * steps / wall / 1e9, "wall": wall, "enc_ms": float(enc_ms.mean()),
            "dec_ms": float(dec_ms.mean()), "enc_gbs": raw / (enc_ms.mean() / 1e3) / 1e9,
            "dec_gbs": raw / (dec_ms.mean() / 1e3) / 1e9, "dec_p50": float(np.percentile(dec_ms, 50)),
            "dec_p99": float(np.percentile(dec_ms, 99)), "chunk_bytes": chunk_bytes, "err": err,
            "dom_name": enc_name, "dom_ms": dom_e, "dom_bytes": eb, "dom_launches": 1,
            "achieved": eb / (dom_e / 1e3) / 1e9,
            "dec_name": dec_name, "dec_dom_ms": dom_d, "dec_dom_bytes": db,
            "dec_launches": int(np.mean([d["dominant_launches"] for d in dec])),
            "dec_achieved": db / (dom_d / 1e3) / 1e9, "raw": raw}


def shape_sweep(ctx, torch, seed, steps=3):
    """Device times of every tunable launch shape (idx2g_tuning), one
    coordinate at a time from the defaults: the 1024^3 float32 workload and
    the 2D float64 8192^2 slice (encode and full decode, library device
    times, mean of `steps` after one warmup) and the 64-field mix (one batch
    encode + one batch decode, wall time, mean of `steps` after one warmup)."""
    from idx2_amd import api
    cands = ([{}] + [{"enc_wg_per_cu": v} for v in (1, 2, 4)] + [{"dec_wg_per_cu": v} for v in (1, 2, 4)]
             + [{"walk_batch": v} for v in (2048, 8192)] + [{"walk_wide_max": v} for v in (1, 1023, 32767, 1 << 30)])
    fields, bufs = batched_fields(torch, ctx, seed)
    outs = [torch.empty_like(b) for b in bufs]
    rows = []
    for c in cands:
        log(f"shape sweep: {c or 'defaults'}")
        api.set_tuning(0, **c)
        m3 = plain_bench(ctx, torch, CONFIGS["c2"], steps, 1, seed)
        m2 = plain_bench(ctx, torch, CONFIGS["c2b"], steps, 1, seed)
        mix = []
        for rep in range(steps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dss = api.Dataset.encode_batch(ctx, fields)
            api.decode_batch(ctx, dss, [o.data_ptr() for o in outs])
            torch.cuda.synchronize()
            if rep:
                mix.append(time.perf_counter() - t0)
            for ds in dss:
                ds.free()
        rows.append({"shape": c or "defaults", "c2_encode_ms": round(m3["enc_ms"], 3),
                     "c2_decode_ms": round(m3["dec_ms"], 3), "c2b_encode_ms": round(m2["enc_ms"], 3),
                     "c2b_decode_ms": round(m2["dec_ms"], 3), "mix_ms": round(float(np.mean(mix)) * 1e3, 3)})
    api.set_tuning(0)
    del bufs, outs
    return rows


def end_to_end(ctx, torch, cfg, seed, reps=2):
    """Encode of the HBM-resident grid followed by idx2g_dataset_write (host
    zstd of the exponent and address sections, L*.bin and .idx2 files into a
    tmpfs directory when the host has one): what the reference's Encode +
    file flush delivers."""
    from idx2_amd import api
    nx, ny, nz = cfg["dims"]
    tdt = torch.float64 if cfg["dtype"] == "float64" else torch.float32
    inp = torch.empty(nx * ny * nz, dtype=tdt, device="cuda")
    ctx.synth_field(inp.data_ptr(), cfg["dims"], cfg["dtype"], seed)
    torch.cuda.synchronize()
    raw = inp.numel() * inp.element_size()
    base = "/dev/shm" if os.path.isdir("/dev/shm") else None
    enc_s, wr_s, files = [], [], 0
    for _ in range(reps):
        with tempfile.TemporaryDirectory(dir=base) as wd:
            t0 = time.perf_counter()
            ds = api.Dataset.encode(ctx, dev_ptr=inp.data_ptr(), dims=cfg["dims"], dtype=cfg["dtype"], name="e2e",
                                    field="f", brick=cfg["brick"], nlevels=cfg["nlevels"], tolerance=cfg["tolerance"])
            t1 = time.perf_counter()
            ds.write(wd)
            t2 = time.perf_counter()
            ds.free()
            files = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(wd) for f in fs)
        enc_s.append(t1 - t0)
        wr_s.append(t2 - t1)
    del inp
    e, w = min(enc_s), min(wr_s)
    return {"encode_ms": round(e * 1e3, 2), "write_ms": round(w * 1e3, 2), "gbps": round(raw / (e + w) / 1e9, 3),
            "file_bytes": int(files), "dir": base or tempfile.gettempdir(),
            "note": "encode of the HBM-resident grid + idx2g_dataset_write (zstd sections and files), best of "
                    f"{reps}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=10000, help="serving-loop requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=400.0, help="serving-loop arrival rate (requests/s)")
    ap.add_argument("--no-progressive", dest="progressive", action="store_false",
                    help="skip the progressive decode sweep")
    ap.add_argument("--no-batched", dest="batched", action="store_false",
                    help="skip the 64-field batched-vs-per-field comparison")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shapes instead of tuning them first")
    ap.add_argument("--no-host-path", dest="host_path", action="store_false",
                    help="skip the host-buffer (PCIe-inclusive) encode/decode")
    ap.add_argument("--no-slice", dest="slice2d", action="store_false",
                    help="skip the 2D float64 8192^2 slice")
    ap.add_argument("--no-end-to-end", dest="end_to_end", action="store_false",
                    help="skip the encode + file write timing")
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="skip the two rocprofv3 --pmc passes that measure the dominant launch's HBM traffic")
    ap.add_argument("--pmc-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--shape-sweep", action="store_true",
                    help="only time every tunable launch shape on the 1024^3, 8192^2 and 64-field workloads")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one plain device copy of MIB MiB (HBM counter calibration under rocprofv3) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch
    if args.calibrate_copy > 0:
        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n}))
        return
    if args.pmc_probe:
        pmc_probe(torch, CONFIGS[args.config], args.seed)
        return
    from idx2_amd import api

    ctx = api.Context(0)
    if args.shape_sweep:
        print(json.dumps({"shape_sweep": shape_sweep(ctx, torch, args.seed)}))
        ctx.close()
        return
    cfg = CONFIGS[args.config]
    tuning = None
    if args.autotune:
        log("autotune")
        t0 = time.perf_counter()
        shapes, best, dflt = api.autotune(ctx)
        tuning = {"shapes": shapes, "probe_ms_tuned": round(best, 3), "probe_ms_default": round(dflt, 3),
                  "tuning_s": round(time.perf_counter() - t0, 2)}
    log(f"plain path: {cfg['workload']}")
    m = plain_bench(ctx, torch, cfg, args.steps, args.warmup, args.seed)
    measured = None
    if args.traffic:
        log("HBM traffic of the dominant launch (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)")
        measured = measure_traffic(m["dom_name"], args.config)
    if measured is not None:
        traffic = measured["traffic"]
        traffic_source = ("measured in this run: rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE over one encode of "
                          "the workload, each scaled by the same counter over a plain %d MiB device copy" % COPY_MIB)
    else:
        traffic = profile_counter(m["dom_name"], "traffic_bytes_per_launch") if args.config == "c2" else None
        traffic_source = ("profiles/k_encode_b32_1024cube_f32.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the "
                          "same launch, copied: the profiler was not available in this run)")
    valu = profile_counter(m["dom_name"], "valu_insts_per_launch") if args.config == "c2" else None
    valu_source = "SQ_INSTS_VALU from profiles/k_encode_b32_1024cube_f32.md (copied)"
    if measured is not None:
        valu = measured["valu"]
        valu_source = "SQ_INSTS_VALU of the same launch, rocprofv3 --pmc in this run"
    bricks0 = 1
    for d in range(3):
        bricks0 *= cfg["dims"][d] // cfg["brick"][d]
    fp64_ops = fp64_ops_full_brick() * bricks0 if tuple(cfg["brick"]) == (32, 32, 32) else None
    res = {
        "metric": METRIC,
        "value": round(m["value"], 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["wall"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian+sine+noise field generated on the GPU)",
        "config": {"workload": cfg["workload"], "dims": list(cfg["dims"]), "field_dtype": cfg["dtype"],
                   "levels": cfg["nlevels"], "tolerance": cfg["tolerance"], "brick": list(cfg["brick"])},
        "encode_gbps": round(m["enc_gbs"], 3),
        "decode_gbps": round(m["dec_gbs"], 3),
        "encode_ms": round(m["enc_ms"], 3),
        "decode_ms": round(m["dec_ms"], 3),
        "decode_p50_ms": round(m["dec_p50"], 3),
        "decode_p99_ms": round(m["dec_p99"], 3),
        "compressed_bytes": int(m["chunk_bytes"]),
        "max_abs_err": m["err"],
        "roofline": {"bound": "hbm", "achieved": round(m["achieved"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(m["achieved"] / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": m["dom_name"],
                     "kernel_ms": round(m["dom_ms"], 3), "launches": m["dom_launches"],
                     "algorithmic_bytes": int(m["dom_bytes"]),
                     "traffic_source": traffic_source,
                     "traffic_counters": measured["counters"] if measured is not None else None},
        "roofline_decode": {"bound": "hbm", "achieved": round(m["dec_achieved"], 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(m["dec_achieved"] / HBM_PEAK_GBS, 5),
                            "kernel": m["dec_name"], "kernel_ms": round(m["dec_dom_ms"], 3),
                            "launches": m["dec_launches"], "algorithmic_bytes": int(m["dec_dom_bytes"])},
    }
    if fp64_ops is not None:
        tf = fp64_ops / (m["dom_ms"] / 1e3) / 1e12
        res["compute"] = {"fp64": {"bound": "fp64", "ops_per_brick": fp64_ops_full_brick(), "ops": int(fp64_ops),
                                   "achieved": round(tf, 3), "peak": 78.6, "unit": "TFLOP/s",
                                   "frac": round(tf / 78.6, 5), "kernel": m["dom_name"],
                                   "note": "FP64 operations of the reference's transform of the launch's bricks "
                                           "(no FMA: the order of operations is the reference's)"}}
        if valu is not None:
            rate = valu / (m["dom_ms"] / 1e3) / 1e9
            res["compute"]["valu"] = {"bound": "valu", "insts_per_launch": valu, "achieved": round(rate, 1),
                                      "peak": 614.4, "unit": "G wave-instructions/s",
                                      "frac": round(rate / 614.4, 4),
                                      "source": valu_source + "; peak = 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per "
                                                              "wave64 instruction"}
    if tuning is not None:
        res["autotune"] = tuning
    if args.slice2d and args.config == "c2":
        log("2D slice")
        s2 = plain_bench(ctx, torch, CONFIGS["c2b"], 2, 1, args.seed)
        res["slice_2d"] = {"workload": CONFIGS["c2b"]["workload"], "gbps": round(s2["value"], 3),
                           "encode_ms": round(s2["enc_ms"], 3), "decode_ms": round(s2["dec_ms"], 3),
                           "kernel": s2["dom_name"], "kernel_ms": round(s2["dom_ms"], 3),
                           "kernel_achieved_gbps": round(s2["achieved"], 2), "max_abs_err": s2["err"]}
    if args.host_path and args.config == "c2":
        log("host-buffer path")
        res["host_path"] = host_path_bench(ctx, torch, cfg, args.seed)
    if args.end_to_end:
        log("end to end: encode + file write")
        res["end_to_end"] = end_to_end(ctx, torch, cfg, args.seed)
    if args.serve_requests > 0:
        log(f"serving loop: {args.serve_requests} requests at {args.serve_rate}/s")
        # the device blocks earlier legs left in the library's and torch's caches
        # go back first: the serving loop's budget and client buffers need the room
        api.release_device_cache(0)
        torch.cuda.empty_cache()
        res["serving"] = serve_bench(ctx, torch, args.serve_requests, args.serve_rate, args.seed)
    if args.progressive:
        log("progressive sweep")
        res["progressive"] = progressive_bench(ctx, torch, args.seed)
    if args.batched:
        log("batched 64-field mix")
        res["batched"] = batched_bench(ctx, torch, args.seed)
    if not args.no_cpu_baseline:
        log("CPU baseline")
        res["cpu_baseline"] = cpu_baseline(ctx, torch, args.seed)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
