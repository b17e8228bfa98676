"""Benchmark of the idx2 encode/decode path on one MI355X.

A step is one encode of the whole grid (input resident in HBM -> chunks
assembled in HBM) followed by one full-resolution decode (chunks in HBM ->
grid in HBM).  value = raw input bytes / (encode + decode time) in GB/s.
Prints one JSON line.
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
# bounded CPU sample of the same workload for the reference binary
CPU_SAMPLE = {
    "c2": dict(dims=(256, 256, 256), dtype="float32", nlevels=3, tolerance=1e-4, brick=(32, 32, 32)),
    "c2b": dict(dims=(2048, 2048, 1), dtype="float64", nlevels=3, tolerance=1e-4, brick=(32, 32, 1)),
    "c1": dict(dims=(384, 384, 256), dtype="float64", nlevels=2, tolerance=1e-3, brick=(32, 32, 32)),
}


def cpu_baseline(ctx, torch, cfg_name, seed):
    """Time the reference idx2App (built from the reference sources into
    oracle/_ref) on a bounded sample of the same workload: encode + serial
    full decode, one core."""
    from oracle import oracle as O  # checker/baseline only
    if not O.ref_available():
        return None
    s = CPU_SAMPLE[cfg_name]
    nx, ny, nz = s["dims"]
    tdt = torch.float64 if s["dtype"] == "float64" else torch.float32
    buf = torch.empty(nx * ny * nz, dtype=tdt, device="cuda")
    ctx.synth_field(buf.data_ptr(), s["dims"], s["dtype"], seed)
    torch.cuda.synchronize()
    host = buf.cpu().numpy().reshape(nz, ny, nx)
    raw = host.nbytes
    with tempfile.TemporaryDirectory() as wd:
        t0 = time.perf_counter()
        rc, rdir = O.ref_encode(wd, "B", "f", host, brick=s["brick"], nlevels=s["nlevels"], tolerance=s["tolerance"])
        t1 = time.perf_counter()
        if rc:
            return None
        rc, _ = O.ref_decode(wd, os.path.join(rdir, "B", "f.idx2"))
        t2 = time.perf_counter()
        if rc:
            return None
    return {"value": round(raw / ((t2 - t0)) / 1e9, 6), "unit": "GB/s", "cores": 1, "kind": "reference",
            "sample": f"{nx}x{ny}x{nz} {s['dtype']} of the same generator, brick {s['brick'][0]}, "
                      f"{s['nlevels']} levels, tol {s['tolerance']}: reference idx2App encode "
                      f"{t1 - t0:.2f}s + serial decode {t2 - t1:.2f}s (includes file I/O)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch
    from idx2_amd import api

    cfg = CONFIGS[args.config]
    nx, ny, nz = cfg["dims"]
    dt = cfg["dtype"]
    tdt = torch.float64 if dt == "float64" else torch.float32
    ctx = api.Context(0)
    inp = torch.empty(nx * ny * nz, dtype=tdt, device="cuda")
    ctx.synth_field(inp.data_ptr(), cfg["dims"], dt, args.seed)
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

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    enc, dec = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        es, dsx, chunk_bytes = step()
        enc.append(es)
        dec.append(dsx)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # error check of the last round trip (size-independent property: |x - x'| small)
    err = float((out - inp).abs().max().item())
    enc_ms = np.array([e["total_ms"] for e in enc])
    dec_ms = np.array([d["total_ms"] for d in dec])
    enc_gbs = raw / (enc_ms.mean() / 1e3) / 1e9
    dec_gbs = raw / (dec_ms.mean() / 1e3) / 1e9
    value = raw / ((enc_ms.mean() + dec_ms.mean()) / 1e3) / 1e9
    # dominant kernel: the level-0 brick kernel of encode or decode, whichever is longer
    dom_e = np.mean([e["dominant_ms"] for e in enc])
    dom_d = np.mean([d["dominant_ms"] for d in dec])
    if dom_e >= dom_d:
        dom_name, dom_ms = "k_encode_level (level 0)", dom_e
        dom_bytes = np.mean([e["dominant_bytes"] for e in enc])
        dom_launches = 1
    else:
        dom_name, dom_ms = "k_decode_level (level 0)", dom_d
        dom_bytes = np.mean([d["dominant_bytes"] for d in dec])
        dom_launches = int(np.mean([d["dominant_launches"] for d in dec]))
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian+sine+noise field generated on the GPU)",
        "config": {"workload": cfg["workload"], "dims": list(cfg["dims"]), "field_dtype": dt,
                   "levels": cfg["nlevels"], "tolerance": cfg["tolerance"], "brick": list(cfg["brick"])},
        "encode_gbps": round(enc_gbs, 3),
        "decode_gbps": round(dec_gbs, 3),
        "encode_ms": round(float(enc_ms.mean()), 3),
        "decode_ms": round(float(dec_ms.mean()), 3),
        "decode_p50_ms": round(float(np.percentile(dec_ms, 50)), 3),
        "decode_p99_ms": round(float(np.percentile(dec_ms, 99)), 3),
        "compressed_bytes": int(chunk_bytes),
        "max_abs_err": err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": dom_name,
                     "kernel_ms": round(float(dom_ms), 3), "launches": dom_launches,
                     "algorithmic_bytes": int(dom_bytes)},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(ctx, torch, args.config, args.seed)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
