"""Write golden fixtures from the reference idx2App (built from the reference
sources into oracle/_ref by oracle/ref.mk).

For each case: the input is regenerated from (dims, dtype, seed) by
idx2_amd.synth (its sha256 is recorded to catch generator drift); the
reference's encoded dataset (.idx2 + L*.bin files) is copied verbatim; and for
a set of decode queries the sha256 and a few samples of the reference's
decoded grid are recorded.

Run in the container that has the reference sources:
    make -f oracle/ref.mk && python tests/golden/make_fixtures.py [--only CASE ...]
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from idx2_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {
    "c3d_f32_64": dict(dims=(64, 64, 64), dtype="float32", nlevels=1, tolerance=1e-3,
                       brick=(32, 32, 32), bpc=4096, cpf=64, fpd=64, seed=7),
    "c3d_f64_partial": dict(dims=(48, 40, 36), dtype="float64", nlevels=1, tolerance=1e-4,
                            brick=(32, 32, 32), bpc=2, cpf=2, fpd=2, seed=3),
    "c3d_f32_2lvl": dict(dims=(72, 68, 66), dtype="float32", nlevels=2, tolerance=1e-2,
                         brick=(32, 32, 32), bpc=8, cpf=2, fpd=2, seed=5),
    "c3d_f64_b16": dict(dims=(40, 36, 34), dtype="float64", nlevels=2, tolerance=1e-6,
                        brick=(16, 16, 16), bpc=8, cpf=4, fpd=4, seed=9),
    "c2d_f64": dict(dims=(128, 96, 1), dtype="float64", nlevels=2, tolerance=1e-4,
                    brick=(32, 32, 1), bpc=4096, cpf=64, fpd=64, seed=1),
    "c2d_f32_3lvl": dict(dims=(300, 260, 1), dtype="float32", nlevels=3, tolerance=1e-4,
                         brick=(32, 32, 1), bpc=4, cpf=2, fpd=4, seed=2),
    # bricks above 32 samples per axis (the generic kernels, 64-bit block words)
    "c3d_f32_b64": dict(dims=(128, 80, 70), dtype="float32", nlevels=1, tolerance=1e-3,
                        brick=(64, 64, 64), bpc=4, cpf=2, fpd=2, seed=13),
    "c3d_f64_b64x32x16": dict(dims=(130, 66, 34), dtype="float64", nlevels=2, tolerance=1e-4,
                              brick=(64, 32, 16), bpc=8, cpf=4, fpd=4, seed=17),
    "c2d_f64_b64": dict(dims=(300, 260, 1), dtype="float64", nlevels=2, tolerance=1e-4,
                        brick=(64, 64, 1), bpc=4, cpf=4, fpd=4, seed=19),
}


def queries(dims):
    nx, ny, nz = dims
    is2d = nz == 1
    q = [dict(),
         dict(tolerance=1e-2),
         dict(downsampling=(1, 1, 0 if is2d else 1)),
         dict(downsampling=(1, 0, 0 if is2d else 2), tolerance=1e-3),
         dict(first=(3, 5, 0 if is2d else 2), last=(min(40, nx - 1), min(30, ny - 1), 0 if is2d else min(25, nz - 1))),
         dict(first=(7, 3, 0 if is2d else 1), last=(nx - 2, ny - 5, 0 if is2d else nz - 3), downsampling=(1, 1, 0 if is2d else 1)),
         dict(downsampling=(2, 2, 0 if is2d else 2)),
         ]
    return q


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def make_case(name, spec):
    out = os.path.join(HERE, name)
    shutil.rmtree(out, ignore_errors=True)
    os.makedirs(out)
    data = synth.make_field(spec["dims"], spec["dtype"], seed=spec["seed"])
    with tempfile.TemporaryDirectory() as wd:
        rc, rdir = O.ref_encode(wd, "G", "f", data, brick=spec["brick"], nlevels=spec["nlevels"],
                                tolerance=spec["tolerance"], bricks_per_chunk=spec["bpc"],
                                chunks_per_file=spec["cpf"], files_per_dir=spec["fpd"])
        assert rc == 0, rc
        shutil.copytree(os.path.join(rdir, "G"), os.path.join(out, "G"))
        idx = os.path.join(rdir, "G", "f.idx2")
        qs = []
        for q in queries(spec["dims"]):
            rc, path = O.ref_decode(wd, idx, **q)
            assert rc == 0, (name, q)
            dec = np.fromfile(path, dtype=data.dtype)
            rec = {k: list(v) if isinstance(v, tuple) else v for k, v in q.items()}
            rec["sha256"] = sha(dec)
            rec["size"] = int(dec.size)
            idxs = np.linspace(0, dec.size - 1, 8).astype(np.int64)
            rec["samples"] = [[int(i), float(dec[i])] for i in idxs]
            if not q:
                rec["max_abs_err"] = float(np.max(np.abs(dec.astype(np.float64) - data.reshape(-1).astype(np.float64))))
            qs.append(rec)
    meta = dict(spec)
    meta["input_sha256"] = sha(data)
    meta["queries"] = qs
    with open(os.path.join(out, "case.json"), "w") as f:
        json.dump(meta, f, indent=1)


# A dataset with one level-0 file deleted: the reference opens files as a
# decode reaches them and fails with FileNotFound (idx2Read.cpp:252-270)
# exactly when the query needs a brick of the missing file; the other
# queries decode normally.
MISSING = dict(source="c3d_f32_2lvl", removed="L00/B1/B2.bin",
               queries=[dict(), dict(downsampling=(1, 1, 1)), dict(first=(0, 0, 0), last=(31, 31, 31)),
                        dict(first=(40, 40, 40), last=(71, 67, 65)), dict(first=(0, 0, 0), last=(63, 63, 31)),
                        dict(tolerance=1e-1), dict(first=(5, 0, 0), last=(60, 30, 20), downsampling=(1, 0, 1))])


def make_missing_file_case():
    src = os.path.join(HERE, MISSING["source"], "G")
    with tempfile.TemporaryDirectory() as wd:
        shutil.copytree(src, os.path.join(wd, "G"))
        os.remove(os.path.join(wd, "G", "f", MISSING["removed"]))
        idx = os.path.join(wd, "G", "f.idx2")
        qs = []
        for q in MISSING["queries"]:
            out_dir = os.path.join(wd, "refdec")
            shutil.rmtree(out_dir, ignore_errors=True)
            os.makedirs(out_dir)
            cmd = [O.REF_APP, "--decode", idx, "--out_dir", out_dir, "--out_file", "out.raw",
                   "--downsampling", *[str(d) for d in q.get("downsampling", (0, 0, 0))],
                   "--tolerance", repr(q.get("tolerance", 0.0))]
            if "first" in q:
                cmd += ["--first", *[str(v) for v in q["first"]], "--last", *[str(v) for v in q["last"]]]
            r = subprocess.run(cmd, capture_output=True, text=True)
            rec = {k: list(v) if isinstance(v, tuple) else v for k, v in q.items()}
            if r.returncode == 0:
                dec = np.fromfile(os.path.join(out_dir, "out.raw"), dtype=np.float32)
                rec.update(result="ok", sha256=sha(dec), size=int(dec.size))
            else:
                msg = r.stdout + r.stderr
                rec.update(result="FileNotFound" if "FileNotFound" in msg else "error")
            qs.append(rec)
    with open(os.path.join(HERE, "missing_file.json"), "w") as f:
        json.dump(dict(source=MISSING["source"], removed=MISSING["removed"], queries=qs), f, indent=1)


# Sparse-brick output (idx2App --brick_map, params::out_mode::HashMap): the
# reference's "bricks" file (WriteBricks, idx2SparseBricks.cpp:207-243) and its
# per-level significant-brick counts for decodes of fixture datasets, and of
# a field that is zero on the slab x0 <= x < x1, whose level-0 bricks there
# are not significant and are written from their level-1 parents.
SPARSE = dict(dims=(80, 66, 66), dtype="float32", nlevels=2, tolerance=1e-3, brick=(32, 32, 32), bpc=8, cpf=4,
              fpd=4, seed=11, zero_x=(30, 64))
BRICK_MAP = [("c3d_f32_2lvl", {}), ("c3d_f32_2lvl", dict(downsampling=(1, 1, 1))), ("c3d_f64_b16", {}),
             ("c3d_f32_64", {}), ("c2d_f32_3lvl", {}), ("c2d_f32_3lvl", dict(tolerance=1.0)),
             ("sparse", {}), ("sparse", dict(tolerance=1e-2)), ("sparse", dict(downsampling=(1, 1, 1)))]


def sparse_field(spec):
    data = synth.make_field(spec["dims"], spec["dtype"], seed=spec["seed"])
    x0, x1 = spec["zero_x"]
    data[:, :, x0:x1] = 0
    return data


def make_brick_map_cases():
    out = os.path.join(HERE, "brick_map")
    shutil.rmtree(out, ignore_errors=True)
    os.makedirs(out)
    data = sparse_field(SPARSE)
    recs = []
    with tempfile.TemporaryDirectory() as wd:
        rc, rdir = O.ref_encode(wd, "G", "f", data, brick=SPARSE["brick"], nlevels=SPARSE["nlevels"],
                                tolerance=SPARSE["tolerance"], bricks_per_chunk=SPARSE["bpc"],
                                chunks_per_file=SPARSE["cpf"], files_per_dir=SPARSE["fpd"])
        assert rc == 0, rc
        shutil.copytree(os.path.join(rdir, "G"), os.path.join(out, "G"))
        for src, q in BRICK_MAP:
            in_dir = out if src == "sparse" else os.path.join(HERE, src)
            run = os.path.join(wd, "bm")
            shutil.rmtree(run, ignore_errors=True)
            os.makedirs(run)
            cmd = [O.REF_APP, "--decode", os.path.join(in_dir, "G", "f.idx2"), "--in_dir", in_dir, "--brick_map",
                   "--downsampling", *[str(d) for d in q.get("downsampling", (0, 0, 0))],
                   "--tolerance", repr(q.get("tolerance", 0.0))]
            r = subprocess.run(cmd, cwd=run, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, (src, q, r.stdout[-500:], r.stderr[-500:])
            counts = []
            for line in r.stdout.splitlines():
                if "bricks significant (" in line:  # level L: K out of N bricks significant (P percent)
                    w = line.replace(":", " ").split()
                    counts.append([int(w[1]), int(w[2]), int(w[5])])
            blob = open(os.path.join(run, "bricks"), "rb").read()
            rec = {k: list(v) if isinstance(v, tuple) else v for k, v in q.items()}
            rec.update(source=src, sha256=hashlib.sha256(blob).hexdigest(), size=len(blob), counts=counts)
            recs.append(rec)
    spec = {k: list(v) if isinstance(v, tuple) else v for k, v in SPARSE.items()}
    spec["input_sha256"] = sha(data)
    with open(os.path.join(out, "brick_map.json"), "w") as f:
        json.dump(dict(sparse=spec, decodes=recs), f, indent=1)


def main():
    assert O.ref_available(), "build the reference first: make -f oracle/ref.mk"
    only = sys.argv[sys.argv.index("--only") + 1:] if "--only" in sys.argv else None
    for name, spec in CASES.items():
        if only is None or name in only:
            make_case(name, spec)
            print("wrote", name)
    if only is None:
        make_missing_file_case()
        print("wrote missing_file.json")
        make_brick_map_cases()
        print("wrote brick_map")


if __name__ == "__main__":
    main()
