"""ctypes binding of the C oracle (oracle/idx2_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(idx2_amd) never imports this module.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libidx2oracle.so")
# Same oracle linked against the reference's vendored zstd 1.4.5 (built by
# oracle/ref.mk only where the reference sources exist).
LIB_Z145_PATH = os.path.join(HERE, "_ref", "libidx2oracle_z145.so")
REF_APP = os.path.join(HERE, "_ref", "idx2App")

_libs = {}
_variant = "system"


def use_variant(v):
    """'system' (libzstd.so.1, as the product) or 'z145' (reference zstd)."""
    global _variant
    _variant = v


def build():
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib(variant=None):
    variant = variant or _variant
    if variant not in _libs:
        path = LIB_Z145_PATH if variant == "z145" else LIB_PATH
        if variant != "z145" and not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i3 = ctypes.c_int * 3
        L.idx2o_encode.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p,
                                   ctypes.c_int, i3, i3, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.idx2o_output_dims.argtypes = [ctypes.c_char_p, i3, i3, i3, i3, ctypes.POINTER(ctypes.c_int)]
        L.idx2o_decode.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i3, i3, i3, ctypes.c_double,
                                   ctypes.c_void_p, ctypes.c_int64]
        L.idx2o_extrapolate.argtypes = [i3, i3, ctypes.c_void_p]
        L.idx2o_forward_brick.argtypes = [i3, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.idx2o_inverse_brick.argtypes = [i3, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.idx2o_flift.argtypes = [ctypes.c_int, i3, i3, i3, i3, i3, ctypes.c_void_p]
        L.idx2o_ilift.argtypes = [ctypes.c_int, i3, i3, i3, i3, i3, ctypes.c_void_p]
        L.idx2o_block_forward.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.idx2o_block_inverse.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.idx2o_encode_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), ctypes.c_void_p, ctypes.c_int64]
        L.idx2o_decode_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), ctypes.c_void_p, ctypes.c_int64]
        L.idx2o_cdf53_norms.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.idx2o_zfp_lift.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _libs[variant] = L
    return _libs[variant]


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def encode(out_dir, name, field, data, brick=(32, 32, 32), nlevels=1, tolerance=1e-3,
           bricks_per_chunk=4096, chunks_per_file=64, files_per_dir=64):
    """Encode a numpy array of shape (nz, ny, nx) into out_dir/name/field*."""
    data = np.ascontiguousarray(data)
    dtype = 1 if data.dtype == np.float64 else 0
    nz, ny, nx = data.shape
    shutil.rmtree(os.path.join(out_dir, name, field), ignore_errors=True)
    rc = lib().idx2o_encode(out_dir.encode(), name.encode(), field.encode(),
                            data.ctypes.data_as(ctypes.c_void_p), dtype, _i3((nx, ny, nz)),
                            _i3(brick), nlevels, tolerance, bricks_per_chunk, chunks_per_file,
                            files_per_dir)
    return rc


def decode(idx2_path, in_dir, first=None, last=None, downsampling=(0, 0, 0), tolerance=0.0):
    """Decode a query; returns (rc, array of shape (nz, ny, nx))."""
    f = _i3(first if first is not None else (-1, -1, -1))
    l_ = _i3(last if last is not None else (-1, -1, -1))
    ds = _i3(downsampling)
    dims = _i3((0, 0, 0))
    dt = ctypes.c_int(0)
    rc = lib().idx2o_output_dims(idx2_path.encode(), f, l_, ds, dims, ctypes.byref(dt))
    if rc:
        return rc, None
    shape = (dims[2], dims[1], dims[0])
    out = np.zeros(shape, dtype=np.float64 if dt.value else np.float32)
    rc = lib().idx2o_decode(idx2_path.encode(), in_dir.encode(), f, l_, ds, tolerance,
                            out.ctypes.data_as(ctypes.c_void_p), out.nbytes)
    return rc, out


def ref_available():
    return os.path.exists(REF_APP)


def _pinned(cpus):
    """preexec_fn restricting the child process to the CPUs in `cpus` (None: unchanged)."""
    if cpus is None:
        return None
    return lambda: os.sched_setaffinity(0, cpus)


def ref_write_raw(workdir, name, field, data):
    """The raw input file the reference idx2App reads, named as it expects."""
    data = np.ascontiguousarray(data)
    nz, ny, nx = data.shape
    tname = "Float64" if data.dtype == np.float64 else "Float32"
    raw = os.path.join(workdir, f"{name}-{field}-[{nx}-{ny}-{nz}]-{tname}.raw")
    os.makedirs(workdir, exist_ok=True)
    data.tofile(raw)
    return raw


def ref_encode_raw(raw, out_dir, brick=(32, 32, 32), nlevels=1, tolerance=1e-3, bricks_per_chunk=4096,
                   chunks_per_file=64, files_per_dir=64, cpus=None):
    """Run the reference idx2App (built from the reference sources into
    oracle/_ref) on a raw file written by ref_write_raw."""
    cmd = [REF_APP, "--encode", raw, "--num_levels", str(nlevels), "--tolerance", repr(tolerance),
           "--brick_size", *[str(b) for b in brick], "--bricks_per_chunk", str(bricks_per_chunk),
           "--chunks_per_file", str(chunks_per_file), "--files_per_dir", str(files_per_dir),
           "--out_dir", out_dir]
    r = subprocess.run(cmd, capture_output=True, text=True, preexec_fn=_pinned(cpus))
    return r.returncode


def ref_encode(workdir, name, field, data, brick=(32, 32, 32), nlevels=1, tolerance=1e-3,
               bricks_per_chunk=4096, chunks_per_file=64, files_per_dir=64):
    """Write the raw file and run the reference idx2App's encode on it."""
    raw = ref_write_raw(workdir, name, field, data)
    out_dir = os.path.join(workdir, "ref")
    rc = ref_encode_raw(raw, out_dir, brick, nlevels, tolerance, bricks_per_chunk, chunks_per_file, files_per_dir)
    return rc, out_dir


def ref_decode(workdir, idx2_path, first=None, last=None, downsampling=(0, 0, 0), tolerance=0.0, parallel=False,
               cpus=None):
    """The reference idx2App's decode (serial, or its thread-pool decoder
    with parallel=True) into workdir/refdec/out.raw; cpus: CPU set the
    process is restricted to."""
    out_dir = os.path.join(workdir, "refdec")
    shutil.rmtree(out_dir, ignore_errors=True)
    os.makedirs(out_dir)
    cmd = [REF_APP, "--decode", idx2_path, "--out_dir", out_dir, "--out_file", "out.raw",
           "--downsampling", *[str(d) for d in downsampling], "--tolerance", repr(tolerance)]
    if first is not None:
        cmd += ["--first", *[str(v) for v in first], "--last", *[str(v) for v in last]]
    if parallel:
        cmd += ["--parallel"]
    r = subprocess.run(cmd, capture_output=True, text=True, preexec_fn=_pinned(cpus))
    if r.returncode != 0:
        return r.returncode, None
    return 0, os.path.join(out_dir, "out.raw")
