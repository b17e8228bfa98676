"""Python binding of libidx2g (include/idx2g.h) over ctypes.

Mirrors the reference's library surface (Source/idx2.h): encode a grid,
open a dataset, ask for the output grid of a query, decode into a buffer.
Device buffers are plain integer addresses (e.g. torch.Tensor.data_ptr()).
The library must have been built (idx2_amd/build.py); there is no fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libidx2g.so")

# idx2g error codes: the values of the reference's enum class idx2_err_code
# (Source/Core/idx2Common.h:36-56 after Source/Core/ErrorCodes.h:6-12;
# idx2_Enum of Source/Core/Enum.h:65-80 puts __Invalid__ last, NoError = 0).
ERRORS = {
    0: "NoError", 1: "UnknownError", 2: "SizeZero", 3: "SizeTooSmall", 4: "SizeMismatched",
    5: "DimensionMismatched", 6: "DimensionsTooMany", 7: "AttributeNotFound", 8: "OptionNotSupported",
    9: "TypeNotSupported", 10: "FileCreateFailed", 11: "FileReadFailed", 12: "FileWriteFailed",
    13: "FileOpenFailed", 14: "FileCloseFailed", 15: "FileSeekFailed", 16: "FileTellFailed",
    17: "ParseFailed", 18: "OutOfMemory", 19: "DimensionsRepeated", 20: "DimensionTooSmall",
    21: "BrickSizeNotPowerOfTwo", 22: "BrickSizeTooBig", 23: "TooManyLevels",
    24: "TooManyTransformPassesPerLevel", 25: "TooManyLevelsOrTransformPasses", 26: "TooManyBricksPerFile",
    27: "TooManyFilesPerDir", 28: "NotSupportedInVersion", 29: "CannotCreateDirectory",
    30: "SyntaxError", 31: "TooManyBricksPerChunk", 32: "TooManyChunksPerFile",
    33: "ChunksPerFileNotPowerOf2", 34: "BricksPerChunkNotPowerOf2", 35: "ChunkNotFound",
    36: "BrickNotFound", 37: "FileNotFound", 38: "UnsupportedScheme", 100: "DeviceError",
}


class Idx2Error(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)} ({code}) {msg}".strip())


class EncodeParams(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("field", ctypes.c_char_p), ("dims", ctypes.c_int32 * 3),
                ("dtype", ctypes.c_int32), ("brick", ctypes.c_int32 * 3), ("nlevels", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("bricks_per_chunk", ctypes.c_int32),
                ("chunks_per_file", ctypes.c_int32), ("files_per_dir", ctypes.c_int32),
                ("bit_planes_per_chunk", ctypes.c_int32)]


class Query(ctypes.Structure):
    _fields_ = [("first", ctypes.c_int32 * 3), ("last", ctypes.c_int32 * 3),
                ("downsampling", ctypes.c_int32 * 3), ("tolerance", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("transform_ms", ctypes.c_double),
                ("assemble_ms", ctypes.c_double), ("dominant_launches", ctypes.c_int64),
                ("dominant_ms", ctypes.c_double), ("dominant_bytes", ctypes.c_int64), ("raw_bytes", ctypes.c_int64),
                ("chunk_bytes", ctypes.c_int64), ("exp_bytes", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_int64), ("batches", ctypes.c_int64), ("batched_requests", ctypes.c_int64),
                ("bypasses", ctypes.c_int64), ("cache_hits", ctypes.c_int64), ("cache_misses", ctypes.c_int64),
                ("cache_inserts", ctypes.c_int64), ("cache_evictions", ctypes.c_int64),
                ("cache_bytes", ctypes.c_int64), ("peak_bytes", ctypes.c_int64), ("peak_running", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("enc_wg_per_cu", ctypes.c_int32), ("dec_wg_per_cu", ctypes.c_int32), ("walk_batch", ctypes.c_int32),
                ("walk_wide_max", ctypes.c_int32)]


EXPORTS = ["idx2g_create", "idx2g_destroy", "idx2g_last_error_message", "idx2g_encode",
           "idx2g_dataset_write", "idx2g_open", "idx2g_dataset_free", "idx2g_dataset_info",
           "idx2g_output_dims", "idx2g_decode", "idx2g_last_stats", "idx2g_dataset_chunk_bytes",
           "idx2g_synth_field", "idx2g_server_create", "idx2g_server_submit", "idx2g_server_wait",
           "idx2g_server_info", "idx2g_server_destroy", "idx2g_decode_footprint", "idx2g_encode_batch",
           "idx2g_decode_batch", "idx2g_validate_params", "idx2g_autotune", "idx2g_set_tuning",
           "idx2g_get_tuning", "idx2g_decode_noalloc", "idx2g_server_counters",
           "idx2g_decode_bricks", "idx2g_brick_pool_count", "idx2g_brick_pool_key", "idx2g_brick_pool_find",
           "idx2g_brick_pool_lookup", "idx2g_brick_pool_resolution", "idx2g_brick_pool_copy", "idx2g_brick_pool_write",
           "idx2g_brick_pool_free", "idx2g_release_device_cache"]

_lib = None


def _share_hip_runtime_with_torch():
    """libidx2g.so needs libamdhip64.so.7.  When torch (which bundles its own
    HIP runtime under that soname) is loaded first, the library binds to that
    same runtime, so device pointers and streams are shared and only one HIP
    runtime lives in the process.  Loading libidx2g.so first would bring in
    a second runtime (/opt/rocm), after which torch cannot initialise HIP."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is None:
        _share_hip_runtime_with_torch()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with idx2_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.idx2g_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.idx2g_destroy.argtypes = [vp]
        L.idx2g_last_error_message.restype = ctypes.c_char_p
        L.idx2g_encode.argtypes = [vp, ctypes.POINTER(EncodeParams), vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.idx2g_dataset_write.argtypes = [vp, ctypes.c_char_p]
        L.idx2g_open.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.idx2g_dataset_free.argtypes = [vp]
        L.idx2g_dataset_info.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                         ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]
        L.idx2g_output_dims.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.idx2g_decode.argtypes = [vp, vp, ctypes.POINTER(Query), vp, ctypes.c_int]
        L.idx2g_decode_noalloc.argtypes = [vp, vp, ctypes.POINTER(Query), vp, ctypes.POINTER(i64)]
        L.idx2g_server_counters.argtypes = [vp, ctypes.POINTER(ServerStats)]
        i3 = ctypes.c_int * 3
        L.idx2g_decode_bricks.argtypes = [vp, vp, ctypes.POINTER(Query), ctypes.POINTER(vp)]
        L.idx2g_brick_pool_count.argtypes = [vp, ctypes.c_int, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.idx2g_brick_pool_key.argtypes = [vp, ctypes.c_int, i3, ctypes.POINTER(ctypes.c_uint64)]
        L.idx2g_brick_pool_find.argtypes = [vp, ctypes.c_int, i3, ctypes.POINTER(vp), i3]
        L.idx2g_brick_pool_lookup.argtypes = [vp, vp, i64, vp, vp]
        L.idx2g_brick_pool_copy.argtypes = [vp, ctypes.c_int, i3, vp, ctypes.c_int]
        L.idx2g_brick_pool_resolution.argtypes = [vp, i3, ctypes.POINTER(ctypes.c_int)]
        L.idx2g_brick_pool_write.argtypes = [vp, ctypes.c_char_p]
        L.idx2g_brick_pool_free.argtypes = [vp]
        L.idx2g_last_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.idx2g_dataset_chunk_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(i64)]
        L.idx2g_synth_field.argtypes = [vp, vp, ctypes.POINTER(i32), i32, ctypes.POINTER(ctypes.c_double),
                                        ctypes.c_uint64, i32, i32]
        L.idx2g_server_create.argtypes = [ctypes.c_int, i64, ctypes.c_int, ctypes.POINTER(vp)]
        L.idx2g_server_submit.argtypes = [vp, vp, ctypes.POINTER(Query), vp, ctypes.c_int, ctypes.POINTER(i64)]
        L.idx2g_server_wait.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(Stats)]
        L.idx2g_server_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(i64)]
        L.idx2g_server_destroy.argtypes = [vp]
        L.idx2g_decode_footprint.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(i64)]
        L.idx2g_validate_params.argtypes = [ctypes.POINTER(EncodeParams)]
        L.idx2g_autotune.argtypes = [vp, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
        L.idx2g_set_tuning.argtypes = [ctypes.c_int, ctypes.POINTER(Tuning)]
        L.idx2g_get_tuning.argtypes = [ctypes.c_int, ctypes.POINTER(Tuning)]
        L.idx2g_release_device_cache.argtypes = [ctypes.c_int]
        L.idx2g_release_device_cache.restype = ctypes.c_int64
        L.idx2g_encode_batch.argtypes = [vp, ctypes.c_int, ctypes.POINTER(EncodeParams), ctypes.POINTER(vp),
                                         ctypes.c_int, ctypes.POINTER(vp)]
        L.idx2g_decode_batch.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(Query),
                                         ctypes.POINTER(vp), ctypes.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise Idx2Error(rc, lib().idx2g_last_error_message().decode(errors="replace"))


def make_query(first=None, last=None, downsampling=(0, 0, 0), tolerance=0.0):
    q = Query()
    f = first if first is not None else (-1, -1, -1)
    l_ = last if last is not None else (-1, -1, -1)
    for d in range(3):
        q.first[d] = int(f[d])
        q.last[d] = int(l_[d])
        q.downsampling[d] = int(downsampling[d])
    q.tolerance = float(tolerance)
    return q


class Context:
    """One GPU, one HIP stream (idx2g_ctx)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(lib().idx2g_create(ctypes.byref(self._h), device))

    def close(self):
        if self._h:
            lib().idx2g_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = Stats()
        _check(lib().idx2g_last_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def synth_field(self, dev_ptr, dims, dtype, seed, z_range=None):
        from . import synth
        prm = synth.field_params(seed)
        arr = (ctypes.c_double * 192)(*prm.tolist())
        d3 = (ctypes.c_int32 * 3)(*[int(x) for x in dims])
        z0, z1 = (0, int(dims[2])) if z_range is None else z_range
        _check(lib().idx2g_synth_field(self._h, ctypes.c_void_p(dev_ptr), d3, 1 if dtype == "float64" else 0, arr,
                                       seed + 0x5EED, z0, z1))


class Dataset:
    """A dataset resident on the GPU (idx2g_dataset)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self._h = handle

    @staticmethod
    def _params(data=None, *, dev_ptr=None, dims=None, dtype=None, name="data", field="field",
                brick=(32, 32, 32), nlevels=1, tolerance=1e-3, bricks_per_chunk=4096, chunks_per_file=64,
                files_per_dir=64):
        """(EncodeParams, input pointer, input is on the device, kept objects)."""
        p = EncodeParams()
        keep = [name.encode(), field.encode()]
        p.name, p.field = keep[0], keep[1]
        if data is not None:
            data = np.ascontiguousarray(data)
            keep.append(data)
            nz, ny, nx = data.shape
            dims = (nx, ny, nz)
            dtype = "float64" if data.dtype == np.float64 else "float32"
            ptr, is_dev = data.ctypes.data_as(ctypes.c_void_p), 0
        else:
            ptr, is_dev = ctypes.c_void_p(dev_ptr), 1
        for d in range(3):
            p.dims[d] = int(dims[d])
            p.brick[d] = int(brick[d])
        p.dtype = 1 if dtype in ("float64", np.float64) else 0
        p.nlevels = nlevels
        p.tolerance = tolerance
        p.bricks_per_chunk = bricks_per_chunk
        p.chunks_per_file = chunks_per_file
        p.files_per_dir = files_per_dir
        p.bit_planes_per_chunk = 1
        return p, ptr, is_dev, keep

    @classmethod
    def encode(cls, ctx, data=None, **kw):
        """Encode a numpy array of shape (nz, ny, nx), or a device buffer (dev_ptr=, dims=, dtype=)."""
        p, ptr, is_dev, keep = cls._params(data, **kw)
        h = ctypes.c_void_p()
        _check(lib().idx2g_encode(ctx._h, ctypes.byref(p), ptr, is_dev, ctypes.byref(h)))
        del keep
        return cls(ctx, h)

    @classmethod
    def encode_batch(cls, ctx, fields):
        """Encode several fields in one pass (idx2g_encode_batch).  fields: list of
        dicts of Dataset.encode arguments, all with `data` (host arrays) or all
        with `dev_ptr` (device buffers).  Returns one Dataset per field."""
        n = len(fields)
        ps = (EncodeParams * n)()
        ptrs = (ctypes.c_void_p * n)()
        keep, devs = [], set()
        for i, f in enumerate(fields):
            f = dict(f)
            p, ptr, is_dev, k = cls._params(f.pop("data", None), **f)
            ps[i] = p
            ptrs[i] = ptr
            devs.add(is_dev)
            keep.append(k)
        if len(devs) != 1:
            raise ValueError("encode_batch: inputs must be all host arrays or all device buffers")
        outs = (ctypes.c_void_p * n)()
        _check(lib().idx2g_encode_batch(ctx._h, n, ps, ptrs, devs.pop(), outs))
        del keep
        return [cls(ctx, ctypes.c_void_p(outs[i])) for i in range(n)]

    @classmethod
    def open(cls, ctx, idx2_path, in_dir=None):
        h = ctypes.c_void_p()
        if in_dir is None:
            in_dir = os.path.dirname(os.path.dirname(os.path.abspath(idx2_path)))
        _check(lib().idx2g_open(ctx._h, idx2_path.encode(), in_dir.encode(), ctypes.byref(h)))
        return cls(ctx, h)

    def write(self, out_dir):
        _check(lib().idx2g_dataset_write(self._h, out_dir.encode()))

    def free(self):
        if self._h:
            lib().idx2g_dataset_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        d, b = (ctypes.c_int32 * 3)(), (ctypes.c_int32 * 3)()
        dt, nl = ctypes.c_int32(), ctypes.c_int32()
        tol, cb = ctypes.c_double(), ctypes.c_int64()
        _check(lib().idx2g_dataset_info(self._h, d, ctypes.byref(dt), ctypes.byref(nl), b, ctypes.byref(tol),
                                        ctypes.byref(cb)))
        return dict(dims=tuple(d), dtype="float64" if dt.value else "float32", nlevels=nl.value, brick=tuple(b),
                    tolerance=tol.value, chunk_bytes=cb.value)

    def output_dims(self, **query):
        q = make_query(**query)
        dims = (ctypes.c_int64 * 3)()
        nbytes = ctypes.c_int64()
        _check(lib().idx2g_output_dims(self._h, ctypes.byref(q), dims, ctypes.byref(nbytes)))
        return tuple(dims), nbytes.value

    def decode(self, out_dev_ptr=None, **query):
        """Decode a query; returns a numpy array (nz, ny, nx), or fills a device buffer."""
        q = make_query(**query)
        dims, nbytes = self.output_dims(**query)
        if out_dev_ptr is not None:
            _check(lib().idx2g_decode(self.ctx._h, self._h, ctypes.byref(q), ctypes.c_void_p(out_dev_ptr), 1))
            return None
        dt = np.float64 if self.info()["dtype"] == "float64" else np.float32
        out = np.empty((dims[2], dims[1], dims[0]), dtype=dt)
        _check(lib().idx2g_decode(self.ctx._h, self._h, ctypes.byref(q), out.ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def decode_bricks(self, **query):
        """Sparse-brick decode (the reference's HashMap output mode): a
        BrickPool of the significant bricks of every decoded level."""
        q = make_query(**query)
        h = ctypes.c_void_p()
        _check(lib().idx2g_decode_bricks(self.ctx._h, self._h, ctypes.byref(q), ctypes.byref(h)))
        return BrickPool(h, self.info()["nlevels"])

    def chunk_bytes(self):
        n = ctypes.c_int64()
        _check(lib().idx2g_dataset_chunk_bytes(self._h, None, ctypes.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        _check(lib().idx2g_dataset_chunk_bytes(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                               ctypes.byref(n)))
        return buf


class BrickPool:
    """Bricks kept by Dataset.decode_bricks (idx2g_brick_pool)."""

    def __init__(self, h, nlevels):
        self._h = h
        self.nlevels = nlevels

    def count(self, level):
        """(kept bricks, bricks of the level's grid)"""
        k, t = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().idx2g_brick_pool_count(self._h, level, ctypes.byref(k), ctypes.byref(t)))
        return k.value, t.value

    def key(self, level, brick3):
        k = ctypes.c_uint64()
        _check(lib().idx2g_brick_pool_key(self._h, level, (ctypes.c_int * 3)(*brick3), ctypes.byref(k)))
        return k.value

    def find(self, level, brick3):
        """(device address, dims (x, y, z)) of a kept brick, None when not kept."""
        ptr, dims = ctypes.c_void_p(), (ctypes.c_int * 3)()
        rc = lib().idx2g_brick_pool_find(self._h, level, (ctypes.c_int * 3)(*brick3), ctypes.byref(ptr), dims)
        if rc == 7:  # AttributeNotFound
            return None
        _check(rc)
        return ptr.value, tuple(dims)

    def samples(self, level, brick3):
        """A kept brick's samples as a numpy float64 array (z, y, x); None when not kept."""
        f = self.find(level, brick3)
        if f is None:
            return None
        dims = f[1]
        out = np.empty((dims[2], dims[1], dims[0]), dtype=np.float64)
        _check(lib().idx2g_brick_pool_copy(self._h, level, (ctypes.c_int * 3)(*brick3),
                                           out.ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def lookup(self, ctx, n, keys_dev_ptr, out_dev_ptr):
        """Device-side lookup of n brick keys (u64 device array) into device addresses (u64 device array)."""
        _check(lib().idx2g_brick_pool_lookup(ctx._h, self._h, n, ctypes.c_void_p(keys_dev_ptr),
                                             ctypes.c_void_p(out_dev_ptr)))

    def resolution(self, brick3):
        r = ctypes.c_int()
        _check(lib().idx2g_brick_pool_resolution(self._h, (ctypes.c_int * 3)(*brick3), ctypes.byref(r)))
        return r.value

    def write(self, path):
        _check(lib().idx2g_brick_pool_write(self._h, path.encode()))

    def free(self):
        if self._h:
            _check(lib().idx2g_brick_pool_free(self._h))
            self._h = None


def autotune(ctx):
    """Pick launch shapes for ctx's device by timing probe encodes/decodes
    (idx2g_autotune); returns (shapes, tuned ms, default ms)."""
    t = Tuning()
    best, dflt = ctypes.c_double(), ctypes.c_double()
    _check(lib().idx2g_autotune(ctx._h, ctypes.byref(t), ctypes.byref(best), ctypes.byref(dflt)))
    return {k: getattr(t, k) for k, _ in Tuning._fields_}, best.value, dflt.value


def set_tuning(device=0, enc_wg_per_cu=0, dec_wg_per_cu=0, walk_batch=0, walk_wide_max=0):
    t = Tuning(enc_wg_per_cu, dec_wg_per_cu, walk_batch, walk_wide_max)
    _check(lib().idx2g_set_tuning(device, ctypes.byref(t)))


def release_device_cache(device=0):
    """Hand the device blocks the library keeps for reuse back to the runtime
    (idx2g_release_device_cache); returns the bytes released."""
    r = lib().idx2g_release_device_cache(device)
    if r < 0:
        raise Idx2Error(-r, lib().idx2g_last_error_message().decode(errors="replace"))
    return r


def get_tuning(device=0):
    t = Tuning()
    _check(lib().idx2g_get_tuning(device, ctypes.byref(t)))
    return {k: getattr(t, k) for k, _ in Tuning._fields_}


def validate_params(**kw):
    """Error code idx2g_encode would return for these Dataset.encode keyword
    arguments (dims=, dtype=, brick=, nlevels=, ...), checked on the host."""
    p, _, _, keep = Dataset._params(dev_ptr=0, **kw)
    rc = lib().idx2g_validate_params(ctypes.byref(p))
    del keep
    return rc


def max_levels(dims, brick, limit=8):
    """Largest level count the layout accepts for a grid (0 if none)."""
    best = 0
    for nl in range(1, limit + 1):
        if validate_params(dims=dims, dtype="float32", brick=brick, nlevels=nl) == 0:
            best = nl
    return best


def decode_batch(ctx, datasets, out_dev_ptrs, queries=None):
    """Decode query i of datasets[i] into device buffer out_dev_ptrs[i] for every i
    in one pass (idx2g_decode_batch).  queries: list of make_query() keyword
    dicts, or None for whole volumes."""
    n = len(datasets)
    hs = (ctypes.c_void_p * n)(*[d._h.value if isinstance(d._h, ctypes.c_void_p) else d._h for d in datasets])
    outs = (ctypes.c_void_p * n)(*[int(p) for p in out_dev_ptrs])
    qs = None
    if queries is not None:
        qs = (Query * n)()
        for i, q in enumerate(queries):
            qs[i] = make_query(**q)
    _check(lib().idx2g_decode_batch(ctx._h, n, hs, qs, outs, 1))


class Server:
    """Decode serving loop (idx2g_server): queued requests on resident
    datasets, admitted against an HBM budget, several running at once."""

    def __init__(self, device=0, hbm_budget=200 << 30, slots=4):
        self._h = ctypes.c_void_p()
        self._held = {}  # request id -> (dataset, output owner) kept alive until wait()
        _check(lib().idx2g_server_create(device, int(hbm_budget), int(slots), ctypes.byref(self._h)))

    def submit(self, ds, out_dev_ptr, out_owner=None, **query):
        """Queue a decode of `query` on `ds` into a device buffer; returns the
        request id.  `ds` and `out_owner` (e.g. the tensor behind out_dev_ptr)
        stay referenced until wait() returns for the request."""
        q = make_query(**query)
        rid = ctypes.c_int64()
        _check(lib().idx2g_server_submit(self._h, ds._h, ctypes.byref(q), ctypes.c_void_p(out_dev_ptr), 1,
                                         ctypes.byref(rid)))
        self._held[rid.value] = (ds, out_owner)
        return rid.value

    def wait(self, rid):
        """Block until request `rid` is done: (latency_ms, queue_ms, stats)."""
        lat, qms = ctypes.c_double(), ctypes.c_double()
        st = Stats()
        try:
            _check(lib().idx2g_server_wait(self._h, int(rid), ctypes.byref(lat), ctypes.byref(qms),
                                           ctypes.byref(st)))
        finally:
            self._held.pop(int(rid), None)
        return lat.value, qms.value, {k: getattr(st, k) for k, _ in Stats._fields_}

    def counters(self):
        """Batches, bypasses and coarse-level cache activity of the serving loop."""
        c = ServerStats()
        _check(lib().idx2g_server_counters(self._h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in ServerStats._fields_}

    def info(self):
        pb, pr, done = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int64()
        _check(lib().idx2g_server_info(self._h, ctypes.byref(pb), ctypes.byref(pr), ctypes.byref(done)))
        return dict(peak_bytes=pb.value, peak_running=pr.value, completed=done.value)

    def close(self):
        if self._h:
            lib().idx2g_server_destroy(self._h)  # drains the queue first
            self._h = ctypes.c_void_p()
        self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_footprint(ds, **query):
    q = make_query(**query)
    n = ctypes.c_int64()
    _check(lib().idx2g_decode_footprint(ds._h, ctypes.byref(q), ctypes.byref(n)))
    return n.value
