"""Test helper: split an idx2 L*.bin file into its sections (layout of
Source/Core/idx2Write.cpp:96-385) so two files can be compared section by
section, with zstd-compressed sections compared after decompression (the
system libzstd.so.1 may differ in compressed bytes from the zstd 1.4.5 the
reference vendors; the decompressed content is what must match)."""
import ctypes
import struct

_z = None


def _zstd():
    global _z
    if _z is None:
        z = ctypes.CDLL("libzstd.so.1")
        z.ZSTD_getFrameContentSize.restype = ctypes.c_ulonglong
        z.ZSTD_getFrameContentSize.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        z.ZSTD_decompress.restype = ctypes.c_size_t
        z.ZSTD_decompress.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
        _z = z
    return _z


def unzstd(b: bytes) -> bytes:
    z = _zstd()
    n = z.ZSTD_getFrameContentSize(b, len(b))
    out = ctypes.create_string_buffer(max(int(n), 1))
    r = z.ZSTD_decompress(out, n, b, len(b))
    assert r == n, (r, n)
    return out.raw[:n]


def _varbytes(b: bytes, count=None):
    vals, v, sh, i = [], 0, 0, 0
    while i < len(b) and (count is None or len(vals) < count):
        byte = b[i]
        v |= (byte & 0x7F) << sh
        sh += 7
        i += 1
        if not byte & 0x80:
            vals.append(v)
            v, sh = 0, 0
    return vals


def sections(data: bytes):
    """Return dict with bit-plane chunks, their addresses, exponent chunks
    (decompressed) and their addresses."""
    pos = len(data)

    def rd():
        nonlocal pos
        pos -= 4
        return struct.unpack("<i", data[pos:pos + 4])[0]

    S = rd()
    expbeg = len(data) - S
    n = rd()
    asz = rd()
    pos -= asz
    eaddr = struct.unpack(f"<{n}Q", unzstd(data[pos:pos + asz])) if asz else ()
    ssz = rd()
    pos -= ssz
    esizes = _varbytes(data[pos:pos + ssz], n)
    exps, off = [], expbeg
    for s in esizes:
        exps.append(unzstd(data[off:off + s]))
        off += s
    pos = expbeg
    nch = rd()
    casz = rd()
    pos -= casz
    caddr = struct.unpack(f"<{nch}Q", unzstd(data[pos:pos + casz])) if casz else ()
    cssz = rd()
    pos -= cssz
    csizes = _varbytes(data[pos:pos + cssz], nch)
    chunks, off = [], 0
    for s in csizes:
        chunks.append(data[off:off + s])
        off += s
    return dict(chunks=chunks, chunk_addrs=list(caddr), exps=exps, exp_addrs=list(eaddr),
                chunk_region=data[:off])


def same_content(a: bytes, b: bytes):
    """Files equal up to the bytes of their zstd-compressed sections."""
    sa, sb = sections(a), sections(b)
    return (sa["chunks"] == sb["chunks"] and sa["chunk_addrs"] == sb["chunk_addrs"]
            and sa["exps"] == sb["exps"] and sa["exp_addrs"] == sb["exp_addrs"])
