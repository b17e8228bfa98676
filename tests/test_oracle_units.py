"""The oracle's pieces against what the mathematics fixes: filter norms,
lifting stencils on constants/ramps/impulses, the zfp lifting matrices, the
coefficient order rule, and the embedded coder by exhaustive round trips
and its closed-form length."""
import ctypes
import itertools

import numpy as np
import pytest


def _i3(v):
    return (ctypes.c_int * 3)(*v)


def test_cdf53_norms_match_filter_energies(oracle_lib):
    L = oracle_lib.lib()
    s = np.zeros(16)
    w = np.zeros(16)
    L.idx2o_cdf53_norms(s.ctypes.data, w.ctypes.data)
    # level 0: energies of the CDF 5/3 synthesis filters
    assert s[0] == pytest.approx(np.sqrt(0.25 + 1 + 0.25), rel=1e-15)
    assert w[0] == pytest.approx(np.sqrt((1 + 4 + 36 + 4 + 1) / 64), rel=1e-15)
    # closed forms of the recurrences: (2^(2i+3)+1)/3 / 2^(i+1) and (11+12*4^i)/2^(i+5)
    for i in range(16):
        assert s[i] == pytest.approx(np.sqrt(((2 ** (2 * i + 3) + 1) / 3) / 2 ** (i + 1)), rel=1e-14)
        assert w[i] == pytest.approx(np.sqrt((11 + 12 * 4 ** i) / 2 ** (i + 5)), rel=1e-14)


def _lift1d(O, line, fwd=True, M=None):
    """Lift a 1D line with the oracle's strided lifting (axis X)."""
    n = len(line)
    N = max(n, (M or n) + 1)
    vol = np.zeros(N, dtype=np.float64)
    vol[:n] = line
    m = M if M is not None else N - 1
    f = O.lib().idx2o_flift if fwd else O.lib().idx2o_ilift
    f(0, _i3((0, 0, 0)), _i3((n, 1, 1)), _i3((1, 1, 1)), _i3((m, 1, 1)), _i3((N, 1, 1)), vol.ctypes.data)
    return vol


def test_lift_constant_and_ramp(oracle_lib):
    for n in (3, 5, 9, 17, 33):
        c = _lift1d(oracle_lib, np.full(n, 2.5))
        assert np.all(c[1:n:2] == 0) and np.all(c[0:n:2] == 2.5)
        x = np.arange(n, dtype=np.float64)
        r = _lift1d(oracle_lib, 3.0 * x - 1.0)
        assert np.all(r[1:n:2] == 0)
        assert np.all(r[0:n:2] == (3.0 * x - 1.0)[0:n:2])


def test_lift_impulses_give_53_stencils(oracle_lib):
    n = 17
    odd = np.zeros(n); odd[7] = 1
    r = _lift1d(oracle_lib, odd)
    exp = np.zeros(n); exp[7] = 1; exp[6] = exp[8] = 0.25
    assert np.array_equal(r[:n], exp)
    even = np.zeros(n); even[8] = 1
    r = _lift1d(oracle_lib, even)
    exp = np.zeros(n); exp[6] = exp[10] = -0.125; exp[7] = exp[9] = -0.5; exp[8] = 0.75
    assert np.array_equal(r[:n], exp)


def test_even_length_extrapolates_linearly(oracle_lib):
    n = 16
    x = np.arange(n, dtype=np.float64)
    r = _lift1d(oracle_lib, 2.0 * x + 1.0, M=32)
    assert r[16] == 2.0 * 16 + 1.0      # extrapolated sample 2B - A
    assert np.all(r[1:16:2] == 0)       # every predict residual vanishes


def test_lift_inverse_roundtrip(oracle_lib):
    rng = np.random.default_rng(0)
    for n in (3, 9, 33):
        v = rng.standard_normal(n)
        f = _lift1d(oracle_lib, v)
        b = _lift1d(oracle_lib, f[:n], fwd=False)
        assert np.max(np.abs(b[:n] - v)) < 1e-14


def test_extrapolate_linear_field(oracle_lib):
    n3 = (33, 33, 33)
    z, y, x = np.meshgrid(np.arange(33.), np.arange(33.), np.arange(33.), indexing="ij")
    f = 0.5 * x - 0.25 * y + 0.125 * z + 3.0
    vol = np.where((x < 32) & (y < 32) & (z < 32), f, 0.0).copy()
    oracle_lib.lib().idx2o_extrapolate(_i3((32, 32, 32)), _i3(n3), vol.ctypes.data)
    assert np.max(np.abs(vol - f)) < 1e-12


def test_zfp_lifts_are_the_textbook_matrices(oracle_lib):
    L = oracle_lib.lib()
    Mf = np.array([[4, 4, 4, 4], [5, 1, -1, -5], [-4, 4, 4, -4], [-2, 6, -6, 2]]) / 16
    Mi = np.array([[4, 6, -4, -1], [4, 2, 4, 5], [4, -2, 4, -5], [4, -6, -4, 1]]) / 4
    rng = np.random.default_rng(1)
    for _ in range(200):
        v = rng.integers(-2 ** 20, 2 ** 20, 4) * 64
        p = v.astype(np.int64).copy()
        L.idx2o_zfp_lift(p.ctypes.data, 1)
        assert np.array_equal(p, (Mf @ v).astype(np.int64))
        q = (rng.integers(-2 ** 20, 2 ** 20, 4) * 64).astype(np.int64)
        p = q.copy()
        L.idx2o_zfp_lift(p.ctypes.data, 0)
        assert np.array_equal(p, (Mi @ q).astype(np.int64))


def test_coefficient_order_first_entries(oracle_lib):
    """A constant block has only the DC coefficient; a ramp along x, y or z
    adds exactly the first-degree coefficient of that axis, which the order
    places at positions 1, 2, 3 (degree 1 entries, x then y then z)."""
    L = oracle_lib.lib()
    k, j, i = np.meshgrid(np.arange(4.), np.arange(4.), np.arange(4.), indexing="ij")
    for field, pos in ((np.ones((4, 4, 4)), 0), (i, 1), (j, 2), (k, 3)):
        vals = (0.25 + 0.0625 * field).reshape(-1).copy()
        u = np.zeros(64, dtype=np.uint64)
        L.idx2o_block_forward(vals.ctypes.data, 3, 1, u.ctypes.data)
        nz = set(np.nonzero(u)[0].tolist())
        assert nz == ({0} if pos == 0 else {0, pos}), (pos, nz)


def test_block_roundtrip(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(5)
    for nd in (0, 1, 2, 3):
        nv = 1 << (2 * nd)
        vals = np.zeros(64)
        vals[:nv] = rng.standard_normal(nv)
        u = np.zeros(64, dtype=np.uint64)
        e = L.idx2o_block_forward(vals.ctypes.data, nd, 1, u.ctypes.data)
        back = np.zeros(64)
        L.idx2o_block_inverse(u.ctypes.data, nd, e, back.ctypes.data)
        assert np.max(np.abs(back[:nv] - vals[:nv])) < 1e-13


def _closed_form_len(X, N, nv):
    ln = N
    cur = N
    if cur >= nv:
        return ln, cur
    while True:
        rest = X >> cur if cur < 64 else 0
        if rest == 0:
            return ln + 1, cur
        t = (rest & -rest).bit_length() - 1
        q = cur + t
        ln += 1 + t + (1 if q < nv - 1 else 0)
        cur = q + 1
        if cur >= nv:
            return ln, cur


@pytest.mark.parametrize("nv", [1, 4, 16])
def test_coder_exhaustive_roundtrip(oracle_lib, nv):
    L = oracle_lib.lib()
    buf = np.zeros(64, dtype=np.uint8)
    for X in range(1 << nv):
        for N in range(nv + 1):
            blk = np.array([((X >> i) & 1) << 5 for i in range(64)], dtype=np.uint64)
            buf[:] = 0
            n = ctypes.c_int(N)
            nbits = L.idx2o_encode_plane(blk.ctypes.data, nv, 5, ctypes.byref(n), buf.ctypes.data, 3)
            ln, nout = _closed_form_len(X, N, nv)
            assert nbits == ln and n.value == nout
            out = np.zeros(64, dtype=np.uint64)
            m = ctypes.c_int(N)
            used = L.idx2o_decode_plane(out.ctypes.data, nv, 5, ctypes.byref(m), buf.ctypes.data, 3)
            assert used == nbits and m.value == nout
            # bits below N are sent verbatim, so every set bit comes back
            assert all(((int(out[i]) >> 5) & 1) == ((X >> i) & 1) for i in range(nv))


def test_coder_random_64(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(3)
    buf = np.zeros(64, dtype=np.uint8)
    for _ in range(2000):
        X = int(rng.integers(0, 2 ** 63)) * int(rng.integers(1, 3)) & (2 ** 64 - 1)
        X &= int(rng.integers(0, 2 ** 63))
        N = int(rng.integers(0, 65))
        blk = np.array([((X >> i) & 1) << 40 for i in range(64)], dtype=np.uint64)
        buf[:] = 0
        n = ctypes.c_int(N)
        nbits = L.idx2o_encode_plane(blk.ctypes.data, 64, 40, ctypes.byref(n), buf.ctypes.data, 0)
        assert (nbits, n.value) == _closed_form_len(X, N, 64)
        out = np.zeros(64, dtype=np.uint64)
        m = ctypes.c_int(N)
        L.idx2o_decode_plane(out.ctypes.data, 64, 40, ctypes.byref(m), buf.ctypes.data, 0)
        assert sum(((int(out[i]) >> 40) & 1) << i for i in range(64)) == X


def test_quantize_closed_form(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(4)
    for _ in range(500):
        v = float(rng.standard_normal() * 10.0 ** rng.integers(-8, 8))
        for is64 in (0, 1):
            vals = np.zeros(64)
            vals[0] = v
            u = np.zeros(64, dtype=np.uint64)
            e = L.idx2o_block_forward(vals.ctypes.data, 0, is64, u.ctypes.data)
            m = abs(v) if is64 else float(np.float32(abs(v)))
            exp_e = int(np.frexp(m)[1])
            assert e == exp_e
            q = int(np.trunc(v * 2.0 ** (62 - e)))
            mask = 0xAAAAAAAAAAAAAAAA
            assert int(u[0]) == ((q + mask) ^ mask) & (2 ** 64 - 1)
    # zero block: exponent -bias
    vals = np.zeros(64)
    u = np.zeros(64, dtype=np.uint64)
    assert L.idx2o_block_forward(vals.ctypes.data, 3, 0, u.ctypes.data) == -127
    assert L.idx2o_block_forward(vals.ctypes.data, 3, 1, u.ctypes.data) == -1023


def test_delayed_transpose_branch_is_unreachable():
    """The reference's DecodeSubband (Source/Core/idx2Decode.cpp:140-203)
    decodes a plane with zfp Decode (direct deposit) when
    NBitPlanesDecoded = Exponent(Tolerance) - 6 - EMax + 1 is at most 8, and
    otherwise with DecodeTest + TransposeRecursive (Zfp.h:960-985,
    1139-1708).  With one bit plane per chunk the plane loop breaks at the
    first plane whose RealBp = Bp + EMax is TooHighPrecision
    (NBitPlanes - 6 > RealBp - Exponent(Tolerance) + 1), so every decoded plane
    has EMax >= Exponent(Tolerance) - 6 and NBitPlanesDecoded <= 1: the second
    branch never decodes a plane, which is why the oracle and the GPU path
    implement only the direct deposit.  Checked here by running the
    reference's loop bounds over every exponent a float32 / float64 block or
    a tolerance can have."""
    nbp = 64  # the first plane, Bp = 63, is in range for float32 and float64 data alike
    for etol in range(-1080, 40):  # Exponent() is frexp's, clamped to [-1023, 1024] (Math.h:523-534)
        for emax in range(-1080, 1030):
            n_decoded = etol - 6 - emax + 1
            first_plane_decoded = not (nbp - 6 > (nbp - 1 + emax) - etol + 1)
            if first_plane_decoded:
                assert n_decoded <= 8, (etol, emax)
            else:  # the loop breaks before any Decode call (BitPlanesPerChunk = 1)
                assert emax < etol - 6
    # the tolerances of the workloads: frexp exponents as Exponent() takes them
    import math
    for tol in (1e-2, 1e-3, 1e-4, 1e-6, 1e-9):
        etol = math.frexp(tol)[1]
        assert all(etol - 6 - e + 1 <= 8 for e in range(etol - 6, 1030))
