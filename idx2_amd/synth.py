"""Deterministic synthetic scalar fields for the idx2 encode/decode workloads.

Field (with x = (i + 0.5) / N_d on each axis, x_z = 0.5 for 2D grids):

    f(x) = sum_{g<32} a_g exp(-|x - c_g|^2 / (2 s_g^2))
         + 0.1 sum_{j<8} sin(2 pi k_j . x + phi_j)
         + eps(x)

    c_g ~ U[0,1]^3, s_g ~ U[0.02, 0.2], a_g ~ N(0,1)
    k_j integer components ~ U{-8..8} (not all zero), phi_j ~ U[0, 2 pi)
    eps ~ N(0, 1e-3) per sample

Every random number comes from a counter-based hash (splitmix64 finaliser of
seed and counter), and normal deviates are drawn as sum of 12 uniforms minus 6
(Irwin-Hall), so no library RNG is involved.  exp and sin are evaluated with
Cody-Waite range reduction and Taylor polynomials written with IEEE +, -, *, /
only, in a fixed operation order.  The same arithmetic runs in the HIP
generator kernel (idx2_amd/csrc/synth.hip) compiled with -ffp-contract=off, so
the bytes produced on the CPU and on the GPU are identical.
"""
import numpy as np

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB

NOISE_SIGMA = 1e-3
NGAUSS = 32
NSINES = 8

# Cody-Waite constants (split so that n * HI is exact for |n| < 2^20)
LN2_HI = 6.93147180369123816490e-01
LN2_LO = 1.90821492927058770002e-10
INV_LN2 = 1.44269504088896338700e+00
PIO2_1 = 1.57079632673412561417e+00
PIO2_2 = 6.07710050650619224932e-11
PIO2_3 = 2.02226624879595063154e-21
TWO_OVER_PI = 6.36619772367581382433e-01
TWO_PI = 6.28318530717958647693e+00


def _mix(x: int) -> int:
    x &= MASK64
    x = ((x ^ (x >> 30)) * M1) & MASK64
    x = ((x ^ (x >> 27)) * M2) & MASK64
    return x ^ (x >> 31)


def _u01_scalar(seed: int, ctr: int) -> float:
    h = _mix((seed * GOLDEN + ctr) & MASK64)
    return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0)


def _normal_scalar(seed: int, ctr: int) -> float:
    s = 0.0
    for i in range(12):
        s = s + _u01_scalar(seed, ctr * 12 + i)
    return s - 6.0


def _mix_np(x):
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(M1)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(M2)
    return x ^ (x >> np.uint64(31))


def _u01_np(seed: int, ctr):
    with np.errstate(over="ignore"):
        key = np.uint64((seed * GOLDEN) & MASK64) + ctr.astype(np.uint64)
    h = _mix_np(key)
    return ((h >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def exp_neg(x):
    """exp(x) for x <= 0 (arrays), basic ops only; exactly 0 below -700."""
    x = np.asarray(x, dtype=np.float64)
    n = np.rint(x * INV_LN2)
    r = (x - n * LN2_HI) - n * LN2_LO
    p = 1.0 / 479001600.0
    for c in (39916800.0, 3628800.0, 362880.0, 40320.0, 5040.0, 720.0, 120.0, 24.0, 6.0, 2.0, 1.0, 1.0):
        p = p * r + 1.0 / c
    res = np.ldexp(p, np.maximum(n, -1000).astype(np.int32))
    return np.where(x < -700.0, 0.0, res)


def sin_(t):
    """sin(t) for |t| < 2^20 (arrays), basic ops only."""
    t = np.asarray(t, dtype=np.float64)
    q = np.rint(t * TWO_OVER_PI)
    r = ((t - q * PIO2_1) - q * PIO2_2) - q * PIO2_3
    r2 = r * r
    # sin(r) = r (1 - r^2/3! + r^4/5! - ...) up to r^15
    s = -1.0 / 1307674368000.0
    for c in (6227020800.0, -39916800.0, 362880.0, -5040.0, 120.0, -6.0):
        s = s * r2 + 1.0 / c
    s = s * r2 + 1.0
    s = s * r
    # cos(r) = 1 - r^2/2! + ... up to r^16
    c_ = 1.0 / 20922789888000.0
    for c in (-87178291200.0, 479001600.0, -3628800.0, 40320.0, -720.0, 24.0, -2.0):
        c_ = c_ * r2 + 1.0 / c
    c_ = c_ * r2 + 1.0
    qi = np.mod(q, 4.0)
    return np.where(qi == 0, s, np.where(qi == 1, c_, np.where(qi == 2, -s, -c_)))


def field_params(seed: int) -> np.ndarray:
    """Parameter table (float64) shared by the CPU and GPU generators.

    Layout: 32 Gaussians x [cx, cy, cz, 1/(2 s^2), a], then 8 sines x
    [kx, ky, kz, phi].  Total 32*5 + 8*4 = 192 doubles.
    """
    p = []
    ctr = 0
    for g in range(NGAUSS):
        cx = _u01_scalar(seed, ctr); cy = _u01_scalar(seed, ctr + 1); cz = _u01_scalar(seed, ctr + 2)
        s = 0.02 + 0.18 * _u01_scalar(seed, ctr + 3)
        a = _normal_scalar(seed, (1 << 20) + g)
        ctr += 4
        p += [cx, cy, cz, 1.0 / (2.0 * s * s), a]
    for j in range(NSINES):
        while True:
            k = [float(int(_u01_scalar(seed, ctr + d) * 17.0) - 8) for d in range(3)]
            ctr += 3
            if any(v != 0.0 for v in k):
                break
        phi = TWO_PI * _u01_scalar(seed, ctr)
        ctr += 1
        p += k + [phi]
    return np.array(p, dtype=np.float64)


def make_field(dims, dtype="float32", seed=0, z_range=None):
    """Generate the field as a C-ordered array of shape (nz, ny, nx) (x fastest).

    dims = (nx, ny, nz).  z_range=(z0, z1) generates only slabs z0..z1-1 (for
    bounded samples of large grids).
    """
    nx, ny, nz = (int(d) for d in dims)
    z0, z1 = (0, nz) if z_range is None else z_range
    prm = field_params(seed)
    ix = np.arange(nx, dtype=np.float64)
    iy = np.arange(ny, dtype=np.float64)
    iz = np.arange(z0, z1, dtype=np.float64)
    x = ((ix + 0.5) / float(nx))[None, None, :]
    y = ((iy + 0.5) / float(ny))[None, :, None]
    z = ((iz + 0.5) / float(nz))[:, None, None]
    shape = (z1 - z0, ny, nx)
    acc = np.zeros(shape, dtype=np.float64)
    for g in range(NGAUSS):
        cx, cy, cz, inv2s2, a = prm[5 * g: 5 * g + 5]
        dx = x - cx
        dy = y - cy
        dz = z - cz
        d2 = (dx * dx + dy * dy) + dz * dz
        acc = acc + a * exp_neg(-(d2 * inv2s2))
    s = np.zeros(shape, dtype=np.float64)
    base = 5 * NGAUSS
    for j in range(NSINES):
        kx, ky, kz, phi = prm[base + 4 * j: base + 4 * j + 4]
        arg = TWO_PI * ((kx * x + ky * y) + kz * z) + phi
        s = s + sin_(arg)
    acc = acc + 0.1 * s
    # noise: counter = linear sample index, 12 uniforms per sample
    lin = (np.arange(z0, z1, dtype=np.uint64)[:, None, None] * np.uint64(ny * nx)
           + np.arange(ny, dtype=np.uint64)[None, :, None] * np.uint64(nx)
           + np.arange(nx, dtype=np.uint64)[None, None, :])
    nsum = np.zeros(shape, dtype=np.float64)
    nseed = seed + 0x5EED
    for i in range(12):
        nsum = nsum + _u01_np(nseed, lin * np.uint64(12) + np.uint64(i))
    acc = acc + NOISE_SIGMA * (nsum - 6.0)
    return acc.astype(np.float32 if dtype in ("float32", np.float32) else np.float64)
