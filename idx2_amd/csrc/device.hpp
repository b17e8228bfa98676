// Device building blocks for the idx2 path on gfx950 (wave64).
//
// Everything that feeds the integer stream is bit-exact with the reference:
// FP64 lifting in the reference's operation order (this file must be compiled
// with -ffp-contract=off), exact f64 <-> i64 conversions written out below,
// and the zfp-style integer stage.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace idx2g {

// Phase clocks of the on-chip encode kernel (profiling builds only, compiled
// with -DIDX2G_PHASE_TIMERS; see idx2g_phase_times): thread 0 of each
// workgroup adds the wall-clock ticks (100 MHz) since its previous mark, taken
// right after a workgroup barrier, to the phase's counter.
#ifdef IDX2G_PHASE_TIMERS
__device__ unsigned long long g_phase_ticks[32];
__shared__ long long g_phase_prev;
#define IDX2G_PHASE(n)                                                                          \
  do {                                                                                          \
    if (threadIdx.x == 0) {                                                                     \
      const long long t_ = wall_clock64();                                                      \
      atomicAdd(&::idx2g::g_phase_ticks[n], (unsigned long long)(t_ - ::idx2g::g_phase_prev)); \
      ::idx2g::g_phase_prev = t_;                                                               \
    }                                                                                           \
  } while (0)
#define IDX2G_PHASE_START()                                         \
  do {                                                              \
    if (threadIdx.x == 0) ::idx2g::g_phase_prev = wall_clock64();   \
  } while (0)
#else
#define IDX2G_PHASE(n) \
  do {                 \
  } while (0)
#define IDX2G_PHASE_START() \
  do {                      \
  } while (0)
#endif

typedef uint64_t u64;
typedef int64_t i64;
typedef uint32_t u32;

#define IDX2G_NBMASK 0xaaaaaaaaaaaaaaaaull  // negabinary mask, traits<u64>::NBinaryMask
constexpr int kBpBiasDev = 1024;        // BitPlaneKeyBias_, Source/Core/idx2Lookup.h:11

struct DGrid {
  int f[3], d[3], s[3];
};

// Constants of one dataset shared by all kernels.
struct BrickConst {
  int N[3];        // extended brick dims (33^3, 33x33x1, 17^3, ...)
  int M[3];        // storage limit of the extrapolation: N-1 (or 1)
  int nsb;         // subbands per brick
  DGrid sb[8];     // subband grids inside the extended brick
  int ext_levels;  // Log2Floor(max N): levels of ExtrapolateCdf53
  u64 order;       // transform order, base 4 ("XYZ++")
  int dtype;       // 0 f32, 1 f64
  int endbp;       // 32 (f32) or 64 (f64): planes Bp >= 64 - endbp are coded
  int nvox;        // N[0]*N[1]*N[2]
};

__device__ __forceinline__ i64 vox(const BrickConst& k, int x, int y, int z) {
  return ((i64)z * k.N[1] + y) * k.N[0] + x;
}

// --------------------------------------------------------------------------
// CDF 5/3 lifting of one line, reference operation order
// (FLiftCdf53 / ILiftCdf53, Source/Core/Wavelet.h:177-310, lift_option Normal).
// L points at the line start (position 0 of the lifting axis), sd is the
// element stride of the axis, P/D/S the line grid and M the storage limit.
// --------------------------------------------------------------------------
__device__ __forceinline__ void flift_line(double* __restrict__ L, long sd, int P, int D, int S, int M) {
  const int x0 = min(P + S * D, M), x1 = min(P + S * (D - 1), M);
  const int x2 = P + S * (D - 2), x3 = P + S * (D - 3);
  const bool ext = (D & 1) == 0;
  if (ext) {
    double A = L[x2 * sd], B = L[x1 * sd];
    L[x0 * sd] = 2 * B - A;
  }
  for (int x = P + S; x < P + S * (D - 2); x += 2 * S) {
    double l = L[(x - S) * sd], r = L[(x + S) * sd];
    L[x * sd] = L[x * sd] - (l + r) / 2;
  }
  if (!ext) {
    double l = L[x1 * sd], r = L[x3 * sd];
    L[x2 * sd] = L[x2 * sd] - (l + r) / 2;
  } else if (x1 < M) {
    L[x1 * sd] = 0;
  }
  for (int x = P + S; x < P + S * (D - 2); x += 2 * S) {
    double v = L[x * sd];
    L[(x - S) * sd] = L[(x - S) * sd] + v / 4;
    L[(x + S) * sd] = L[(x + S) * sd] + v / 4;
  }
  if (!ext) {
    double v = L[x2 * sd];
    L[x3 * sd] = L[x3 * sd] + v / 4;
    L[x1 * sd] = L[x1 * sd] + v / 4;
  }
}

__device__ __forceinline__ void ilift_line(double* __restrict__ L, long sd, int P, int D, int S, int M) {
  const int x0 = min(P + S * D, M), x1 = min(P + S * (D - 1), M);
  const int x2 = P + S * (D - 2), x3 = P + S * (D - 3);
  const bool ext = (D & 1) == 0;
  for (int x = P + S; x < P + S * (D - 2); x += 2 * S) {
    double v = L[x * sd];
    L[(x - S) * sd] = L[(x - S) * sd] - v / 4;
    L[(x + S) * sd] = L[(x + S) * sd] - v / 4;
  }
  if (!ext) {
    double v = L[x2 * sd];
    L[x3 * sd] = L[x3 * sd] - v / 4;
    L[x1 * sd] = L[x1 * sd] - v / 4;
  } else {
    double A = L[x0 * sd], B = L[x2 * sd];
    L[x1 * sd] = (A + B) / 2;
  }
  for (int x = P + S; x < P + S * (D - 2); x += 2 * S) {
    double l = L[(x - S) * sd], r = L[(x + S) * sd];
    L[x * sd] = L[x * sd] + (l + r) / 2;
  }
  if (!ext) {
    double l = L[x1 * sd], r = L[x3 * sd];
    L[x2 * sd] = L[x2 * sd] + (l + r) / 2;
  }
}

// The same two liftings with one line point per lane: lane i holds point i
// of its line (i = D: the extrapolated point of an even-length line), lanes
// of one line are consecutive.  Every lane of the wave calls these (DPP
// reads); lanes outside a line pass any value and are ignored.  The
// operation order per point is flift_line's / ilift_line's.
__device__ __forceinline__ double dpp_prev_f64(double v) {  // lane - 1 (DPP wave_shr:1)
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double dpp_next_f64(double v) {  // lane + 1 (DPP wave_shl:1)
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x130, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double flift_lanes(double v, int i, int D) {
  const bool ext = (D & 1) == 0;
  const int hi_odd = ext ? D - 3 : D - 2;  // last odd point of the predict/update loops
  double l = dpp_prev_f64(v);
  const double ll = dpp_prev_f64(l);
  if (ext && i == D) v = 2 * l - ll;  // L[x0] = 2 B - A
  l = dpp_prev_f64(v);
  double r = dpp_next_f64(v);
  if ((i & 1) && i <= hi_odd) v = v - (l + r) / 2;
  if (ext && i == D - 1) v = 0;
  l = dpp_prev_f64(v);
  r = dpp_next_f64(v);
  if (!(i & 1) && i <= D - 1) {
    if (i >= 1 && i - 1 <= hi_odd) v = v + l / 4;
    if (i + 1 <= hi_odd) v = v + r / 4;
  }
  return v;
}

__device__ __forceinline__ double ilift_lanes(double v, int i, int D) {
  const bool ext = (D & 1) == 0;
  const int hi_odd = ext ? D - 3 : D - 2;
  double l = dpp_prev_f64(v), r = dpp_next_f64(v);
  if (!(i & 1) && i <= D - 1) {
    if (i >= 1 && i - 1 <= hi_odd) v = v - l / 4;
    if (i + 1 <= hi_odd) v = v - r / 4;
  }
  l = dpp_prev_f64(v);
  r = dpp_next_f64(v);
  if (ext && i == D - 1) v = (r + l) / 2;  // L[x1] = (L[x0] + L[x2]) / 2
  if ((i & 1) && i <= hi_odd) v = v + (l + r) / 2;
  return v;
}

// Lifts the reference applies to the line at logical index i of axis ax of
// grid g: its loops clamp the coordinates of the other two axes to the
// storage limit M (Wavelet.h:195-199, Min(y, M.y)), so the logical lines
// past M all lift the line at M, one after the other (on bricks whose axes
// shrink at different rates, at the coarse levels of the extrapolation).
// The first of them returns their count, the others 0, the rest 1.
__device__ __forceinline__ int line_lifts(const DGrid& g, int ax, int i, const int M[3]) {
  if (g.f[ax] + g.s[ax] * i < M[ax]) return 1;
  if (i > 0 && g.f[ax] + g.s[ax] * (i - 1) >= M[ax]) return 0;
  return g.d[ax] - i;
}

// One lifting pass of a whole grid along axis d.  The other two axes follow
// the reference macros: X -> loops (Y, Z), Y -> (X, Z), Z -> (X, Y);
// coordinates clamp to M, and a line that several logical lines clamp to is
// lifted once per logical line, in sequence.  Lines of at most 64 points that
// never reach the storage limit run one point per lane, several lines per
// wave; others one line per thread.
__device__ __forceinline__ void lift_pass(const BrickConst& k, double* F, bool fwd, int d, const DGrid& g,
                                          const int M[3]) {
  if (g.d[d] == 1) return;
  const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
  const long st[3] = {1, (long)k.N[0], (long)k.N[0] * k.N[1]};
  const int nl = g.d[a] * g.d[b];
  const int D = g.d[d], Dn = D + ((D & 1) == 0);
  if (Dn <= 64 && g.f[d] + g.s[d] * (Dn - 1) <= M[d]) {
    const int per = 64 / Dn, lane = __lane_id(), slot = lane / Dn, i = lane - slot * Dn;
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j0 = wave * per; j0 < nl; j0 += nw * per) {
      const int li = j0 + slot;
      bool act = slot < per && li < nl;
      int reps = 0;
      double* q = F;
      if (act) {
        const int ia = li % g.d[a], ib = li / g.d[a];
        reps = line_lifts(g, a, ia, M) * line_lifts(g, b, ib, M);
        const int ca = min(g.f[a] + g.s[a] * ia, M[a]);
        const int cb = min(g.f[b] + g.s[b] * ib, M[b]);
        q = F + ca * st[a] + cb * st[b] + (long)(g.f[d] + g.s[d] * i) * st[d];
        act = reps > 0;
      }
      double v = act ? *q : 0.0;
      int rmax = reps;  // over the wave (the lifts' DPP reads need every lane)
      for (int o = 32; o > 0; o >>= 1) rmax = max(rmax, __shfl_xor(rmax, o, 64));
      for (int r = 0; r < rmax; ++r) {
        const double u = fwd ? flift_lanes(v, i, D) : ilift_lanes(v, i, D);
        v = r < reps ? u : v;
      }
      if (act) *q = v;
    }
    return;
  }
  for (int li = threadIdx.x; li < nl; li += blockDim.x) {
    int ia = li % g.d[a], ib = li / g.d[a];
    const int reps = line_lifts(g, a, ia, M) * line_lifts(g, b, ib, M);
    int ca = min(g.f[a] + g.s[a] * ia, M[a]);
    int cb = min(g.f[b] + g.s[b] * ib, M[b]);
    double* L = F + ca * st[a] + cb * st[b];
    for (int r = 0; r < reps; ++r) {
      if (fwd) flift_line(L, st[d], g.f[d], g.d[d], g.s[d], M[d]);
      else ilift_line(L, st[d], g.f[d], g.d[d], g.s[d], M[d]);
    }
  }
}

// Multi-level transform with the run-time order (ForwardCdf53 /
// InverseCdf53 without normalisation, Source/Core/Wavelet.cpp:209-404):
// level k lifts the axes of the order on grid G, whose extent grows by one
// extrapolated sample on even extents before the next axis.  step(i) replays
// the order to the i-th lifting (no per-thread stack); returns the count.
__device__ __forceinline__ int runtime_step(const BrickConst& k, const int dims3[3], int nlevels, int want,
                                            int* axis, DGrid* out) {
  int D3[3] = {dims3[0], dims3[1], dims3[2]}, R3[3] = {dims3[0], dims3[1], dims3[2]}, S3[3] = {1, 1, 1};
  DGrid G;
  for (int d = 0; d < 3; ++d) { G.f[d] = 0; G.d[d] = dims3[d]; G.s[d] = 1; }
  u64 order = k.order, prev = order;
  int level = 0, n = 0;
  while (level < nlevels) {
    int D = int(order & 3);
    order >>= 2;
    if (D == 3) {
      if (order == 3) order = prev; else prev = order;
      for (int d = 0; d < 3; ++d) { G.s[d] = S3[d]; G.d[d] = D3[d]; R3[d] = D3[d]; }
      ++level;
    } else {
      if (n == want) { *axis = D; *out = G; }
      ++n;
      R3[D] = D3[D] + ((D3[D] & 1) == 0);
      for (int d = 0; d < 3; ++d) G.d[d] = R3[d];
      D3[D] = (R3[D] + 1) >> 1;
      S3[D] <<= 1;
    }
  }
  return n;
}

__device__ void transform_runtime(const BrickConst& k, double* F, bool fwd, const int dims3[3], const int M[3],
                                  int nlevels) {
  int ax = 0;
  DGrid g;
  const int n = runtime_step(k, dims3, nlevels, -1, &ax, &g);
  for (int j = 0; j < n; ++j) {
    int i = fwd ? j : n - 1 - j;
    runtime_step(k, dims3, nlevels, i, &ax, &g);
    lift_pass(k, F, fwd, ax, g, M);
    __syncthreads();
  }
}

// ExtrapolateCdf53 (Source/Core/Wavelet.cpp:67-77): forward levels on the
// data extent, inverse levels on the full extended brick.
__device__ __forceinline__ void extrapolate(const BrickConst& k, double* F, const int data3[3]) {
  transform_runtime(k, F, true, data3, k.M, k.ext_levels);
  transform_runtime(k, F, false, k.N, k.M, k.ext_levels);
}

// One-level transform of the extended brick (subband ForwardCdf53 /
// InverseCdf53, Wavelet.cpp:80-203): lifting on the full extent (odd, so no
// extrapolation) along X,Y,Z (inverse Z,Y,X), then per-subband weights.
__device__ __forceinline__ void forward_brick(const BrickConst& k, double* F, const double* w, int coarsest) {
  DGrid g;
  for (int d = 0; d < 3; ++d) { g.f[d] = 0; g.d[d] = k.N[d]; g.s[d] = 1; }
  for (int d = 0; d < 3; ++d) {
    if (k.N[d] == 1) continue;
    lift_pass(k, F, true, d, g, k.N);
    __syncthreads();
  }
  for (int s = coarsest ? 0 : 1; s < k.nsb; ++s) {
    const DGrid& sg = k.sb[s];
    const int n = sg.d[0] * sg.d[1] * sg.d[2];
    const double W = w[s];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      int x = i % sg.d[0], y = (i / sg.d[0]) % sg.d[1], z = i / (sg.d[0] * sg.d[1]);
      i64 v = vox(k, sg.f[0] + sg.s[0] * x, sg.f[1] + sg.s[1] * y, sg.f[2] + sg.s[2] * z);
      F[v] = F[v] * W;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void inverse_brick(const BrickConst& k, double* F, const double* winv, int coarsest) {
  for (int s = coarsest ? 0 : 1; s < k.nsb; ++s) {
    const DGrid& sg = k.sb[s];
    const int n = sg.d[0] * sg.d[1] * sg.d[2];
    const double W = winv[s];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      int x = i % sg.d[0], y = (i / sg.d[0]) % sg.d[1], z = i / (sg.d[0] * sg.d[1]);
      i64 v = vox(k, sg.f[0] + sg.s[0] * x, sg.f[1] + sg.s[1] * y, sg.f[2] + sg.s[2] * z);
      F[v] = F[v] * W;
    }
  }
  __syncthreads();
  DGrid g;
  for (int d = 0; d < 3; ++d) { g.f[d] = 0; g.d[d] = k.N[d]; g.s[d] = 1; }
  for (int d = 2; d >= 0; --d) {
    if (k.N[d] == 1) continue;
    lift_pass(k, F, false, d, g, k.N);
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// exact conversions and exponents
// --------------------------------------------------------------------------
// Exponent<f64> (Source/Core/Math.h:523-534): frexp exponent, clamped at the
// denormal bound; -1023 for zero.
__device__ __forceinline__ int exponent_f64(double v) {
  if (v == 0) return -1023;
  int e = int((__double_as_longlong(v) >> 52) & 0x7ff);
  return e == 0 ? -1022 : e - 1022;
}
__device__ __forceinline__ int exponent_f32(float v) {
  if (v == 0) return -127;
  int e = int((__float_as_uint(v) >> 23) & 0xff);
  return e == 0 ? -126 : e - 126;
}
// 2^e as a double (0 below the subnormal range, inf above)
__device__ __forceinline__ double pow2d(int e) {
  if (e > 1023) return __longlong_as_double(0x7ff0000000000000ll);
  if (e >= -1022) return __longlong_as_double((i64)(e + 1023) << 52);
  if (e >= -1074) return __longlong_as_double(1ll << (e + 1074));
  return 0.0;
}
// 1.0 / pow2d(k) without the division (Dequantize's Scale, Function.h:146):
// for k in [-1023, 1023] the quotient is exactly 2^-k (2^-1023 subnormal);
// k > 1023 divides by inf and gives 0; k < -1023 divides by a subnormal or
// by 0 and gives inf, as pow2d(-k) does.
__device__ __forceinline__ double inv_pow2d(int k) { return k > 1023 ? 0.0 : pow2d(-k); }
// (i64)x for |x| < 2^63: truncation toward zero, split exactly into 32-bit
// halves (t and hi*2^32 are exact, so lo is exact).
// NaN and |x| >= 2^63 give 0x8000000000000000 (x86's cvttsd2si); computed
// without a branch (the hardware conversions saturate, and the select drops
// their result for such x).
__device__ __forceinline__ i64 f64_to_i64(double x) {
  const double t = trunc(x);
  const double hi = floor(t * 2.3283064365386963e-10);
  const double lo = t - hi * 4294967296.0;
  const i64 r = (i64)(((u64)(u32)__double2int_rz(hi) << 32) | (u64)__double2uint_rz(lo));
  return fabs(x) < 9.2233720368547758e18 ? r : (i64)0x8000000000000000ll;
}
// (double)q, correctly rounded: hi*2^32 is exact and the sum rounds once.
__device__ __forceinline__ double i64_to_f64(i64 q) {
  return (double)(int)(q >> 32) * 4294967296.0 + (double)(u32)(q & 0xffffffffu);
}

// --------------------------------------------------------------------------
// wave64 helpers
// --------------------------------------------------------------------------
__device__ __forceinline__ i64 shfl_i64(i64 v, int src) {
  int lo = __shfl((int)(u32)v, src, 64);
  int hi = __shfl((int)(u32)((u64)v >> 32), src, 64);
  return (i64)(((u64)(u32)hi << 32) | (u32)lo);
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = v < w ? w : v;
  }
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
// Global-memory view of a pointer.  Loads and stores through it use the
// global instructions, which count only against vmcnt; the flat ones used
// for a pointer of unknown address space also count against lgkmcnt, so a
// later wait for an LDS result (every barrier has one) would wait for them.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gm(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}
// Store of a trivially copyable struct through the global view.
template <class T>
__device__ __forceinline__ void gm_store(T* p, const T& v) {
  static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "32-bit words");
  auto d = gm(reinterpret_cast<uint32_t*>(p));
  const uint32_t* s = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) d[i] = s[i];
}

__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
  return v;
}
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// --------------------------------------------------------------------------
// zfp-style block transform with one coefficient per lane
// (FLift/ILift, ForwardZfp/InverseZfp, Source/Core/Zfp.h:144-345).
// Lane index = x + 4y + 16z over the block's non-unit axes.
// --------------------------------------------------------------------------
__device__ __forceinline__ void zfp_flift4(i64& x, i64& y, i64& z, i64& w) {
  x += w; x >>= 1; w -= x;
  z += y; z >>= 1; y -= z;
  x += z; x >>= 1; z -= x;
  w += y; w >>= 1; y -= w;
  w += y >> 1; y -= w >> 1;
}
__device__ __forceinline__ void zfp_ilift4(i64& x, i64& y, i64& z, i64& w) {
  y += w >> 1; w -= y >> 1;
  y += w; w <<= 1; w -= y;
  z += x; x <<= 1; x -= z;
  y += z; z <<= 1; z -= y;
  w += x; x <<= 1; x -= w;
}
// lift along the axis whose lane stride is `st` (1, 4 or 16)
__device__ __forceinline__ i64 lane_lift(i64 v, int lane, int st, bool fwd) {
  int pos = (lane / st) & 3;
  int base = lane - pos * st;
  i64 a = shfl_i64(v, base), b = shfl_i64(v, base + st), c = shfl_i64(v, base + 2 * st),
      e = shfl_i64(v, base + 3 * st);
  if (fwd) zfp_flift4(a, b, c, e);
  else zfp_ilift4(a, b, c, e);
  return pos == 0 ? a : pos == 1 ? b : pos == 2 ? c : e;
}
__device__ __forceinline__ i64 zfp_forward_lanes(i64 v, int lane, int nd) {
  if (nd >= 1) v = lane_lift(v, lane, 1, true);
  if (nd >= 2) v = lane_lift(v, lane, 4, true);
  if (nd >= 3) v = lane_lift(v, lane, 16, true);
  return v;
}
__device__ __forceinline__ i64 zfp_inverse_lanes(i64 v, int lane, int nd) {
  if (nd >= 3) v = lane_lift(v, lane, 16, false);
  if (nd >= 2) v = lane_lift(v, lane, 4, false);
  if (nd >= 1) v = lane_lift(v, lane, 1, false);
  return v;
}

// Coefficient orders (Perm3/Perm2, Zfp.h:414-521) and their inverses, filled
// by the host at load time.
struct PermTables {
  unsigned char p3[64], ip3[64], p2[16], ip2[16];
};

// --------------------------------------------------------------------------
// embedded coder of one bit plane (Encode/Decode, Zfp.h:769-813, 988-1059),
// restated with 64-bit windows: plane word X, NV values, N significant so far.
// --------------------------------------------------------------------------
// number of bits the plane emits; updates n
__device__ __forceinline__ int plane_len(u64 X, int& n, int nv) {
  int len = n;
  int cur = n;
  if (cur >= nv) return len;
  u64 rest = cur >= 64 ? 0 : X >> cur;
  for (;;) {
    if (rest == 0) { len += 1; break; }
    int t = __ffsll((long long)rest) - 1;  // zeros before the next one
    int q = cur + t;
    len += 1 + t + (q < nv - 1 ? 1 : 0);
    cur = q + 1;
    rest = (t + 1 >= 64) ? 0 : rest >> (t + 1);
    if (cur >= nv) break;
  }
  n = cur;
  return len;
}

// Emit the plane's bits (LSB first) into w[0..2]; returns the length.
__device__ __forceinline__ int plane_emit(u64 X, int n, int nv, u64 w[3]) {
  w[0] = w[1] = w[2] = 0;
  int pos = 0;
  auto put = [&](u64 bits, int cnt) {  // cnt <= 64
    if (cnt <= 0) return;
    if (cnt < 64) bits &= (1ull << cnt) - 1;
    int wi = pos >> 6, off = pos & 63;
    w[wi] |= bits << off;
    if (off && off + cnt > 64) w[wi + 1] |= bits >> (64 - off);
    pos += cnt;
  };
  put(X, n);
  int cur = n;
  if (cur >= nv) return pos;
  u64 rest = cur >= 64 ? 0 : X >> cur;
  for (;;) {
    if (rest == 0) { put(0, 1); break; }
    int t = __ffsll((long long)rest) - 1;
    int q = cur + t;
    put(1, 1);          // group is significant
    pos += t;           // t zeros
    if (q < nv - 1) put(1, 1);
    cur = q + 1;
    rest = (t + 1 >= 64) ? 0 : rest >> (t + 1);
    if (cur >= nv) break;
  }
  return pos;
}

}  // namespace idx2g
