// On-chip transform of one full 32^3 brick (extended to 33^3 f64) by a
// workgroup of 1024 threads (16 wave64s), for gfx950.
//
// The extended brick (287 KB) does not fit in LDS, but it fits in the
// workgroup's registers: thread (x, y), x, y < 32, keeps the z-column
// c[0..32] in 66 VGPRs.  The three faces that do not belong to a thread live
// in LDS: EX = (x = 32, y < 32), EY = (y = 32, x < 32) and the corner column
// CC = (32, 32).  Per lifting pass:
//   * Z: in registers (columns are thread-private);
//   * X: across the lanes of a wave (lanes 0-31 hold x of row y = 2w, lanes
//     32-63 row 2w + 1) with DPP wave shifts; lane x = 31 also handles x = 32;
//   * Y: across waves, through two LDS row buffers (even rows, odd rows) in
//     z-chunks;
//   * the face lines that are not covered above are lifted one line per wave
//     with one point per lane.
// The coarse levels 1-4 of the extrapolation (17^3 and smaller) run on an
// LDS copy of the even subgrid: level 1 one line per thread, levels 2-4 by
// one wave without workgroup barriers.
//
// Arithmetic is the reference's (FLiftCdf53/ILiftCdf53, Source/Core/
// Wavelet.h:177-310; ExtrapolateCdf53 and the subband transform,
// Wavelet.cpp:67-203) element by element: predicts read only even samples,
// updates read only (predicted) odd samples, so every element sees exactly
// the operands and operation order of the sequential loops.  Division by 2
// and 4 are the exact products by 0.5 and 0.25.  Missing neighbour terms are
// skipped (never added as zeros) so that even the sign of zero matches.
#pragma once

#include "device.hpp"

namespace idx2g {
namespace b32 {

constexpr int kN = 33;
constexpr int kTh = 1024;
constexpr int kZC = 8;  // z planes per Y-pass chunk
enum { kFwd = 0, kExt = 1, kInv = 2 };

struct Lds {
  double ex[kN * 32];  // [z][y], x = 32
  double ey[kN * 32];  // [z][x], y = 32
  double cc[kN + 7];   // [z], x = y = 32 (padded)
  union {
    struct {
      double h0[kZC][16][32];  // even rows y = 2w, per wave w
      double h1[kZC][16][32];  // odd rows y = 2w + 1
    } halo;
    double sub[17 * 17 * 17];  // even subgrid [k][j][i] = (2i, 2j, 2k)
  } u;
};

// ---- wave shifts of a double (DPP wave_shr:1 / wave_shl:1) ----
__device__ __forceinline__ double lane_prev(double v) {  // value of lane - 1
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_next(double v) {  // value of lane + 1
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x130, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Keep a value materialised here: stops the compiler from sinking a whole
// unrolled pass into its consumers (which multiplies live registers).
__device__ __forceinline__ void pin(double& v) { asm volatile("" : "+v"(v)); }

// Orders one wave's LDS accesses (stores before, loads after); the wave's
// LDS instructions execute in issue order.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- one line held one point per lane: lane index i in [0, n) ----
// kFwd/kInv: D = n odd points.  kExt: D = 32 data points plus the
// extrapolated point 32 (n = 33).  Every lane of the wave must call this.
template <int MODE>
__device__ __forceinline__ double lift_lanes(double v, int i, int n) {
  const bool odd = i & 1;
  if (MODE == kFwd) {
    double l = lane_prev(v), r = lane_next(v);
    if (odd) v = v - (l + r) * 0.5;
    l = lane_prev(v);
    r = lane_next(v);
    if (!odd) {
      if (i > 0) v = v + l * 0.25;
      if (i < n - 1) v = v + r * 0.25;
    }
  } else if (MODE == kExt) {
    const double l = lane_prev(v), r = lane_next(v), ll = lane_prev(l);
    if (i == 32) v = 2.0 * l - ll;  // x0 = 2 L[31] - L[30]
    if (odd && i <= 29) v = v - (l + r) * 0.5;
    if (i == 31) v = 0.0;
    const double l2 = lane_prev(v), r2 = lane_next(v);
    if (!odd && i <= 30) {
      if (i > 0) v = v + l2 * 0.25;
      if (i < 30) v = v + r2 * 0.25;
    }
  } else {
    double l = lane_prev(v), r = lane_next(v);
    if (!odd) {
      if (i > 0) v = v - l * 0.25;
      if (i < n - 1) v = v - r * 0.25;
    }
    l = lane_prev(v);
    r = lane_next(v);
    if (odd) v = v + (l + r) * 0.5;
  }
  return v;
}

// ---- Z lifting of a register column (positions multiple of S) ----
template <int S>
__device__ __forceinline__ void zcol_fwd(double (&c)[kN]) {
#pragma unroll
  for (int o = S; o < kN; o += 2 * S) { c[o] = c[o] - (c[o - S] + c[o + S]) * 0.5; pin(c[o]); }
#pragma unroll
  for (int e = 0; e < kN; e += 2 * S) {
    double v = c[e];
    if (e > 0) v = v + c[e - S] * 0.25;
    if (e < 32) v = v + c[e + S] * 0.25;
    c[e] = v;
    pin(c[e]);
  }
}
template <int S>
__device__ __forceinline__ void zcol_inv(double (&c)[kN]) {
#pragma unroll
  for (int e = 0; e < kN; e += 2 * S) {
    double v = c[e];
    if (e > 0) v = v - c[e - S] * 0.25;
    if (e < 32) v = v - c[e + S] * 0.25;
    c[e] = v;
    pin(c[e]);
  }
#pragma unroll
  for (int o = S; o < kN; o += 2 * S) { c[o] = c[o] + (c[o - S] + c[o + S]) * 0.5; pin(c[o]); }
}
__device__ __forceinline__ void zcol_ext(double (&c)[kN]) {
  c[32] = 2.0 * c[31] - c[30];
#pragma unroll
  for (int o = 1; o <= 29; o += 2) { c[o] = c[o] - (c[o - 1] + c[o + 1]) * 0.5; pin(c[o]); }
  c[31] = 0.0;
#pragma unroll
  for (int e = 0; e <= 30; e += 2) {
    double v = c[e];
    if (e > 0) v = v + c[e - 1] * 0.25;
    if (e < 30) v = v + c[e + 1] * 0.25;
    c[e] = v;
    pin(c[e]);
  }
}

// ---- X lifting of one z plane of the register columns ----
// lane x of row y; ex = EX(z, y) (the x = 32 point, used by lane x = 31).
template <int MODE>
__device__ __forceinline__ double xlift(double v, int x, double* ex) {
  // branch-free except for the lane-31 store: every lane of a row reads the
  // same EX element (an LDS broadcast)
  const bool odd = x & 1, last = x == 31;
  if (MODE == kFwd) {
    const double e = *ex;
    double l = lane_prev(v), r = lane_next(v);
    r = last ? e : r;
    const double pv = v - (l + r) * 0.5;
    v = odd ? pv : v;
    l = lane_prev(v);
    r = lane_next(v);
    double t = v + l * 0.25;
    t = x > 0 ? t : v;
    t = t + r * 0.25;
    v = odd ? v : t;
    const double ne = e + v * 0.25;
    if (last) *ex = ne;
  } else if (MODE == kExt) {
    double l = lane_prev(v), r = lane_next(v);
    const double ne = 2.0 * v - l;
    if (last) *ex = ne;
    const double pv = v - (l + r) * 0.5;
    v = (odd && x <= 29) ? pv : v;
    v = last ? 0.0 : v;
    l = lane_prev(v);
    r = lane_next(v);
    double t = v + l * 0.25;
    t = x > 0 ? t : v;
    const double t2 = t + r * 0.25;
    t = x < 30 ? t2 : t;
    v = odd ? v : t;
  } else {
    double e = *ex;
    double l = lane_prev(v), r = lane_next(v);
    double t = v - l * 0.25;
    t = x > 0 ? t : v;
    t = t - r * 0.25;
    const double ne = e - v * 0.25;  // lane 31: update-undo of x = 32 with the odd L[31]
    v = odd ? v : t;
    l = lane_prev(v);
    r = lane_next(v);
    r = last ? ne : r;
    const double pv = v + (l + r) * 0.5;
    v = odd ? pv : v;
    if (last) *ex = ne;
  }
  return v;
}

struct Pos {
  int tid, x, y, w, lane;
  bool orow;  // odd row (lanes 32-63)
};

// Thread coordinates.  Called once per brick: the asm makes them opaque so
// that the compiler does not hoist the hundreds of addresses derived from
// them out of a brick loop (they would all stay live in registers).
__device__ __forceinline__ Pos pos() {
  Pos p;
  p.tid = threadIdx.x;
  asm volatile("" : "+v"(p.tid));
  p.lane = p.tid & 63;
  p.w = __builtin_amdgcn_readfirstlane(p.tid >> 6);  // uniform: the wave
  p.x = p.tid & 31;
  p.y = p.tid >> 5;
  p.orow = p.lane >= 32;
  return p;
}

// X pass at level 0: interior rows in registers (z < nz), plus the y = 32
// rows (EY + corner) one line per wave (not part of the extrapolation's
// first pass, whose X lines cover y < 32 only).
template <int MODE>
__device__ __forceinline__ void xpass(double (&c)[kN], Lds& S, const Pos& p) {
  constexpr int nz = MODE == kExt ? 32 : 33;
#pragma unroll
  for (int z = 0; z < nz; ++z) {
    c[z] = xlift<MODE>(c[z], p.x, &S.ex[z * 32 + p.y]);
    pin(c[z]);
  }
  if (MODE != kExt) {
    for (int z = p.w; z < kN; z += kTh / 64) {
      double* q = p.lane < 32 ? &S.ey[z * 32 + p.lane] : &S.cc[z];
      double v = p.lane < 33 ? *q : 0.0;
      v = lift_lanes<MODE>(v, p.lane, 33);
      if (p.lane < 33) *q = v;
    }
  }
}

// Y pass at level 0: interior columns through the row buffers, the x = 32
// column (EX + corner) one line per wave.
template <int MODE>
__device__ __forceinline__ void ypass(double (&c)[kN], Lds& S, const Pos& p) {
  constexpr int nz = MODE == kExt ? 32 : 33;
  // w (the wave, i.e. the pair of rows) is uniform across the wave: in a
  // scalar register the row-15 face work is a branch taken by one wave only
  const int x = p.x, w = __builtin_amdgcn_readfirstlane(p.w);
  // x = 32 column lines
  for (int z = w; z < nz; z += kTh / 64) {
    double* q = p.lane < 32 ? &S.ex[z * 32 + p.lane] : &S.cc[z];
    double v = p.lane < 33 ? *q : 0.0;
    v = lift_lanes<MODE>(v, p.lane, 33);
    if (p.lane < 33) *q = v;
  }
#pragma unroll
  for (int z0 = 0; z0 < nz; z0 += kZC) {
    // phase 1
#pragma unroll
    for (int zc = 0; zc < kZC; ++zc) {
      const int z = z0 + zc;
      if (z < nz) {
        if (MODE == kInv) {
          if (p.orow) {
            S.u.halo.h1[zc][w][x] = c[z];
            if (w == 15) S.ey[z * 32 + x] = S.ey[z * 32 + x] - c[z] * 0.25;
          }
        } else {
          if (!p.orow) S.u.halo.h0[zc][w][x] = c[z];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int zc = 0; zc < kZC; ++zc) {
      const int z = z0 + zc;
      if (z < nz) {
        if (MODE == kInv) {
          if (!p.orow) {
            double v = c[z];
            if (w > 0) v = v - S.u.halo.h1[zc][w - 1][x] * 0.25;
            v = v - S.u.halo.h1[zc][w][x] * 0.25;
            c[z] = v;
            S.u.halo.h0[zc][w][x] = v;
          }
        } else if (MODE == kFwd) {
          if (p.orow) {
            const double below = S.u.halo.h0[zc][w][x];
            const double above = w < 15 ? S.u.halo.h0[zc][w + 1][x] : S.ey[z * 32 + x];
            const double v = c[z] - (below + above) * 0.5;
            c[z] = v;
            S.u.halo.h1[zc][w][x] = v;
            if (w == 15) S.ey[z * 32 + x] = above + v * 0.25;
          }
        } else {
          if (p.orow) {
            const double below = S.u.halo.h0[zc][w][x];
            double v;
            if (w == 15) {
              S.ey[z * 32 + x] = 2.0 * c[z] - below;
              v = 0.0;
            } else {
              v = c[z] - (below + S.u.halo.h0[zc][w + 1][x]) * 0.5;
            }
            c[z] = v;
            S.u.halo.h1[zc][w][x] = v;
          }
        }
      }
    }
    __syncthreads();
    // phase 2
#pragma unroll
    for (int zc = 0; zc < kZC; ++zc) {
      const int z = z0 + zc;
      if (z < nz) {
        if (MODE == kInv) {
          if (p.orow) {
            const double below = S.u.halo.h0[zc][w][x];
            const double above = w < 15 ? S.u.halo.h0[zc][w + 1][x] : S.ey[z * 32 + x];
            c[z] = c[z] + (below + above) * 0.5;
          }
        } else {
          if (!p.orow) {
            double v = c[z];
            if (w > 0) v = v + S.u.halo.h1[zc][w - 1][x] * 0.25;
            if (MODE == kFwd || w < 15) v = v + S.u.halo.h1[zc][w][x] * 0.25;
            c[z] = v;
          }
        }
      }
    }
#pragma unroll
    for (int zc = 0; zc < kZC; ++zc)
      if (z0 + zc < nz) pin(c[z0 + zc]);
  }
}

// Z pass at level 0: register columns, plus the 65 face columns one line per
// wave (EX: 32 lines, EY: 32 lines, corner: 1).
template <int MODE>
__device__ __forceinline__ void zpass(double (&c)[kN], Lds& S, const Pos& p) {
  if (MODE == kFwd) zcol_fwd<1>(c);
  else if (MODE == kExt) zcol_ext(c);
  else zcol_inv<1>(c);
  for (int j = p.w; j < 65; j += kTh / 64) {
    double* q = j < 32 ? &S.ex[j] : j < 64 ? &S.ey[j - 32] : &S.cc[0];
    const int st = j < 64 ? 32 : 1;
    double v = p.lane < 33 ? q[p.lane * st] : 0.0;
    v = lift_lanes<MODE>(v, p.lane, 33);
    if (p.lane < 33) q[p.lane * st] = v;
  }
}

// One line of D points held in registers: the same operations, in the same
// order, as lift_lanes / zcol_fwd / zcol_inv.
template <int MODE, int D>
__device__ __forceinline__ void line_lift(double (&v)[D]) {
  if (MODE == kFwd) {
#pragma unroll
    for (int o = 1; o < D; o += 2) v[o] = v[o] - (v[o - 1] + v[o + 1]) * 0.5;
#pragma unroll
    for (int e = 0; e < D; e += 2) {
      double t = v[e];
      if (e > 0) t = t + v[e - 1] * 0.25;
      if (e < D - 1) t = t + v[e + 1] * 0.25;
      v[e] = t;
    }
  } else {
#pragma unroll
    for (int e = 0; e < D; e += 2) {
      double t = v[e];
      if (e > 0) t = t - v[e - 1] * 0.25;
      if (e < D - 1) t = t - v[e + 1] * 0.25;
      v[e] = t;
    }
#pragma unroll
    for (int o = 1; o < D; o += 2) v[o] = v[o] + (v[o - 1] + v[o + 1]) * 0.5;
  }
}

// Line t (< D * D) of the subgrid along axis A, lifted in registers.  The
// forward lift streams along the line (odd point o, then the even point o - 1
// it completes), so only a few points are live at a time.
template <int MODE, int A, int SS, int D>
__device__ __forceinline__ void sub_line(double* sub, int t) {
  constexpr int sa = (A == 0 ? 1 : A == 1 ? 17 : 289) * SS;
  constexpr int s1 = (A == 0 ? 17 : 1) * SS;
  constexpr int s2 = (A == 2 ? 17 : 289) * SS;
  double* q = sub + (t % D) * s1 + (t / D) * s2;
  double v[D];
#pragma unroll
  for (int i = 0; i < D; ++i) v[i] = q[i * sa];
  line_lift<MODE, D>(v);
#pragma unroll
  for (int i = 0; i < D; ++i) q[i * sa] = v[i];
}

// All D * D lines of one axis by the lanes of one wave.
template <int MODE, int A, int SS, int D>
__device__ __forceinline__ void sub_lines_wave(double* sub, int lane) {
#pragma unroll
  for (int it = 0; it < (D * D + 63) / 64; ++it)
    if (lane + 64 * it < D * D) sub_line<MODE, A, SS, D>(sub, lane + 64 * it);
}

// Levels 2-4 of the subgrid (9^3 points and fewer) forward then inverse, by
// one wave: 18 passes separated by wave-level ordering only.
__device__ __forceinline__ void sub_coarse_wave(double* sub, int lane) {
  sub_lines_wave<kFwd, 0, 2, 9>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 1, 2, 9>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 2, 2, 9>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 0, 4, 5>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 1, 4, 5>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 2, 4, 5>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 0, 8, 3>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 1, 8, 3>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kFwd, 2, 8, 3>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 2, 8, 3>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 1, 8, 3>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 0, 8, 3>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 2, 4, 5>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 1, 4, 5>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 0, 4, 5>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 2, 2, 9>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 1, 2, 9>(sub, lane);
  wave_lds_order();
  sub_lines_wave<kInv, 0, 2, 9>(sub, lane);
}

// Level 1 (17^3 points, 289 lines per axis), one line per thread.
template <int MODE>
__device__ __forceinline__ void sub_level1_lines(double* sub, const Pos& p) {
  if (MODE == kFwd) {
    if (p.tid < 289) sub_line<kFwd, 0, 1, 17>(sub, p.tid);
    __syncthreads();
    if (p.tid < 289) sub_line<kFwd, 1, 1, 17>(sub, p.tid);
    __syncthreads();
    if (p.tid < 289) sub_line<kFwd, 2, 1, 17>(sub, p.tid);
    __syncthreads();
  } else {
    if (p.tid < 289) sub_line<kInv, 2, 1, 17>(sub, p.tid);
    __syncthreads();
    if (p.tid < 289) sub_line<kInv, 1, 1, 17>(sub, p.tid);
    __syncthreads();
    if (p.tid < 289) sub_line<kInv, 0, 1, 17>(sub, p.tid);
    __syncthreads();
  }
}

// Copy the even subgrid between the registers/faces and LDS.
template <bool kIn>
__device__ __forceinline__ void sub_copy(double (&c)[kN], Lds& S, const Pos& p) {
  if (!(p.x & 1) && !(p.y & 1)) {
    double* q = S.u.sub + (p.y >> 1) * 17 + (p.x >> 1);
#pragma unroll
    for (int k = 0; k < 17; ++k) {
      if (kIn) q[k * 289] = c[2 * k];
      else c[2 * k] = q[k * 289];
    }
  }
  // faces: EX (16 even y x 17 even z), EY (16 x 17), corner (17)
  for (int t = p.tid; t < 561; t += kTh) {
    int k, j, i;
    double* f;
    if (t < 272) { k = t >> 4; j = (t & 15); i = 16; f = &S.ex[2 * k * 32 + 2 * j]; }
    else if (t < 544) { k = (t - 272) >> 4; i = (t - 272) & 15; j = 16; f = &S.ey[2 * k * 32 + 2 * i]; }
    else { k = t - 544; i = j = 16; f = &S.cc[2 * k]; }
    double* q = S.u.sub + k * 289 + j * 17 + i;
    if (kIn) *q = *f;
    else *f = *q;
  }
}

// ExtrapolateCdf53 of a full brick (data extent 32 on every axis): forward
// levels 0-4 on the growing data grid, then inverse levels 4-0 on 33^3.
// park: kParkN * kTh doubles of LDS that are free during the subgrid levels;
// the top kParkN points of every column wait there, which leaves room in
// the registers for the lines of level 1.
constexpr int kParkN = 12;
__device__ __forceinline__ void extrapolate(double (&c)[kN], Lds& S, const Pos& p, double* park) {
  xpass<kExt>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(1);
  ypass<kExt>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(2);
  zpass<kExt>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(3);
  sub_copy<true>(c, S, p);
#pragma unroll
  for (int k = 0; k < kParkN; ++k) park[k * kTh + p.tid] = c[kN - kParkN + k];
  __syncthreads();
  sub_level1_lines<kFwd>(S.u.sub, p);
  if (p.w == 0) sub_coarse_wave(S.u.sub, p.lane);
  __syncthreads();
  sub_level1_lines<kInv>(S.u.sub, p);
#pragma unroll
  for (int k = 0; k < kParkN; ++k) c[kN - kParkN + k] = park[k * kTh + p.tid];
  sub_copy<false>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(4);
  zpass<kInv>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(5);
  ypass<kInv>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(6);
  xpass<kInv>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(7);
}

// One-level forward transform of the extended brick (X, Y, Z), no weights.
__device__ __forceinline__ void forward(double (&c)[kN], Lds& S, const Pos& p) {
  xpass<kFwd>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(8);
  ypass<kFwd>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(9);
  zpass<kFwd>(c, S, p);
  __syncthreads();
  IDX2G_PHASE(10);
}

// Inverse one-level transform (Z, Y, X), weights already removed.
__device__ __forceinline__ void inverse(double (&c)[kN], Lds& S, const Pos& p) {
  zpass<kInv>(c, S, p);
  __syncthreads();
  ypass<kInv>(c, S, p);
  __syncthreads();
  xpass<kInv>(c, S, p);
  __syncthreads();
}

// Subband weights by parity: w8[px + 2 py + 4 pz].
__device__ __forceinline__ void weigh(double (&c)[kN], Lds& S, const Pos& p, const double* w8) {
  const double we = w8[(p.x & 1) + 2 * (p.y & 1)], wo = w8[(p.x & 1) + 2 * (p.y & 1) + 4];
#pragma unroll
  for (int z = 0; z < kN; ++z) c[z] = c[z] * ((z & 1) ? wo : we);
  for (int t = p.tid; t < kN * 32; t += kTh) {
    const int z = t >> 5, i = t & 31;
    S.ex[t] = S.ex[t] * w8[2 * (i & 1) + 4 * (z & 1)];
    S.ey[t] = S.ey[t] * w8[(i & 1) + 4 * (z & 1)];
  }
  if (p.tid < kN) S.cc[p.tid] = S.cc[p.tid] * w8[4 * (p.tid & 1)];
  __syncthreads();
}

}  // namespace b32
}  // namespace idx2g
