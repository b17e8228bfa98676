// On-chip transform of 32 full 2D bricks (32x32, extended to 33x33 f64) at a
// time by a workgroup of 1024 threads, for gfx950.
//
// Brick slot s (0..31) is the 32-lane half-wave s: lane x (< 32) keeps the
// y-column c[0..32] of its brick in registers; the x = 32 column EX[s][y]
// lives in LDS.  Per lifting pass:
//   * Y: in registers (columns are thread-private);
//   * X: across the 32 lanes of the slot with DPP wave shifts (lane 31 also
//     handles x = 32 through EX);
//   * the EX lines (33 points) one line per wave, one point per lane.
// The coarse levels 1-4 of the extrapolation run on LDS copies of the slots'
// 17x17 even subgrids.  Operation order per element is the reference's
// (FLiftCdf53/ILiftCdf53, Source/Core/Wavelet.h:177-310; ExtrapolateCdf53
// and the subband transform, Wavelet.cpp:67-203, order "XY++"), as for the
// 3D bricks (brick32.hpp, whose lifting helpers this file reuses).
#pragma once

#include "brick32.hpp"

namespace idx2g {
namespace b2d {

using b32::kN;
using b32::kFwd;
using b32::kExt;
using b32::kInv;
constexpr int kTh = 1024;
constexpr int kSlots = 32;
constexpr int kSub = 17 * 17;

struct Pos2 {
  int tid, x, slot, w, lane;
};

__device__ __forceinline__ Pos2 pos2() {
  Pos2 p;
  p.tid = threadIdx.x;
  asm volatile("" : "+v"(p.tid));
  p.lane = p.tid & 63;
  p.w = __builtin_amdgcn_readfirstlane(p.tid >> 6);  // uniform: the wave
  p.x = p.tid & 31;
  p.slot = p.tid >> 5;
  return p;
}

// X pass: rows y < ny of every slot (kExt: the data rows y < 32; else all 33).
template <int MODE>
__device__ __forceinline__ void xpass(double (&c)[kN], double* ex, const Pos2& p) {
  constexpr int ny = MODE == kExt ? 32 : 33;
  double* e = ex + p.slot * kN;
#pragma unroll
  for (int y = 0; y < ny; ++y) {
    c[y] = b32::xlift<MODE>(c[y], p.x, &e[y]);
    b32::pin(c[y]);
  }
}

// Y pass: register columns, then the EX lines (one per wave, 33 lanes).
template <int MODE>
__device__ __forceinline__ void ypass(double (&c)[kN], double* ex, const Pos2& p) {
  if (MODE == kFwd) b32::zcol_fwd<1>(c);
  else if (MODE == kExt) b32::zcol_ext(c);
  else b32::zcol_inv<1>(c);
#pragma unroll
  for (int y = 0; y < kN; ++y) b32::pin(c[y]);
  __syncthreads();  // EX written by the X pass
  for (int s = p.w; s < kSlots; s += kTh / 64) {
    double* q = ex + s * kN;
    double v = p.lane < kN ? q[p.lane] : 0.0;
    v = b32::lift_lanes<MODE>(v, p.lane, kN);
    if (p.lane < kN) q[p.lane] = v;
  }
}

// Line t of the first ns slots' LDS subgrids (sub[slot][j * 17 + i] =
// (2i, 2j)) along axis A with index stride SS and D points, lifted in
// registers by one thread (t < ns * D).
template <int MODE, int A, int SS, int D>
__device__ __forceinline__ void sub_line(double* sub, int t) {
  constexpr int sa = A == 0 ? SS : SS * 17, so = A == 0 ? SS * 17 : SS;
  const int s = t / D, o = t - s * D;
  double* q = sub + s * kSub + o * so;
  double v[D];
#pragma unroll
  for (int i = 0; i < D; ++i) v[i] = q[i * sa];
  b32::line_lift<MODE, D>(v);
#pragma unroll
  for (int i = 0; i < D; ++i) q[i * sa] = v[i];
}

template <int MODE, int L>
__device__ __forceinline__ void sub_level(double* sub, const Pos2& p, int ns) {
  constexpr int SS = 1 << (L - 1);
  constexpr int D = L == 1 ? 17 : L == 2 ? 9 : L == 3 ? 5 : 3;
  const int nl = ns * D;  // <= 32 * 17 < kTh
  if (MODE == kFwd) {
    if (p.tid < nl) sub_line<MODE, 0, SS, D>(sub, p.tid);
    __syncthreads();
    if (p.tid < nl) sub_line<MODE, 1, SS, D>(sub, p.tid);
    __syncthreads();
  } else {
    if (p.tid < nl) sub_line<MODE, 1, SS, D>(sub, p.tid);
    __syncthreads();
    if (p.tid < nl) sub_line<MODE, 0, SS, D>(sub, p.tid);
    __syncthreads();
  }
}

// Even subgrid between the registers / EX and LDS.
template <bool kIn>
__device__ __forceinline__ void sub_copy(double (&c)[kN], double* ex, double* sub, const Pos2& p) {
  if (!(p.x & 1)) {
    double* q = sub + p.slot * kSub + (p.x >> 1);
#pragma unroll
    for (int j = 0; j < 17; ++j) {
      if (kIn) q[j * 17] = c[2 * j];
      else c[2 * j] = q[j * 17];
    }
  }
  if (!kIn) {
#pragma unroll
    for (int y = 0; y < kN; ++y) b32::pin(c[y]);
  }
  // x = 32 column: EX[s][2j] <-> sub[s][j][16]
  for (int t = p.tid; t < kSlots * 17; t += kTh) {
    const int s = t / 17, j = t - s * 17;
    if (kIn) sub[s * kSub + j * 17 + 16] = ex[s * kN + 2 * j];
    else ex[s * kN + 2 * j] = sub[s * kSub + j * 17 + 16];
  }
}

// ExtrapolateCdf53 of full bricks (data extent 32 on both axes): forward
// levels 0-4 on the growing data grid, then inverse levels 4-0 on 33x33.
// The coarse levels run on the first ns slots only (the others hold no
// brick this iteration).
__device__ __forceinline__ void extrapolate(double (&c)[kN], double* ex, double* sub, const Pos2& p, int ns) {
  xpass<kExt>(c, ex, p);
  ypass<kExt>(c, ex, p);
  __syncthreads();
  sub_copy<true>(c, ex, sub, p);
  __syncthreads();
  sub_level<kFwd, 1>(sub, p, ns);
  sub_level<kFwd, 2>(sub, p, ns);
  sub_level<kFwd, 3>(sub, p, ns);
  sub_level<kFwd, 4>(sub, p, ns);
  sub_level<kInv, 4>(sub, p, ns);
  sub_level<kInv, 3>(sub, p, ns);
  sub_level<kInv, 2>(sub, p, ns);
  sub_level<kInv, 1>(sub, p, ns);
  sub_copy<false>(c, ex, sub, p);
  __syncthreads();
  ypass<kInv>(c, ex, p);
  __syncthreads();
  xpass<kInv>(c, ex, p);
  __syncthreads();
}

// One-level forward transform of the extended bricks (X, Y), no weights.
__device__ __forceinline__ void forward(double (&c)[kN], double* ex, const Pos2& p) {
  xpass<kFwd>(c, ex, p);
  ypass<kFwd>(c, ex, p);
  __syncthreads();
}

// Inverse one-level transform (Y, X), weights already removed.
__device__ __forceinline__ void inverse(double (&c)[kN], double* ex, const Pos2& p) {
  ypass<kInv>(c, ex, p);
  __syncthreads();
  xpass<kInv>(c, ex, p);
  __syncthreads();
}

}  // namespace b2d
}  // namespace idx2g
