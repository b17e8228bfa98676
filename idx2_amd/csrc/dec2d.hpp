// Decode of full 2D 32x32 bricks, 32 per workgroup iteration (DecodeBrick /
// DecodeSubband, Source/Core/idx2Decode.cpp:85-321, for 2D bricks): the
// stream walks run in k_walk (2D shape); here the plane words of
// up to kGroup bricks' same subband are turned back into values in one pass
// (a 2D subband has at most 25 blocks) into an LDS array [brick][j][i], from
// which every brick slot takes its samples straight into its register
// columns; then the inverse weights and the inverse transform run on chip
// (brick2d.hpp).
#pragma once

#include "code2d.hpp"
#include "dec32.hpp"

namespace idx2g {
namespace b2d {

// Parking area of the odd rows (the hy = 1 subbands) of a workgroup's 32
// slots, in doubles.
constexpr int kDecPark = kSlots * (17 * 16 + 16 * 16);

struct DecLds2 {
  double ex[kSlots * kN];      // x = 32 column of every slot
  double V[17 * 17 * kGroup];  // one subband of the group's bricks, [segment][j][i]
  i64 T[kTh / 64][4 * TDec::q];  // TDec layout
  short E[kSlots][4][32];
  unsigned char nbps[kSlots][4][32];
  unsigned char p3[64], p2[16];
};

// Subband sb of the group's G bricks (slots s0 ..): plane words of slot s in
// the region a.xs + ((bi0 + s) * 8 + sb) * xs_region(npl) (the walk's
// layout, b32::xs_at); values to
// P[(g * dy + j) * dx + i] for g = slot - s0.
template <int DT>
__device__ __forceinline__ void inverse_blocks_2d(const DecArgs& a, int sb, int s0, int G, int dx, int dy, DecLds2& D,
                                                  const b32::Pos& p, const uint32_t* sb_blocks, int nb1, int bi0,
                                                  double* P) {
  const BrickConst& k = a.k;
  constexpr int npl = DT ? 64 : 32;
  const int lane = p.lane, q = lane >> 4, r = lane & 15;
  const int nblk = G * nb1;
  i64* T = &D.T[p.w][0];
  auto xs_word = [&](int blk, int pp) {  // plane word pp of group block blk
    const int g = blk / nb1, lb = blk - g * nb1;
    return gm(a.xs)[((size_t)(bi0 + s0 + g) * 8 + sb) * b32::xs_region(npl) + b32::xs_at(lb, pp, npl)];
  };
  auto nbps_of = [&](int blk) -> int {
    const int g = blk / nb1, lb = blk - g * nb1;
    return D.nbps[s0 + g][sb][lb];
  };
  for (int g4 = p.w; g4 * 4 < nblk; g4 += kTh / 64) {
    const int b = g4 * 4 + q;
    const bool bin = b < nblk;
    const int sg = bin ? b / nb1 : 0, lb = bin ? b - sg * nb1 : 0;
    const uint32_t bw = bin ? sb_blocks[lb] : (1u << 24);
    const int o3[3] = {int(bw & 31), int((bw >> 5) & 31), int((bw >> 10) & 31)};
    const int bd[3] = {int((bw >> 15) & 7), int((bw >> 18) & 7), int((bw >> 21) & 7)};
    const bool ux = bd[0] > 1, uy = bd[1] > 1;
    const int nd = ux + uy;
    const int e = bin ? D.E[s0 + sg][sb][lb] : kSkip;
    bool live = bin && e != kSkip;
    if (live) {
      const int stop = max(64 - k.endbp, 57 + a.etol - e);
      const int np = stop > 63 ? 0 : 64 - stop;
      bool bypass = false;
      for (int d = 0; d < 3; ++d) bypass |= (o3[d] & a.spmask[sb][d]) != 0;
      live = np > 0 && !bypass;
    }
    const int nd_s[4] = {__builtin_amdgcn_readlane(nd, 0), __builtin_amdgcn_readlane(nd, 16),
                         __builtin_amdgcn_readlane(nd, 32), __builtin_amdgcn_readlane(nd, 48)};
    // f32, and f64 rounds whose blocks have at most 32 planes decoded: two
    // blocks share a transpose of their top 32 planes
    const int nbq = bin ? nbps_of(b) : 0;
    const bool narrow = DT == 0 || max(max(__builtin_amdgcn_readlane(nbq, 0), __builtin_amdgcn_readlane(nbq, 16)),
                                       max(__builtin_amdgcn_readlane(nbq, 32), __builtin_amdgcn_readlane(nbq, 48))) <= 32;
    if (narrow) {
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) {
        const int slot = 2 * pr + (lane >> 5);
        const int blk = g4 * 4 + slot;
        const int pp = 31 - (lane & 31);
        const u64 X = (blk < nblk && pp < nbps_of(blk)) ? xs_word(blk, pp) : 0ull;
        const u64 R = b32::bit_transpose(X, lane);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int sl = 2 * pr + h, ndq = nd_s[sl];
          const u64 u = h == 0 ? (R << 32) : (R & 0xFFFFFFFF00000000ull);
          const int c = lane;
          if (c < (1 << (2 * ndq))) {
            const int dst = ndq == 3 ? D.p3[c] : ndq == 2 ? D.p2[c] : c;
            T[sl * TDec::q + dst] = (i64)((u ^ IDX2G_NBMASK) - IDX2G_NBMASK);
          }
        }
      }
    } else {
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const int blk = g4 * 4 + qq;
        const int pp = 63 - lane;
        const u64 X = (blk < nblk && pp < nbps_of(blk)) ? xs_word(blk, pp) : 0ull;
        const u64 u = b32::bit_transpose(X, lane);
        const int ndq = nd_s[qq], c = lane;
        if (c < (1 << (2 * ndq))) {
          const int dst = ndq == 3 ? D.p3[c] : ndq == 2 ? D.p2[c] : c;
          T[qq * TDec::q + dst] = (i64)((u ^ IDX2G_NBMASK) - IDX2G_NBMASK);
        }
      }
    }
    b32::wave_sync_lds();
    i64* Tq = T + q * TDec::q;
    {  // second axis (2D blocks: inverse lifts run y, x)
      const int j0 = r & 3, j2 = r >> 2;
      i64 y4[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) y4[i] = Tq[j0 + 4 * i + TDec::c * j2];
      if (nd >= 2) {
        zfp_ilift4(y4[0], y4[1], y4[2], y4[3]);
#pragma unroll
        for (int i = 0; i < 4; ++i) Tq[j0 + 4 * i + TDec::c * j2] = y4[i];
      }
    }
    b32::wave_sync_lds();
    // row mode: first axis, dequantise (Dequantize, Function.h:139-149)
    {
      const int s0a = ux ? 1 : dx;  // first non-unit axis
      const int s1a = (ux && uy) ? dx : 0;
      const int i1 = r & 3, i2 = r >> 2;
      const bool rowok = bin && ((nd == 2 && i2 == 0) || (nd <= 1 && r == 0));  // zeros when not decoded
      const int n0 = nd >= 1 ? 4 : 1;
      i64 w4[4];
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0) w4[i0] = Tq[i0 + 4 * i1 + TDec::c * i2];
      if (nd >= 1) zfp_ilift4(w4[0], w4[1], w4[2], w4[3]);
      const double scale = inv_pow2d((63 - nd) - 1 - e);
      const int vbase = (sg * dy + o3[1]) * dx + o3[0] + (nd >= 2 ? i1 * s1a : 0);
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0)
        if (rowok && i0 < n0) P[vbase + i0 * s0a] = live ? scale * i64_to_f64(w4[i0]) : 0.0;
    }
    b32::wave_sync_lds();
  }
}

// The group's subband from V to the workgroup's parking area:
// P[(s * dy + j) * dx + i] for slot s = s0 + g (P: the subband's area).
__device__ __forceinline__ void park_group(const DecLds2& D, double* P, const Pos2& p, int dx, int dy, int s0, int G) {
  const auto q = gm(P) + (size_t)s0 * dx * dy;
  for (int t = p.tid; t < G * dx * dy; t += kTh) q[t] = D.V[t];
}

// Odd rows of every slot from the parking area (areas at off[hx] for hy = 1)
// into the register columns and the x = 32 column.
__device__ __forceinline__ void unpark_odd(double (&c)[kN], double* ex, const double* P, const int* off,
                                           const Pos2& p) {
  const int hx = p.x & 1;
  const int dx = 17 - hx;
  const auto q = gm(P) + off[hx] + (p.slot * 16) * dx + (p.x >> 1);
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) c[2 * jj + 1] = q[jj * dx];
  for (int t = p.tid; t < kSlots * 16; t += kTh) {
    const int s = t >> 4, jj = t & 15;
    ex[s * kN + 2 * jj + 1] = gm(P)[off[0] + (s * 16 + jj) * 17 + 16];
  }
}

// Samples (2i + hx, 2j + HY) of the group's subband from V into the register
// columns of slot s0 + g (lanes of matching x parity) and the x = 32 column.
template <int HY>
__device__ __forceinline__ void take_subband(double (&c)[kN], DecLds2& D, const Pos2& p, int hx, int s0, int G) {
  constexpr int dy = 17 - HY;
  const int dx = 17 - hx, g = p.slot - s0;
  if (g >= 0 && g < G && (p.x & 1) == hx) {
    const double* q = D.V + (g * dy) * dx + (p.x >> 1);
#pragma unroll
    for (int jj = 0; jj < dy; ++jj) c[2 * jj + HY] = q[jj * dx];
  }
  if (hx == 0)
    for (int t = p.tid; t < G * dy; t += kTh) {
      const int gg = t / dy, jj = t - gg * dy;
      D.ex[(s0 + gg) * kN + 2 * jj + HY] = D.V[(gg * dy + jj) * dx + 16];
    }
}

}  // namespace b2d
}  // namespace idx2g
