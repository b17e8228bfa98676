// Coding stages for 2D bricks, several bricks per pass (EncodeSubbandBlocks
// and the per-brick stream layout, Source/Core/idx2Encode.cpp:141-256 and
// 84-137).  A 2D subband has at most 25 4x4 blocks, too few to fill a
// 1024-thread workgroup, so the same subband of up to kGroup bricks is staged
// as one volume (brick = z slice) and coded together: the block stage runs
// once for all their blocks, while the key scans, stream layouts, arena
// slices, entries and headers stay per brick (segment).
#pragma once

#include "brick2d.hpp"
#include "code32.hpp"

namespace idx2g {
namespace b2d {

constexpr int kGroup = 5;          // bricks coded per pass (5 x 25 blocks <= kMaxBlocks)
constexpr int kParkOff = 8192;     // doubles: f64 plane words come first in the scratch slot
constexpr int kParkLen = kSlots * kN * kN;
constexpr int kSlotDoubles = kParkOff + kParkLen;  // scratch slot of one workgroup

struct CodeLds2 {
  union {
    double V[17 * 17 * kGroup];
    uint32_t stream[b32::kStreamWords];
  } v;
  i64 T[kTh / 64][4 * TEnc::q];  // TEnc layout
  uint32_t pool[16384];           // as CodeLds: (len | N << 8 | offset << 15) per (block, plane), f32 planes after
  uint32_t gblk[128];             // the group's block table (z = segment)
  short E[128];
  unsigned char np[128];
  uint32_t kbits[kGroup][kMaxKeys];
  uint32_t kbyte[kGroup][kMaxKeys];
  short kfirst[kGroup][kMaxKeys];
  unsigned char p3[64], p2[16];
  uint16_t e8[256];
  int klo_s[kGroup], khi_s[kGroup];
  unsigned long long base_s[kGroup], ebase_s[kGroup];
  uint32_t total_s[kGroup], woff_s[kGroup];
  int over;
};

struct Lds2 {
  double ex[kSlots * kN];  // x = 32 column of every slot
  union {
    double sub[kSlots * kSub];
    CodeLds2 code;
  } u;
};

// Code subband sb of the group's G bricks.  src: their parked subband data,
// [segment][j][i] (dx * dy per brick); sb_blocks: the subband's 2D block
// table (nb1 blocks); seg_dense: dense ids of the G bricks.
template <int DT>
__device__ __forceinline__ void code_group(const EncArgs& a, int sb, int G, const int* seg_dense, const double* src,
                                           CodeLds2& C, const b32::Pos& p, u64* XG, const uint32_t* sb_blocks, int nb1,
                                           b32::ArenaCursor& cur) {
  const BrickConst& k = a.k;
  constexpr int npl = DT ? 64 : 32;
  const int lane = p.lane;
  const int hx = k.sb[sb].f[0], hy = k.sb[sb].f[1];
  const int dx = 17 - hx, dy = 17 - hy;
  const int nblk = G * nb1;
  {  // two loads in flight per thread (the loop would wait for each in turn)
    const int n = G * dx * dy;
    for (int t0 = p.tid; t0 < n; t0 += 2 * kTh) {
      const int t1 = t0 + kTh;
      const double v0 = src[t0];
      const double v1 = t1 < n ? src[t1] : 0.0;
      C.v.V[t0] = v0;
      if (t1 < n) C.v.V[t1] = v1;
    }
  }
  for (int b = p.tid; b < nblk; b += kTh) {
    const int g = b / nb1;
    C.gblk[b] = sb_blocks[b - g * nb1] | ((uint32_t)g << 10);
  }
  if (p.tid < kGroup) {
    C.klo_s[p.tid] = 0x7fffffff;
    C.khi_s[p.tid] = -0x7fffffff;
  }
  if (p.tid == 0) C.over = 0;
  __syncthreads();
  b32::block_stage<DT, true, CodeLds2>(a, sb, dx, dy, C.gblk, nblk, C, p, 0, XG, seg_dense);
  __syncthreads();
  // ---- per-brick key offsets over its blocks (Morton order) ----
  for (int g = 0; g < G; ++g) {
    const int klo = C.klo_s[g], khi = C.khi_s[g];
    const int nk = khi >= klo ? khi - klo + 1 : 0;
    if (nk > kMaxKeys) continue;  // flagged below
    for (int kk = p.w; kk < nk; kk += kTh / 64) {
      const int key = klo + kk;
      const int b = g * nb1 + lane;
      const bool in = lane < nb1;
      const int e = in ? C.E[b] : kSkip;
      const int pp = 63 - (key - kBpBiasDev - e);
      const bool coded = in && e != kSkip && pp >= 0 && pp < C.np[b];
      const uint32_t len = coded ? (b32::lno_at(C, npl, b, pp) & 0xff) : 0u;
      const uint32_t incl = b32::wave_scan_add(len, lane);
      if (coded) b32::lno_at(C, npl, b, pp) |= (incl - len) << 15;
      const u64 m = __ballot(coded);
      if (lane == 0) {
        C.kbits[g][kk] = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        C.kfirst[g][kk] = (short)(m ? __ffsll((long long)m) - 1 : -1);
      }
    }
  }
  __syncthreads();
  // ---- per-brick byte layout + arena reservation (wave 0) ----
  if (p.w == 0) {
    uint32_t woff = 0;
    for (int g = 0; g < G; ++g) {
      const int klo = C.klo_s[g], khi = C.khi_s[g];
      const int nk = khi >= klo ? khi - klo + 1 : 0;
      const int dense = seg_dense[g];
      if (nk > kMaxKeys) {
        if (lane == 0) {
          atomicOr(&a.counters[2], 1ull);
          gm_store(a.hdrs + (size_t)dense * k.nsb + sb, SbHdr{0, 0, 0});
          C.total_s[g] = 0;
          C.woff_s[g] = woff;
          C.base_s[g] = 0;
          C.ebase_s[g] = 0;
        }
        continue;
      }
      uint32_t carry = 0;
      for (int c0 = 0; c0 < nk; c0 += 64) {
        const int kk = c0 + lane;
        const uint32_t by = kk < nk ? (C.kbits[g][kk] + 7) >> 3 : 0u;
        const uint32_t incl = b32::wave_scan_add(by, lane);
        if (kk < nk) C.kbyte[g][kk] = incl - by + carry;
        carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
      }
      if (lane == 0) {
        const unsigned long long alloc = (carry + 3ull) & ~3ull;
        if (alloc > cur.aleft) {
          const unsigned long long take = alloc > b32::kArenaChunk ? alloc : b32::kArenaChunk;
          cur.abase = atomicAdd(&a.counters[0], take);
          cur.aleft = take;
        }
        if ((unsigned long long)nk > cur.eleft) {
          const unsigned long long take =
              (unsigned long long)nk > b32::kEntryChunk ? (unsigned long long)nk : b32::kEntryChunk;
          cur.ebase = atomicAdd(&a.counters[1], take);
          cur.eleft = take;
        }
        const unsigned long long bse = cur.abase, ebs = cur.ebase;
        cur.abase += alloc;
        cur.aleft -= alloc;
        cur.ebase += nk;
        cur.eleft -= nk;
        cur.bytes += carry;
        const bool over = bse + alloc > a.arena_cap || ebs + nk > a.entry_cap;
        if (over) {
          atomicOr(&a.counters[2], 2ull);
          C.over = 1;
        }
        C.base_s[g] = bse;
        C.ebase_s[g] = ebs;
        C.total_s[g] = carry;
        C.woff_s[g] = woff;
        SbHdr h;
        h.klo = klo;
        h.nkeys = over ? 0 : nk;
        h.ebase = ebs;
        gm_store(a.hdrs + (size_t)dense * k.nsb + sb, h);
      }
      woff += (carry + 3) >> 2;
    }
  }
  __syncthreads();
  if (C.over) return;
  uint32_t words_all = 0;
  for (int g = 0; g < G; ++g) words_all += (C.total_s[g] + 3) >> 2;
  const bool staged = words_all <= (uint32_t)b32::kStreamWords;
  for (int g = 0; g < G; ++g) {
    const int klo = C.klo_s[g], khi = C.khi_s[g];
    const int nk = khi >= klo ? khi - klo + 1 : 0;
    if (nk > kMaxKeys) continue;
    for (int kk = p.tid; kk < nk; kk += kTh) {
      SbEntry e;
      e.off = C.base_s[g] + C.kbyte[g][kk];
      e.size = (C.kbits[g][kk] + 7) >> 3;
      e.first_block = (uint16_t)(C.kfirst[g][kk] < 0 ? 0xffff : C.kfirst[g][kk]);
      e.key = (uint16_t)(klo + kk);
      gm_store(a.entries + C.ebase_s[g] + kk, e);
    }
    if (!staged) {
      uint32_t* gw = (uint32_t*)(a.arena + C.base_s[g]);
      for (uint32_t i = p.tid; i < ((C.total_s[g] + 3) >> 2); i += kTh) gw[i] = 0;
    }
  }
  // ---- emission ----
  if (staged) {
    for (uint32_t i = p.tid; i < words_all; i += kTh) C.v.stream[i] = 0;
  } else {
    __threadfence();
  }
  __syncthreads();
  const int per_w = (nblk + kTh / 64 - 1) / (kTh / 64);  // blocks per wave, coded planes packed over lanes
  const int wb0 = min(nblk, p.w * per_w);
  b32::for_coded_planes(C, wb0, min(per_w, nblk - wb0), lane, [&](int b, int pp) {
    const int e = C.E[b];
    const int g = b / nb1;
    const uint32_t bw = C.gblk[b];
    const int nd = (((bw >> 15) & 7) > 1) + (((bw >> 18) & 7) > 1) + (((bw >> 21) & 7) > 1);
    const uint32_t lno = b32::lno_at(C, npl, b, pp);
    u64 lo, hi;
    b32::plane_bits(b32::plane_load(C, XG, npl, b, pp), int((lno >> 8) & 0x7f), 1 << (2 * nd), C.e8, lo, hi);
    const int kk = (63 - pp) + e + kBpBiasDev - C.klo_s[g];
    const u64 pos = (u64)C.kbyte[g][kk] * 8 + (lno >> 15);
    if (staged) b32::or_bits128<true>(C.v.stream + C.woff_s[g], pos, lo, hi);
    else b32::or_bits128<false>((uint32_t*)(a.arena + C.base_s[g]), pos, lo, hi);
  });
  __syncthreads();
  if (staged)
    for (int g = 0; g < G; ++g) {
      uint32_t* gw = (uint32_t*)(a.arena + C.base_s[g]);
      const uint32_t nw = (C.total_s[g] + 3) >> 2;
      for (uint32_t i = p.tid; i < nw; i += kTh) gw[i] = C.v.stream[C.woff_s[g] + i];
    }
  __syncthreads();
}

}  // namespace b2d
}  // namespace idx2g
