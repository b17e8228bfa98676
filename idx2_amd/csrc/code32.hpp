// zfp block stage and bit-plane coding of one transformed 32^3 brick
// (EncodeSubbandBlocks, Source/Core/idx2Encode.cpp:141-256; zfp Encode,
// Source/Core/Zfp.h:770-813), lane-parallel, for the 1024-thread brick
// workgroup of brick32.hpp.  Per subband:
//   1. stage the subband's f64 values from the register columns and the LDS
//      faces into a dense LDS array V;
//   2. block stage, 4 blocks per wave: "row mode" (one lane per 4-value row)
//      for the common exponent, quantisation and the zfp lifts (rows along
//      the first axis in registers, the other axes through a per-wave LDS
//      transpose buffer), then "coefficient mode" (one lane per coefficient)
//      for the coefficient order and negabinary, then a 64x64 bit-matrix
//      transpose across the wave that leaves one bit plane per lane;
//   3. per bit plane: significance count N by a suffix-max scan over the
//      planes above it, emitted length by closed form (no bit loop);
//   4. per key (bit-plane channel), one wave per key: an exclusive scan of
//      the plane lengths over the blocks in Morton order gives each plane's
//      bit offset in the stream, and the same lane emits the plane's bits
//      there, into an LDS stream buffer (the global arena when the streams
//      do not fit), which is then copied into the arena.
#pragma once

#include "brick32.hpp"
#include "kernels.hpp"

namespace idx2g {
namespace b32 {

constexpr int kStreamWords = 4096;  // LDS stream buffer (16 KB)
// per (block, plane) words of the pool: the length word of every plane plus,
// for float32, the plane word (two 32-bit words) after them
constexpr int kPoolWords = kMaxBlocks * 32 * 3;
constexpr unsigned long long kArenaChunk = 256ull << 10;  // arena bytes a workgroup takes at a time
constexpr unsigned long long kEntryChunk = 2048;          // stream entries a workgroup takes at a time

// A workgroup's private slices of the level arena and entry table, refilled
// with one global atomic when exhausted (no atomic round trip per subband).
struct ArenaCursor {
  unsigned long long abase, aleft, ebase, eleft;
  unsigned long long bytes;  // stream bytes coded, added to counters[3] once at the end
};

// The workgroup's stream byte count into counters[3] (all jobs of a launch
// share the counters), once per kernel instead of one atomic per subband.
__device__ __forceinline__ void flush_cursor(const ArenaCursor& cur, unsigned long long* counters) {
  __syncthreads();
  if (threadIdx.x == 0 && cur.bytes) atomicAdd(&counters[3], cur.bytes);
}

struct CodeLds {
  struct {
    double V[kVXY * 17];  // one subband, kVX / kVY layout
  } v;
  // the streams of one subband; all zero between subbands (the copy into
  // the arena clears what it copies), so the next subband's staging into V
  // does not wait for that copy
  uint32_t stream[kStreamWords];
  i64 T[kTh / 64][4 * TEnc::q];   // row-mode transpose buffers, per wave (TEnc layout)
  // per (block, plane): len | N << 8, stride npl; for f32 the plane words
  // follow (f64 plane words go to global scratch)
  uint32_t pool[kPoolWords];
  short E[128];                   // block exponent (kSkip: not coded)
  unsigned char np[128];          // coded planes of the block
  // per key stream, indexed by key % kMaxKeys (a subband's keys span fewer
  // than kMaxKeys): bits, first block with a plane in it
  uint32_t kbits[kMaxKeys];
  int kfirst[kMaxKeys];
  unsigned char p3[64], p2[16];
  uint16_t e8[256];               // byte -> bits with a 1 inserted after every 1
  int klo, khi;
  // arena slice and overflow flag of the subband with sequence number q in
  // slot q & 1: read after the emission's barrier, when the next subband may
  // already write its own slot
  unsigned long long base_s[2];
  int over_s[2];
};

// ---- wave helpers ----
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, true);
}

// Value of lane ^ s for a 32-bit quantity, s in {1, 2, 4, 8, 16, 32}.
// All DPP moves run with the full wave active (a DPP read of a lane that is
// switched off returns 0), hence the asm that keeps both row shifts
// outside of any branch.
template <int S>
__device__ __forceinline__ uint32_t xlane(uint32_t v, int lane) {
  if (S == 1) return dpp_u<0xB1>(v);                  // quad_perm [1,0,3,2]
  if (S == 2) return dpp_u<0x4E>(v);                  // quad_perm [2,3,0,1]
  if (S == 8) return dpp_u<0x128>(v);               // row_ror:8: lane ^ 8 inside a row
  if (S == 4) {
    uint32_t up = dpp_u<0x114>(v), dn = dpp_u<0x104>(v);  // row_shr:4, row_shl:4
    asm volatile("" : "+v"(up), "+v"(dn));
    return (lane & 4) ? up : dn;
  }
  if (S == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? (uint32_t)p[0] : (uint32_t)p[1];
  }
  const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return (lane & 32) ? (uint32_t)p[0] : (uint32_t)p[1];
}

template <int CTRL>
__device__ __forceinline__ int dppi(int v) { return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true); }

// Max over each 16-lane row, in every lane of the row.
__device__ __forceinline__ int row_max(int v) {
  v = max(v, dppi<0xB1>(v));   // quad_perm [1,0,3,2]
  v = max(v, dppi<0x4E>(v));   // quad_perm [2,3,0,1]
  v = max(v, dppi<0x141>(v));  // row_half_mirror
  v = max(v, dppi<0x140>(v));  // row_mirror
  return v;
}

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ uint32_t wave_scan_add(uint32_t v, int lane) {
  v += (uint32_t)dppi<0x111>((int)v);  // row_shr:1 (0 shifted in)
  v += (uint32_t)dppi<0x112>((int)v);
  v += (uint32_t)dppi<0x114>((int)v);
  v += (uint32_t)dppi<0x118>((int)v);
  const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
  const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
  const uint32_t t2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
  const int row = lane >> 4;
  v += row == 0 ? 0u : row == 1 ? t0 : row == 2 ? t0 + t1 : t0 + t1 + t2;
  return v;
}

// Exclusive suffix max (over the higher lanes) inside segments of seg lanes
// (32 or 64) of non-negative values.
__device__ __forceinline__ int seg_suffix_max_excl(int v, int lane, int seg) {
  v = max(v, dppi<0x101>(v));  // row_shl:1 (0 shifted in)
  v = max(v, dppi<0x102>(v));
  v = max(v, dppi<0x104>(v));
  v = max(v, dppi<0x108>(v));
  const int r1 = __builtin_amdgcn_readlane(v, 16), r2 = __builtin_amdgcn_readlane(v, 32),
            r3 = __builtin_amdgcn_readlane(v, 48);
  const int row = lane >> 4;
  int add;
  if (seg == 32) add = row == 0 ? r1 : row == 2 ? r3 : 0;
  else add = row == 0 ? max(r1, max(r2, r3)) : row == 1 ? max(r2, r3) : row == 2 ? r3 : 0;
  v = max(v, add);
  int nx = dppi<0x101>(v);
  const int s1 = __builtin_amdgcn_readlane(v, 16), s2 = __builtin_amdgcn_readlane(v, 32),
            s3 = __builtin_amdgcn_readlane(v, 48);
  if ((lane & 15) == 15) nx = row == 0 ? s1 : row == 1 ? s2 : row == 2 ? s3 : 0;
  return ((lane & (seg - 1)) + 1 < seg) ? nx : 0;
}

// One butterfly stage of the 64x64 bit transpose: lanes L (bit s of the
// lane clear) and U = L ^ s exchange the bit columns whose bit s differs:
//   L: (r & M) | (P << s & ~M),  U: (r & ~M) | (P >> s & M)   (P: partner's r)
// which is a bit select keep ? r : T with T = M ? P >> s : P << s (the
// partner's word with adjacent s-bit groups swapped) and keep = M (L) or ~M
// (U).  gfx950's three-input bit operation does each select in one
// instruction (0xE4: S2 ? S0 : S1); the s = 32 stage is a single
// permlane32_swap of the two halves, the s = 16 swap a rotate.
template <int S>
__device__ __forceinline__ u64 tstage(u64 r, int lane) {
  uint32_t lo = (uint32_t)r, hi = (uint32_t)(r >> 32);
  if constexpr (S == 32) {
    const auto p = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
    return ((u64)(uint32_t)p[1] << 32) | (uint32_t)p[0];
  } else {
    constexpr uint32_t M = S == 16 ? 0x0000FFFFu : S == 8 ? 0x00FF00FFu : S == 4 ? 0x0F0F0F0Fu
                         : S == 2 ? 0x33333333u : 0x55555555u;
    const uint32_t plo = xlane<S>(lo, lane), phi = xlane<S>(hi, lane);
    uint32_t tlo, thi;
    if constexpr (S == 16) {
      tlo = __builtin_amdgcn_alignbit(plo, plo, 16);
      thi = __builtin_amdgcn_alignbit(phi, phi, 16);
    } else {
      tlo = __builtin_amdgcn_bitop3_b32(plo >> S, plo << S, M, 0xE4);
      thi = __builtin_amdgcn_bitop3_b32(phi >> S, phi << S, M, 0xE4);
    }
    const uint32_t keep = (lane & S) ? ~M : M;
    lo = __builtin_amdgcn_bitop3_b32(lo, tlo, keep, 0xE4);
    hi = __builtin_amdgcn_bitop3_b32(hi, thi, keep, 0xE4);
    return ((u64)hi << 32) | lo;
  }
}

// Lane j returns sum_i bit_j(r_i) << i (r_i: the value of lane i).
__device__ __forceinline__ u64 bit_transpose(u64 r, int lane) {
  r = tstage<32>(r, lane);
  r = tstage<16>(r, lane);
  r = tstage<8>(r, lane);
  r = tstage<4>(r, lane);
  r = tstage<2>(r, lane);
  r = tstage<1>(r, lane);
  return r;
}

__device__ __forceinline__ int clz64(u64 x) { return x ? __clzll((long long)x) : 64; }

// Emitted length of one plane given N (closed form of the embedded coder,
// Zfp.h:770-813): verbatim N bits, then one group test per newly
// significant coefficient.  Returns the length, updates n.
__device__ __forceinline__ int plane_len_cf(u64 X, int& n, int nv) {
  if (n >= nv) return nv;
  const u64 Y = X >> n;
  if (Y == 0) return n + 1;
  const int qm = 63 - clz64(X);
  const int m = __popcll(Y);
  const int len = qm + 1 + m - (qm == nv - 1) + (qm + 1 < nv);
  n = qm + 1;
  return len;
}

// Emitted bits of one plane (LSB first) as a 128-bit value, without a bit
// loop.  After the N verbatim bits, the group tests turn the remaining bits
// Y = X >> N, read up to its last one t_m, into Y with a 1 inserted before
// every run "0..01"; with E() = "a 1 inserted after every 1":
//   tail = 1 | E(Y below t_m) << 1 | 1 << (t_m + m)   (m ones in Y)
// where the last 1 is dropped when t_m is the block's last coefficient
// (it is implied) and a 0 follows when coefficients remain.  E() is
// assembled from a 256-entry byte table in a three-level tree.
__device__ __forceinline__ void plane_bits(u64 X, int n, int nv, const uint16_t* e8, u64& lo, u64& hi) {
  lo = n >= 64 ? X : (X & ((1ull << n) - 1));
  hi = 0;
  if (n >= nv) return;
  const u64 Y = X >> n;
  if (Y == 0) return;
  const int tm = 63 - clz64(Y), m = __popcll(Y);
  const u64 Yl = tm ? (Y & ((1ull << tm) - 1)) : 0ull;
  uint32_t p8[8], l8[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t byte = uint32_t(Yl >> (8 * i)) & 0xffu;
    p8[i] = e8[byte];
    l8[i] = 8u + __popc(byte);
  }
  u64 q4[4];
  uint32_t L4[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    q4[j] = (u64)p8[2 * j] | ((u64)p8[2 * j + 1] << l8[2 * j]);
    L4[j] = l8[2 * j] + l8[2 * j + 1];
  }
  const u64 r0 = q4[0] | (q4[1] << L4[0]), r1 = q4[2] | (q4[3] << L4[2]);
  const uint32_t R0 = L4[0] + L4[1];  // 16..64
  // E = r0 | r1 << R0 as 128 bits
  u64 elo = r0 | (R0 < 64 ? (r1 << R0) : 0ull);
  u64 ehi = R0 < 64 ? (r1 >> (64 - R0)) : r1;
  // tail = 1 | E << 1 | value bit of the last one
  u64 tlo = 1ull | (elo << 1), thi = (ehi << 1) | (elo >> 63);
  if (n + tm != nv - 1) {
    const int vb = tm + m;  // <= 127
    if (vb < 64) tlo |= 1ull << vb; else thi |= 1ull << (vb - 64);
  }
  // out = verbatim | tail << n (n < 64)
  lo |= tlo << n;
  hi = n ? ((thi << n) | (tlo >> (64 - n))) : thi;
}

// OR a 128-bit little-endian bit string into 32-bit words at bit position pos.
template <bool kLds>
__device__ __forceinline__ void or_bits128(uint32_t* words, u64 pos, u64 lo, u64 hi) {
  const int sh = int(pos & 31);
  uint32_t* w = words + (pos >> 5);
  const u64 a0 = lo << sh;
  const u64 a1 = sh ? ((hi << sh) | (lo >> (64 - sh))) : hi;
  const uint32_t a2 = sh ? (uint32_t)(hi >> (64 - sh)) : 0u;
  const uint32_t d[5] = {(uint32_t)a0, (uint32_t)(a0 >> 32), (uint32_t)a1, (uint32_t)(a1 >> 32), a2};
#pragma unroll
  for (int i = 0; i < 5; ++i)
    if (d[i]) atomicOr(&w[i], d[i]);
}

// The coded (block, plane) pairs of blocks [b0, b0 + nb) (nb <= 16), packed
// over the wave's lanes: a block's coded planes are pp = 0 .. np - 1, so the
// pairs are numbered by an exclusive scan of np and item i belongs to the
// block whose range holds it.  f(b, pp) runs once per coded pair.
template <class CL, class F>
__device__ __forceinline__ void for_coded_planes(const CL& C, int b0, int nb, int lane, F&& f) {
  const int bl = b0 + lane;
  const int np = (lane < nb && C.E[bl] != kSkip) ? (int)C.np[bl] : 0;
  const uint32_t inc = wave_scan_add((uint32_t)np, lane);
  const int excl = (int)inc - np;
  const int tot = __builtin_amdgcn_readlane((int)inc, 63);
  for (int i0 = 0; i0 < tot; i0 += 64) {
    const int it = i0 + lane;
    int j = 0;  // blocks that end at or before item it
    for (int q = 0; q < nb; ++q) j += __builtin_amdgcn_readlane((int)inc, q) <= it;
    const int start = __shfl(excl, j < nb ? j : 0, 64);
    if (it < tot) f(b0 + j, it - start);
  }
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- 1. stage subband sb of the transformed brick into V ----
template <int HZ>
__device__ __forceinline__ void stage_cols(const double (&c)[kN], double* V, int i, int j, int dz) {
  double* q = V + j * kVX + i;
#pragma unroll
  for (int k = 0; k < 17; ++k)
    if (k < dz) q[k * kVXY] = c[HZ + 2 * k];
}

// Faces of subband (hx, hy, hz) from the x = 32 / y = 32 planes and the corner
// line: x = 32 (hx = 0), y = 32 (hy = 0), corner.
__device__ __forceinline__ void stage_faces(const Lds& S, double* V, const Pos& p, int hx, int hy, int hz) {
  const int dz = 17 - hz;
  for (int t = p.tid; t < 17 * 16 * 2 + 17; t += kTh) {
    if (t < 272) {
      if (hx) continue;
      const int k = t >> 4, j = t & 15;
      if (k < dz) V[(k * kVY + j) * kVX + 16] = S.ex[(hz + 2 * k) * 32 + hy + 2 * j];
    } else if (t < 544) {
      if (hy) continue;
      const int k = (t - 272) >> 4, i = (t - 272) & 15;
      if (k < dz) V[(k * kVY + 16) * kVX + i] = S.ey[(hz + 2 * k) * 32 + hx + 2 * i];
    } else {
      if (hx || hy) continue;
      const int k = t - 544;
      if (k < dz) V[(k * kVY + 16) * kVX + 16] = S.cc[hz + 2 * k];
    }
  }
}

// z-even subband (hx, hy, 0) into V: interior from the register columns.
__device__ __forceinline__ void stage_subband_even(const double (&c)[kN], const Lds& S, double* V, const Pos& p, int hx,
                                                   int hy) {
  if ((p.x & 1) == hx && (p.y & 1) == hy) stage_cols<0>(c, V, p.x >> 1, p.y >> 1, 17);
  stage_faces(S, V, p, hx, hy, 0);
}

// The z-odd half of the register columns, parked in the workgroup's scratch
// volume subband by subband: Z1[((hx + 2 hy) * 16 + k) * 256 + j * 16 + i]
// holds column (2i + hx, 2j + hy) at z = 2k + 1.
constexpr int kOddOff = 16384;  // doubles into the scratch volume
__device__ __forceinline__ void park_odd(const double (&c)[kN], double* Z1, const Pos& p) {
  auto q = gm(Z1) + ((p.x & 1) + 2 * (p.y & 1)) * 16 * 256 + (p.y >> 1) * 16 + (p.x >> 1);
#pragma unroll
  for (int k = 0; k < 16; ++k) q[k * 256] = c[2 * k + 1];
}

// z-odd subband (hx, hy, 1) into V from the parked columns.
__device__ __forceinline__ void stage_subband_odd(const double* Z1, const Lds& S, double* V, const Pos& p, int hx,
                                                  int hy) {
  const auto src = gm(Z1) + (hx + 2 * hy) * 16 * 256;
  static_assert(16 * 256 == 4 * kTh, "four samples per thread");
  // the four loads first, so that their latencies overlap
  double v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = src[p.tid + r * kTh];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int t = p.tid + r * kTh;
    const int k = t >> 8, j = (t >> 4) & 15, i = t & 15;
    V[(k * kVY + j) * kVX + i] = v[r];
  }
  stage_faces(S, V, p, hx, hy, 1);
}

template <class CL>
__device__ __forceinline__ uint32_t& lno_at(CL& C, int npl, int b, int pp) { return C.pool[b * npl + pp]; }
// Plane word of (block b, plane pp): in LDS after the pool entries for f32
// (npl 32), in the workgroup's global scratch XG for f64.
template <class CL>
__device__ __forceinline__ void plane_store(CL& C, u64* XG, int npl, int b, int pp, u64 X) {
  if (npl == 32) reinterpret_cast<u64*>(C.pool + kMaxBlocks * 32)[b * 32 + pp] = X;
  else gm(XG)[b * 64 + pp] = X;
}
template <class CL>
__device__ __forceinline__ u64 plane_load(const CL& C, const u64* XG, int npl, int b, int pp) {
  return npl == 32 ? reinterpret_cast<const u64*>(C.pool + kMaxBlocks * 32)[b * 32 + pp] : gm(XG)[b * 64 + pp];
}

// Coefficient c of block slot qq (coefficient order, then negabinary),
// ForwardShuffle (Zfp.h:534-559).
template <class CL>
__device__ __forceinline__ u64 coef_u(const i64* Tq, const CL& C, int nd, int c) {
  const int nv = 1 << (2 * nd);
  if (c >= nv) return 0;
  const int src = nd == 3 ? C.p3[c] : nd == 2 ? C.p2[c] : c;
  return ((u64)Tq[src] + IDX2G_NBMASK) ^ IDX2G_NBMASK;
}

// Plane mode after the transpose: lane j holds plane word X of block slot
// `slot` (of the wave's group) and plane pp.  seg: lanes per block (32 for
// f32 pairs, 64 for f64).  kAcc: also add the plane's length to its key
// stream's total and lower the stream's first block (e: block exponent).
template <bool kAcc = false, class CL>
__device__ __forceinline__ void plane_mode(u64 X, int lane, int seg, int blk, bool blk_ok, int np, int nd, int pp,
                                           CL& C, u64* XG, int npl, int e = 0) {
  const int k = lane & (seg - 1);
  const bool coded = blk_ok && pp < np;
  const int t = (coded && X) ? 64 - clz64(X) : 0;
  int n = seg_suffix_max_excl(t, lane, seg);
  (void)k;
  if (coded) {
    const int n_in = n;
    const int len = plane_len_cf(X, n, 1 << (2 * nd));
    lno_at(C, npl, blk, pp) = (uint32_t)len | ((uint32_t)n_in << 8);
    plane_store(C, XG, npl, blk, pp, X);
    if constexpr (kAcc) {
      const int ring = ((63 - pp) + e + kBpBiasDev) & (kMaxKeys - 1);
      atomicAdd(&C.kbits[ring], (uint32_t)len);
      atomicMin(&C.kfirst[ring], blk);
    }
  }
}

// Block stage of one subband staged in C.v.V (sample (i, j, k) at
// (k * dy + j) * dx + i: dx, dy are the layout's row stride and rows per
// plane): per block exponent, planes and their (length, N); 4 blocks per wave.
// kSeg: the blocks belong to several bricks (block z = brick slot of the
// group, seg_dense[z] its dense id): exponent rows and key ranges per brick.
template <int DT, bool kSeg = false, class CL = CodeLds>
__device__ __forceinline__ void block_stage(const EncArgs& a, int sb, int dx, int dy, const uint32_t* blocks, int nblk, CL& C,
                            const Pos& p, int dense, u64* XG, const int* seg_dense = nullptr) {
  const BrickConst& k = a.k;
  const int lane = p.lane, q = lane >> 4, r = lane & 15;
  i64* T = &C.T[p.w][0];
  for (int g = p.w; g * 4 < nblk; g += kTh / 64) {
    const int b = g * 4 + q;
    const bool bin = b < nblk;
    const uint32_t bw = bin ? blocks[b] : (1u << 24);
    const int o3[3] = {int(bw & 31), int((bw >> 5) & 31), int((bw >> 10) & 31)};
    const int bd[3] = {int((bw >> 15) & 7), int((bw >> 18) & 7), int((bw >> 21) & 7)};
    const bool skip = !bin || ((bw >> 24) & 1);
    const int ci = int(bw >> 25);
    // V strides of the block's non-unit axes, in order (x before y before z)
    const bool ux = bd[0] > 1, uy = bd[1] > 1, uz = bd[2] > 1;
    const int nd = ux + uy + uz;
    const int dxy = dx * dy;
    const int s0 = ux ? 1 : (uy ? dx : dxy);
    const int s1 = ux ? (uy ? dx : dxy) : (uy ? dxy : 0);
    const int s2 = (ux && uy && uz) ? dxy : 0;
    const int i1 = r & 3, i2 = r >> 2;
    const bool rowok = !skip && (nd == 3 || (nd == 2 && i2 == 0) || (nd <= 1 && r == 0));
    const int n0 = nd >= 1 ? 4 : 1;
    const int vbase = (o3[2] * dy + o3[1]) * dx + o3[0] + (nd >= 2 ? i1 * s1 : 0) + (nd == 3 ? i2 * s2 : 0);
    // common exponent as QuantizeF32/F64 (Function.h:103-105): the largest
    // |value| first (per row, then over the block's rows), one exponent per
    // row; padding entries are 0 and do not change the maximum
    double v[4], m = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < 4; ++i0) {
      const bool ok = rowok && i0 < n0;
      v[i0] = ok ? C.v.V[vbase + i0 * s0] : 0.0;
      m = fmax(m, fabs(v[i0]));
    }
    int emx = rowok ? (DT ? exponent_f64(m) : exponent_f32((float)m)) : -4000;
    emx = row_max(emx);
    const int emax = emx;
    // block scalars
    const int stop = max(64 - k.endbp, 57 + a.etol - emax);
    const int np = skip ? 0 : (stop > 63 ? 0 : 64 - stop);
    if (r == 0 && bin) {
      C.E[b] = skip ? kSkip : (short)emax;
      C.np[b] = (unsigned char)np;
      if (!skip) {
        // exponent sub-channel (EncodeBrickSubbandExponents, idx2Encode.cpp:41-80)
        const int bdense = kSeg ? seg_dense[o3[2]] : dense;
        const long long row = a.exp_off[sb] + ((long long)bdense * a.ncoded[sb] + ci) * (DT ? 2 : 1);
        if (DT) {
          const uint16_t e16 = (uint16_t)(emax + 1023);
          gm(a.exps)[row] = (uint8_t)(e16 & 0xff);
          gm(a.exps)[row + 1] = (uint8_t)(e16 >> 8);
        } else {
          gm(a.exps)[row] = (uint8_t)(emax + 127);
        }
        if (np > 0) {
          if constexpr (kSeg) {
            atomicMax(&C.khi_s[o3[2]], 63 + emax + kBpBiasDev);
            atomicMin(&C.klo_s[o3[2]], stop + emax + kBpBiasDev);
          } else {
            atomicMax(&C.khi, 63 + emax + kBpBiasDev);
            atomicMin(&C.klo, stop + emax + kBpBiasDev);
          }
        }
      }
    }
    if (__builtin_amdgcn_ballot_w64(np > 0) == 0) continue;  // no block codes a plane
    const double scale = pow2d((63 - nd) - 1 - emax);
    i64 w4[4];
#pragma unroll
    for (int i0 = 0; i0 < 4; ++i0) w4[i0] = f64_to_i64(scale * v[i0]);
    if (nd >= 1) zfp_flift4(w4[0], w4[1], w4[2], w4[3]);
    i64* Tq = T + q * TEnc::q;
#pragma unroll
    for (int i0 = 0; i0 < 4; ++i0) Tq[i0 + 4 * i1 + TEnc::c * i2] = w4[i0];
    wave_sync_lds();
    {  // second axis: lane (j0, j2) lifts positions j0 + 4 i + 16 j2
      const int j0 = r & 3, j2 = r >> 2;
      i64 y4[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) y4[i] = Tq[j0 + 4 * i + TEnc::c * j2];
      if (nd >= 2) {
        zfp_flift4(y4[0], y4[1], y4[2], y4[3]);
#pragma unroll
        for (int i = 0; i < 4; ++i) Tq[j0 + 4 * i + TEnc::c * j2] = y4[i];
      }
    }
    wave_sync_lds();
    {  // third axis: lane (j0, j1) lifts positions j0 + 4 j1 + 16 i
      const int j0 = r & 3, j1 = r >> 2;
      i64 y4[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) y4[i] = Tq[j0 + 4 * j1 + TEnc::c * i];
      if (nd == 3) {
        zfp_flift4(y4[0], y4[1], y4[2], y4[3]);
#pragma unroll
        for (int i = 0; i < 4; ++i) Tq[j0 + 4 * j1 + TEnc::c * i] = y4[i];
      }
    }
    wave_sync_lds();
    // coefficient mode + bit transpose + plane mode
    const int nd_s[4] = {__builtin_amdgcn_readlane(nd, 0), __builtin_amdgcn_readlane(nd, 16),
                         __builtin_amdgcn_readlane(nd, 32), __builtin_amdgcn_readlane(nd, 48)};
    const int np_s[4] = {__builtin_amdgcn_readlane(np, 0), __builtin_amdgcn_readlane(np, 16),
                         __builtin_amdgcn_readlane(np, 32), __builtin_amdgcn_readlane(np, 48)};
    const int e_s[4] = {__builtin_amdgcn_readlane(emax, 0), __builtin_amdgcn_readlane(emax, 16),
                        __builtin_amdgcn_readlane(emax, 32), __builtin_amdgcn_readlane(emax, 48)};
    // f32, and f64 rounds whose blocks code at most 32 planes: only the top
    // 32 bits of the coefficients are coded, two blocks share a transpose
    const bool narrow = DT == 0 || max(max(np_s[0], np_s[1]), max(np_s[2], np_s[3])) <= 32;
    if (narrow) {
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) {
        const u64 uA = coef_u(T + (2 * pr) * TEnc::q, C, nd_s[2 * pr], lane);
        const u64 uB = coef_u(T + (2 * pr + 1) * TEnc::q, C, nd_s[2 * pr + 1], lane);
        const u64 R = (uA >> 32) | (uB & 0xFFFFFFFF00000000ull);
        const u64 X = bit_transpose(R, lane);
        const int slot = 2 * pr + (lane >> 5);
        const int blk = g * 4 + slot;
        const int nds = (lane >> 5) ? nd_s[2 * pr + 1] : nd_s[2 * pr];
        const int nps = (lane >> 5) ? np_s[2 * pr + 1] : np_s[2 * pr];
        const int es = (lane >> 5) ? e_s[2 * pr + 1] : e_s[2 * pr];
        plane_mode<!kSeg>(X, lane, 32, blk, blk < nblk, nps, nds, 31 - (lane & 31), C, XG, DT ? 64 : 32, es);
      }
    } else {
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const u64 u = coef_u(T + qq * TEnc::q, C, nd_s[qq], lane);
        const u64 X = bit_transpose(u, lane);
        const int blk = g * 4 + qq;
        plane_mode<!kSeg>(X, lane, 64, blk, blk < nblk, np_s[qq], nd_s[qq], 63 - lane, C, XG, 64, e_s[qq]);
      }
    }
    wave_sync_lds();
  }
}

// Code every subband of the brick held in c (register columns) and S
// (faces); writes exponent rows, streams into the arena, entries, headers.
// One subband of the brick: stage into V, block stage, key offsets, emission.
// seq: the subband's position in the brick's coding order.
template <int DT, bool kOdd>
__device__ __forceinline__ void code_subband(const EncArgs& a, int sb, int seq, const double (&c)[kN], const double* Z1,
                                             Lds& S, CodeLds& C, const Pos& p, int dense, u64* XG,
                                             const uint32_t* s_blocks, ArenaCursor& cur) {
  const BrickConst& k = a.k;
  constexpr int npl = DT ? 64 : 32;
  const int lane = p.lane;
  const int hx = k.sb[sb].f[0], hy = k.sb[sb].f[1];
  const int dx = 17 - hx, dy = 17 - hy;
  const int nblk = a.blk_n[sb];
  const uint32_t* blocks = s_blocks + sb * 128;
  if (kOdd) stage_subband_odd(Z1, S, C.v.V, p, hx, hy);
  else stage_subband_even(c, S, C.v.V, p, hx, hy);
  if (p.tid == 0) { C.klo = 0x7fffffff; C.khi = -0x7fffffff; }
  if (p.tid < kMaxKeys) { C.kbits[p.tid] = 0; C.kfirst[p.tid] = 0x7fffffff; }
  __syncthreads();
  IDX2G_PHASE(12);
  block_stage<DT>(a, sb, kVX, kVY, blocks, nblk, C, p, dense, XG);
  __syncthreads();
  IDX2G_PHASE(13);
  // ---- 4. per-key offsets over the blocks (Morton order) ----
  const int klo = C.klo, khi = C.khi;
  const int nk = khi >= klo ? khi - klo + 1 : 0;
  if (nk > kMaxKeys) {
    if (p.tid == 0) {
      atomicOr(&a.counters[2], 1ull);
      gm_store(a.hdrs + (size_t)dense * k.nsb + sb, SbHdr{0, 0, 0});
    }
    __syncthreads();
    return;
  }
  // the block stage has summed each key stream's bits: every wave lays the
  // streams out itself (an exclusive scan of at most kMaxKeys byte sizes),
  // so that the emission starts at once; wave 0 also reserves the arena
  // slice and writes the header, which only the copy below needs
  constexpr int M = kMaxKeys - 1;
  static_assert(kMaxKeys == 128, "keys scanned in two 64-lane chunks");
  const uint32_t by0 = lane < nk ? (C.kbits[(klo + lane) & M] + 7) >> 3 : 0u;
  const uint32_t by1 = lane + 64 < nk ? (C.kbits[(klo + lane + 64) & M] + 7) >> 3 : 0u;
  const uint32_t in0 = wave_scan_add(by0, lane);
  const uint32_t in1 = wave_scan_add(by1, lane) + (uint32_t)__builtin_amdgcn_readlane((int)in0, 63);
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)in1, 63);
  const uint32_t ex0 = in0 - by0, ex1 = in1 - by1;  // byte offsets of keys klo + lane, klo + lane + 64
  const uint32_t nwords = (total + 3) >> 2;
  const bool staged = nwords <= (uint32_t)kStreamWords;
  const int slot = seq & 1;
  if (p.w == 0) {
    unsigned long long bse = 0, ebs = 0;
    int over = 0;
    if (lane == 0) {
      const unsigned long long alloc = (total + 3ull) & ~3ull;
      if (alloc > cur.aleft) {
        const unsigned long long take = alloc > kArenaChunk ? alloc : kArenaChunk;
        cur.abase = atomicAdd(&a.counters[0], take);
        cur.aleft = take;
      }
      if ((unsigned long long)nk > cur.eleft) {
        const unsigned long long take = (unsigned long long)nk > kEntryChunk ? (unsigned long long)nk : kEntryChunk;
        cur.ebase = atomicAdd(&a.counters[1], take);
        cur.eleft = take;
      }
      bse = cur.abase;
      ebs = cur.ebase;
      cur.abase += alloc;
      cur.aleft -= alloc;
      cur.ebase += nk;
      cur.eleft -= nk;
      cur.bytes += total;
      over = bse + alloc > a.arena_cap || ebs + nk > a.entry_cap;
      if (over) atomicOr(&a.counters[2], 2ull);
      C.base_s[slot] = bse;
      C.over_s[slot] = over;
      SbHdr h;
      h.klo = klo;
      h.nkeys = over ? 0 : nk;
      h.ebase = ebs;
      gm_store(a.hdrs + (size_t)dense * k.nsb + sb, h);
    }
    // the stream entries, from lane 0's reservation
    bse = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(bse >> 32), 0) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bse, 0);
    ebs = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(ebs >> 32), 0) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)ebs, 0);
    over = __builtin_amdgcn_readlane(over, 0);
    if (!over)
#pragma unroll
      for (int h2 = 0; h2 < 2; ++h2) {
        const int kk = lane + 64 * h2;
        if (kk >= nk) continue;
        const int ring = (klo + kk) & M;
        SbEntry e;
        e.off = bse + (h2 ? ex1 : ex0);
        e.size = (C.kbits[ring] + 7) >> 3;
        e.first_block = (uint16_t)(C.kfirst[ring] == 0x7fffffff ? 0xffff : C.kfirst[ring]);
        e.key = (uint16_t)(klo + kk);
        gm_store(a.entries + ebs + kk, e);
      }
  }
  uint32_t* words = C.stream;
  if (!staged) {  // streams larger than the LDS buffer are assembled in the arena
    __syncthreads();
    if (C.over_s[slot]) return;
    words = (uint32_t*)(a.arena + C.base_s[slot]);
    for (uint32_t i = p.tid; i < nwords; i += kTh) words[i] = 0;
    __threadfence();
    __syncthreads();
  }
  // ---- 4+5. per key: exclusive scan of the plane lengths over the blocks
  // in Morton order gives each plane's bit offset in the key's stream, and
  // the same lane emits the plane's bits there ----
  for (int kk = p.w; kk < nk; kk += kTh / 64) {
    const int key = klo + kk;
    const int ku = __builtin_amdgcn_readfirstlane(kk);  // the wave's key (uniform)
    const uint32_t kbyte = (uint32_t)__builtin_amdgcn_readlane((int)(ku < 64 ? ex0 : ex1), ku & 63);
    const u64 sbit = (u64)kbyte * 8;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < nblk; c0 += 64) {
      const int b = c0 + lane;
      const int e = b < nblk ? C.E[b] : kSkip;
      const int pp = 63 - (key - kBpBiasDev - e);
      const bool coded = b < nblk && e != kSkip && pp >= 0 && pp < C.np[b];
      const uint32_t lno = coded ? lno_at(C, npl, b, pp) : 0u;
      const uint32_t len = lno & 0xff;
      const uint32_t incl = wave_scan_add(len, lane);
      if (coded) {
        const uint32_t bw = blocks[b];
        const int nd = (((bw >> 15) & 7) > 1) + (((bw >> 18) & 7) > 1) + (((bw >> 21) & 7) > 1);
        u64 lo, hi;
        plane_bits(plane_load(C, XG, npl, b, pp), int((lno >> 8) & 0x7f), 1 << (2 * nd), C.e8, lo, hi);
        const u64 pos = sbit + incl - len + carry;
        if (staged) or_bits128<true>(words, pos, lo, hi);
        else or_bits128<false>(words, pos, lo, hi);
      }
      carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
  }
  __syncthreads();
  IDX2G_PHASE(14);
  if (C.over_s[slot]) {  // no room in the arena (reported through the counters): leave the buffer zero
    if (staged)
      for (uint32_t i = p.tid; i < nwords; i += kTh) words[i] = 0;
    return;
  }
  uint32_t* gwords = (uint32_t*)(a.arena + C.base_s[slot]);
  // into the arena, leaving the LDS buffer zero for the next subband (its
  // emission comes after at least two more barriers)
  if (staged)
    for (uint32_t i = p.tid; i < nwords; i += kTh) {
      const uint32_t w = words[i];
      words[i] = 0;
      gm(gwords)[i] = w;
    }
}

template <int DT>
__device__ __forceinline__ void code_brick(const EncArgs& a, const double (&c)[kN], Lds& S, CodeLds& C, const Pos& p,
                                           int dense, u64* XG, const uint32_t* s_blocks, ArenaCursor& cur,
                                           const unsigned char* s_perm) {
  // s_perm: the job's coefficient orders, kept in the kernel's own LDS (CodeLds
  // shares LDS with the transform, so it takes them per brick)
  const BrickConst& k = a.k;
  if (p.tid < 64) C.p3[p.tid] = s_perm[p.tid];  // positions in the TEnc layout
  if (p.tid < 16) C.p2[p.tid] = s_perm[64 + p.tid];
  if (p.tid < 256) {
    uint32_t e = 0;
    int o = 0;
    for (int i = 0; i < 8; ++i) {
      if ((p.tid >> i) & 1) { e |= 3u << o; o += 2; }
      else o += 1;
    }
    C.e8[p.tid] = (uint16_t)e;
  }
  // the stream buffer shares LDS with the transform's row buffers: clear it
  // once per brick (the first subband's barriers order this before use)
  static_assert(kStreamWords % (4 * kTh) == 0, "stream buffer cleared in 16-byte units");
  for (int i = p.tid; i < kStreamWords / 4; i += kTh) reinterpret_cast<uint4*>(C.stream)[i] = uint4{0, 0, 0, 0};
  // the z-odd half leaves the registers first; the z-even subbands are coded
  // from the registers, the z-odd ones from the parked copy
  double* Z1 = reinterpret_cast<double*>(XG) + kOddOff;
  park_odd(c, Z1, p);
  int seq = 0;
  for (int sb = 0; sb < 8; ++sb)
    if (k.sb[sb].f[2] == 0) code_subband<DT, false>(a, sb, seq++, c, Z1, S, C, p, dense, XG, s_blocks, cur);
  for (int sb = 0; sb < 8; ++sb)
    if (k.sb[sb].f[2] == 1) code_subband<DT, true>(a, sb, seq++, c, Z1, S, C, p, dense, XG, s_blocks, cur);
}

}  // namespace b32
}  // namespace idx2g
