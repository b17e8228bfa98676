// Decode of one full 32^3 brick by a 1024-thread workgroup, for gfx950
// (DecodeBrick, Source/Core/idx2Decode.cpp:238-321; DecodeSubband,
// idx2Decode.cpp:85-235; zfp Decode, Source/Core/Zfp.h:988-1059):
//   1. the brick's bit-plane streams (all subbands, all keys) are copied
//      from the chunk blob into LDS when they fit;
//   2. eight waves walk the streams, one subband each: lane j owns key
//      khi - j and decodes that plane of block t - j at step t, receiving
//      the block's significance count N from lane j - 1 (the plane above)
//      through a DPP lane shift; plane words go to a per-workgroup scratch;
//   3. per subband, 4 blocks per wave: a 64x64 bit transpose turns plane
//      words back into coefficients, inverse coefficient order and
//      negabinary, inverse zfp lifts (LDS transposes), dequantisation into a
//      dense LDS subband array, then into the register columns and faces;
//   4. inverse subband weights and the inverse CDF5/3 on chip (brick32.hpp);
//   5. the data region goes to the output grid or to the level's dense grid
//      that the next finer level reads its subband-0 interiors from.
#pragma once

#include "code32.hpp"

namespace idx2g {
namespace b32 {


struct DecLds32 {
  union {
    double V[kVXY * 17];  // one subband, kVX / kVY layout
  } v;
  i64 T[kTh / 64][4 * TDec::q];  // TDec layout
  short E[8][128];
  unsigned char nbps[8][128];  // planes decoded per block
  unsigned char p3[64], p2[16];
};

// Plane words of one (brick, subband) in the walk's batch array, along
// diagonals: plane pl (0: the block's top plane) of block b at word
// (b + pl) * npl + pl.  At one step the lanes of a walk segment decode planes
// pl0 + l of blocks b0 - l, which lie on one diagonal when the blocks'
// exponents agree: their 8-byte stores form one contiguous row instead of
// one cache line each.  The on-chip decode reads a block's planes from up to
// npl rows of the subband's region, which its workgroup reads whole.
__host__ __device__ constexpr size_t xs_region(int npl) { return size_t(128 + npl) * npl; }
__host__ __device__ inline int xs_at(int b, int pl, int npl) { return (b + pl) * npl + pl; }

// Past the decode kernel's LDS (faces + DecLds32): the brick's parent
// subband-0 interior, 16^3 doubles [z][y][x], loaded straight into LDS when
// the brick starts and copied into V when subband 0 is decoded.
constexpr size_t dec_parent_off() {
  return (offsetof(Lds, u) +
          (sizeof(DecLds32) > sizeof(Lds) - offsetof(Lds, u) ? sizeof(DecLds32) : sizeof(Lds) - offsetof(Lds, u)) + 15) &
         ~size_t(15);
}
constexpr size_t kDecParentBytes = 16 * 16 * 16 * sizeof(double);

// Issue the loads of the parent interior of brick b3 into PB: rows
// r = z * 16 + y of 16 doubles, eight rows (1 KB) per wave instruction, each
// lane 16 bytes (global_load_lds_dwordx4: no registers held while the loads
// are in flight).  The caller waits (vmcnt) before reading PB.
__device__ __forceinline__ void load_parent_lds(const double* pgrid, const int* pgdims, const int (&b3)[3], double* PB,
                                                const Pos& p) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r0 = p.w * 16 + h * 8;
    const int r = r0 + (p.lane >> 3), z = r >> 4, y = r & 15;
    const double* src = pgrid + ((size_t)(b3[2] * 16 + z) * pgdims[1] + (b3[1] * 16 + y)) * pgdims[0] + b3[0] * 16 +
                        (p.lane & 7) * 2;
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)(PB + r0 * 16), 16, 0, 0);
  }
}

// 128 bits of one lane's stream kept in registers: every lane of a walk
// reads a different stream, so each load instruction touches 64 cache lines;
// a 128-bit refill serves two to three planes instead of two 64-bit windows
// per plane.
// Positions are bits from the lane's stream pointer (a stream is far below
// 2^31 bits).
struct BitBuf {
  u64 b0 = 0, b1 = 0;
  uint32_t base = 1u << 31;  // stream bit of b0's bit 0 (dword aligned); none yet
};

// 64 bits of the stream at bit pos (words padded by >= 16 bytes)
template <class P>
__device__ __forceinline__ u64 bb_peek(P w32, BitBuf& B, uint32_t pos) {
  uint32_t sh = pos - B.base;
  if (sh > 64) {  // outside the buffer (also pos < base): refill at pos's word
    const auto w = w32 + (pos >> 5);
    B.b0 = (u64)w[0] | ((u64)w[1] << 32);
    B.b1 = (u64)w[2] | ((u64)w[3] << 32);
    B.base = pos & ~31u;
    sh = pos & 31;
  }
  return sh == 0 ? B.b0 : sh == 64 ? B.b1 : (B.b0 >> sh) | (B.b1 << (64 - sh));
}

// One bit plane (Decode, Zfp.h:988-1059) from 32-bit words: plane word.
// A 64-bit window is re-read only when the next group test may need bits
// beyond it.
template <class P>
__device__ __forceinline__ u64 decode_plane_win(P w32, BitBuf& B, uint32_t& pos, int& n, int nv) {
  u64 X = 0;
  u64 w = bb_peek(w32, B, pos);
  int avail = 64;
  if (n > 0) {
    if (n >= 64) {
      X = w;
      pos += 64;
      avail = 0;
    } else {
      X = w & ((1ull << n) - 1);
      pos += n;
      w >>= n;
      avail -= n;
    }
  }
  while (n < nv) {
    const int lim = nv - 1 - n;  // bits examined before the implicit last one
    if (avail < lim + 1) {
      w = bb_peek(w32, B, pos);
      avail = 64;
    }
    const bool sig = w & 1;
    pos += 1;
    w >>= 1;
    avail -= 1;
    if (!sig) break;  // group insignificant
    // zeros before the next one; bit 63 of w is clear after the shift, so
    // the sentinel there only stands in for "none" (63 >= lim: the implicit
    // last one)
    const int z = __builtin_ctzll(w | (1ull << 63));
    const bool expl = z < lim;
    const int adv = expl ? z + 1 : lim;  // <= 63
    n = expl ? n + z : nv - 1;
    pos += adv;
    w >>= adv;
    avail -= adv;
    X |= 1ull << n;
    ++n;
  }
  return X;
}

__device__ __forceinline__ int lane_prev_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }

// Subband values of V into the register columns and the faces (inverse of
// stage_subband).
template <int HZ>
__device__ __forceinline__ void unstage_cols(double (&c)[kN], const double* V, int i, int j, int dz) {
  const double* q = V + j * kVX + i;
#pragma unroll
  for (int k = 0; k < 17; ++k)
    if (k < dz) c[HZ + 2 * k] = q[k * kVXY];
}

// Faces of subband (hx, hy, hz) from V back to the x = 32 / y = 32 planes and
// the corner line.
__device__ __forceinline__ void unstage_faces(Lds& S, const double* V, const Pos& p, int hx, int hy, int hz) {
  const int dz = 17 - hz;
  for (int t = p.tid; t < 17 * 16 * 2 + 17; t += kTh) {
    if (t < 272) {
      if (hx) continue;
      const int k = t >> 4, j = t & 15;
      if (k < dz) S.ex[(hz + 2 * k) * 32 + hy + 2 * j] = V[(k * kVY + j) * kVX + 16];
    } else if (t < 544) {
      if (hy) continue;
      const int k = (t - 272) >> 4, i = (t - 272) & 15;
      if (k < dz) S.ey[(hz + 2 * k) * 32 + hx + 2 * i] = V[(k * kVY + 16) * kVX + i];
    } else {
      if (hx || hy) continue;
      const int k = t - 544;
      if (k < dz) S.cc[hz + 2 * k] = V[(k * kVY + 16) * kVX + 16];
    }
  }
}

// z-even subband (hx, hy, 0) from V into the register columns.
__device__ __forceinline__ void unstage_subband_even(double (&c)[kN], Lds& S, const double* V, const Pos& p, int hx,
                                                     int hy) {
  if ((p.x & 1) == hx && (p.y & 1) == hy) unstage_cols<0>(c, V, p.x >> 1, p.y >> 1, 17);
  unstage_faces(S, V, p, hx, hy, 0);
}

// One sample (i, j, k) of z-odd subband (hx, hy, 1) to where it waits for
// the register columns: the workgroup's parking area for i, j < 16
// (Z1[((hx + 2 hy) * 16 + k) * 256 + j * 16 + i] = column (2i + hx, 2j + hy)
// at z = 2k + 1; Z1s is the subband's region), else the x = 32 / y = 32
// faces and the corner line.
__device__ __forceinline__ void park_put(double* Z1s, Lds& S, int hx, int hy, int i, int j, int k, double v) {
  if (i < 16 && j < 16) gm(Z1s)[k * 256 + j * 16 + i] = v;
  else if (j < 16) S.ex[(1 + 2 * k) * 32 + hy + 2 * j] = v;
  else if (i < 16) S.ey[(1 + 2 * k) * 32 + hx + 2 * i] = v;
  else S.cc[1 + 2 * k] = v;
}

// A z-odd subband that the query does not decode: zeros where its samples go.
__device__ __forceinline__ void park_zero(double* Z1s, Lds& S, const Pos& p, int hx, int hy) {
  for (int t = p.tid; t < 16 * 256; t += kTh) gm(Z1s)[t] = 0.0;
  for (int t = p.tid; t < 16 * 17; t += kTh) {
    const int k = t / 17, j = t % 17;
    if (!hx && (j < 16 || !hy)) park_put(Z1s, S, hx, hy, 16, j, k, 0.0);  // x face, corner
    if (!hy && j < 16) park_put(Z1s, S, hx, hy, j, 16, k, 0.0);
  }
}

__device__ __forceinline__ void unpark_odd(double (&c)[kN], const double* Z1, const Pos& p) {
  const auto q = gm(Z1) + ((p.x & 1) + 2 * (p.y & 1)) * 16 * 256 + (p.y >> 1) * 16 + (p.x >> 1);
#pragma unroll
  for (int k = 0; k < 16; ++k) c[2 * k + 1] = q[k * 256];
}

// Steps 1-2, one workgroup per brick (k_walk): stream staging and the
// pipelined walks.  Every subband with planes to decode gets a segment of
// contiguous lanes, one lane per bit-plane key: as many lanes as the
// subband's keys span, at most 32 (a subband spanning more keys walks them
// in groups of 32).  The segments are packed first-fit, widest first, into
// 64-lane bins, and the workgroup's waves take the bins in turn, so a wave
// walks as many subbands as its lanes hold (a level-0 brick at tolerance
// 1e-4 spans 12-22 keys per subband: two bins instead of four waves of two
// 32-lane segments).  Lane l of a segment owns key g0 - l of the current
// group and decodes that plane of block t - l at step t, receiving the
// block's significance count N from lane l - 1 (the plane above) through a
// DPP lane shift.  The walk's shape is a template: 3D bricks (8 subbands of
// <= 128 blocks) and 2D bricks (4 subbands of <= 32 blocks).
constexpr int kWalkSlots = 32;  // staged key streams per subband
constexpr int kWalkCount = 64;  // keys per subband whose stream bytes are counted
// LDS stream buffer per brick.  Bricks whose streams do not fit walk them
// from HBM at nearly the same speed, and a small buffer leaves room for many
// resident workgroups per CU (measured on the 1024^3 float32 decode with
// 256-thread walks: 4 KB 24.4 ms, 20 KB 25.5 ms, 44 KB 29.3 ms).
constexpr int kWalkStreamBytes = 2 * 1024;

template <int NSB, int MB, int TH, int SBYTES, int OCC>
struct WalkShape {
  static constexpr int nsb = NSB;        // subbands
  static constexpr int mb = MB;          // blocks per subband (power of two)
  static constexpr int th = TH;          // threads
  static constexpr int sbytes = SBYTES;  // LDS stream buffer
  // waves per SIMD the register allocation is held to (the walk is VALU
  // bound and latency hidden by occupancy)
  static constexpr int occ = OCC;
  static_assert(NSB <= 8 && TH % 64 == 0 && (NSB * kWalkCount) % TH == 0, "walk shape");
};
using Walk3D = WalkShape<8, 128, 128, kWalkStreamBytes, 7>;
using Walk3DWide = WalkShape<8, 128, 256, kWalkStreamBytes, 7>;
using Walk2D = WalkShape<4, 32, 128, kWalkStreamBytes, 6>;

template <class WS>
struct WalkLds {
  uint32_t sbuf[WS::sbytes / 4 + 8];
  uint32_t sstart[WS::nsb][kWalkSlots];  // byte offset of a staged stream in sbuf (~0u: none)
  short E[WS::nsb][WS::mb];
  unsigned char nd[WS::nsb][WS::mb];     // block dimensions > 1 (the job's block table)
  unsigned char nbps[WS::nsb][WS::mb];
  unsigned char nsave[WS::nsb][WS::mb];
  unsigned char ssave[WS::nsb][WS::mb];
  uint32_t wsum[8];        // per-wave sums of the stream scan, then the bins' fill levels
  uint32_t seg_new[WS::nsb];
  uint32_t seg[WS::nsb];  // bin | first lane << 4 | lanes << 10 | groups << 16 (~0u: not walked)
  int klo[WS::nsb], khi[WS::nsb];
  int total, staged, nbins;
};

// Exponent of block b of subband sb (EncodeBrickSubbandExponents layout),
// kSkip for blocks that are not coded.
__device__ __forceinline__ int exponent_of_word(const DecArgs& a, uint32_t bw, int sb, int dense) {
  if ((bw >> 24) & 1) return kSkip;
  const int ci = int(bw >> 25);
  const long long row = a.exp_off[sb] + ((long long)dense * a.ncoded[sb] + ci) * (a.k.dtype ? 2 : 1);
  return a.k.dtype ? int((uint32_t)a.exps[row] | ((uint32_t)a.exps[row + 1] << 8)) - 1023 : int(a.exps[row]) - 127;
}
__device__ __forceinline__ int block_exponent(const DecArgs& a, const uint32_t* s_blocks, int sb, int b, int dense,
                                              int stride = 128) {
  return exponent_of_word(a, s_blocks[sb * stride + b], sb, dense);
}

// The job's block dimensions into W.nd (once per job).
template <class WS>
__device__ __forceinline__ void walk_load_blocks(const DecArgs& a, WalkLds<WS>& W) {
  for (int i = threadIdx.x; i < WS::nsb * WS::mb; i += WS::th) {
    const int sb = i / WS::mb, b = i % WS::mb;
    const uint32_t bw = b < a.blk_n[sb] ? a.blocks[a.blk_off[sb] + b] : 0u;
    W.nd[sb][b] = (unsigned char)((((bw >> 15) & 7) > 1) + (((bw >> 18) & 7) > 1) + (((bw >> 21) & 7) > 1));
  }
}

// The pipelined walk of every segment of one bin; P: the streams (LDS
// buffer or the chunk blob in HBM).
template <class WS, class P>
__device__ __forceinline__ void walk_bin(const DecArgs& a, WalkLds<WS>& W, int bi, int job, int npl, P w32) {
  const int lane = threadIdx.x & 63;
  const u64 lin = a.lin[bi];
  // the arguments the steps use, read once (the plane-word stores could
  // alias them as far as the compiler knows)
  const int endbp = a.k.endbp, etol = a.etol;
  const int dir_klo = a.dir_klo, dir_nk = a.dir_nk;
  const u64 span = a.span;
  const auto dir = gm(a.dir);
  int sb = -1, sl = 0, wdt = 32, ngs = 0;
#pragma unroll
  for (int s = 0; s < WS::nsb; ++s) {
    const uint32_t d = W.seg[s];
    const int off = int((d >> 4) & 63), n = int((d >> 10) & 63);
    if (d != ~0u && int(d & 15) == job && lane >= off && lane < off + n) {
      sb = s;
      sl = lane - off;
      wdt = n;
      ngs = int(d >> 16);
    }
  }
  const bool sb_ok = sb >= 0;
  const int sbi = sb_ok ? sb : 0;
  const int klo = sb_ok ? W.klo[sbi] : 0, khi = sb_ok ? W.khi[sbi] : -1;
  const int nblk = sb_ok ? a.blk_n[sbi] : 0;
  const int ng = wave_max_i(ngs);
  const auto xs = gm(a.xs) + ((size_t)bi * 8 + sbi) * xs_region(npl);
  constexpr bool kStaged = !std::is_same<P, const __attribute__((address_space(1))) uint32_t*>::value;
  for (int gi = 0; gi < ng; ++gi) {
    const int g0 = khi - wdt * gi;
    const bool gact = sb_ok && gi < ngs;
    const int key = g0 - sl;
    const bool kact = gact && key >= klo;
    const int kk = key - klo;
    bool have = false;
    uint32_t pos = 0;  // bits from ws
    P ws = w32;
    if (kStaged) {
      if (kact && kk < kWalkSlots) {
        const uint32_t st = W.sstart[sbi][kk];
        have = st != ~0u;
        pos = have ? st * 8 : 0;
      }
    } else if (kact && key >= dir_klo && key < dir_klo + dir_nk) {
      const u64 d = dir[((size_t)sbi * dir_nk + (key - dir_klo)) * span + lin];
      have = d != 0;
      const u64 off = have ? dir_offset(d) : 0;
      ws = w32 + (off >> 2);
      pos = uint32_t(off & 3) * 8;
    }
    BitBuf B;
    const int my_steps = gact ? nblk + min(wdt, g0 - klo + 1) : 0;
    const int nsteps = wave_max_i(my_steps);
    int n_out = 0, stop_out = 0;
    int e_next = (sl == 0 && nblk > 0) ? W.E[sbi][0] : kSkip;
    int nd_next = W.nd[sbi][0];
    for (int t = 0; t < nsteps; ++t) {
      const int n_left = lane_prev_i(n_out);
      const int s_left = lane_prev_i(stop_out);
      const int b = t - sl;
      const bool inb = b >= 0 && b < nblk;
      const int e = e_next;
      const int nd = nd_next;
      {
        const int bn = b + 1;
        const bool inn = bn >= 0 && bn < nblk;
        e_next = inn ? W.E[sbi][bn] : kSkip;
        nd_next = W.nd[sbi][inn ? bn : 0];
      }
      // every lane runs the plane decoder, with no group to decode (nv 0)
      // where it has no plane at this step: the stream state then stays in
      // place instead of being copied around a branch (the decoder only
      // peeks at the lane's current position, inside the padded streams)
      const bool act = kact && inb && e != kSkip;
      const int bs = inb ? b : 0;
      const int n_in = sl == 0 ? (gi == 0 ? 0 : W.nsave[sbi][bs]) : n_left;
      const int st_in = sl == 0 ? (gi == 0 ? 0 : W.ssave[sbi][bs]) : s_left;
      const int bp = key - kBpBiasDev - e;
      const int stop = max(64 - endbp, 57 + etol - e);
      const bool inplane = act && bp <= 63 && bp >= stop && !st_in;
      const bool dec = inplane && have;
      int n = dec ? n_in : 0;
      const u64 X = decode_plane_win(ws, B, pos, n, dec ? 1 << (2 * nd) : 0);
      if (act) {
        n_out = dec ? n : (bp > 63 ? 0 : n_in);
        stop_out = inplane && !have ? 1 : st_in;  // missing chunk: the block stops decoding planes
        if (dec) {
          xs[xs_at(b, 63 - bp, npl)] = X;
          W.nbps[sbi][b] = (unsigned char)(64 - bp);
        }
        if (sl == wdt - 1 || key == klo) {
          W.nsave[sbi][b] = (unsigned char)n_out;
          W.ssave[sbi][b] = (unsigned char)stop_out;
        }
      }
    }
  }
}

// Plane words and counts go to the batch arrays in the 3D layout
// ([brick][8][128][npl] and [brick][8][128]) whatever the shape.
template <class WS>
__device__ __forceinline__ void walk_wg(const DecArgs& a, WalkLds<WS>& W, int bi) {
  constexpr int NSB = WS::nsb, MB = WS::mb, TH = WS::th;
  const BrickConst& k = a.k;
  const int npl = k.dtype ? 64 : 32;
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);  // w: uniform
  const u64 lin = a.lin[bi];
  const int dense = a.dense[bi];
  if (tid < NSB) { W.klo[tid] = 0x7fffffff; W.khi[tid] = -0x7fffffff; }
  __syncthreads();
  for (int i = tid; i < NSB * MB; i += TH) {
    const int sb = i / MB, b = i % MB;
    W.nbps[sb][b] = 0;
    W.nsave[sb][b] = 0;
    W.ssave[sb][b] = 0;
    if (sb >= k.nsb || b >= a.blk_n[sb]) continue;
    const int e = exponent_of_word(a, a.blocks[a.blk_off[sb] + b], sb, dense);
    W.E[sb][b] = (short)e;
    if (e == kSkip) continue;
    const int stop = max(64 - k.endbp, 57 + a.etol - e);
    if (stop <= 63 && ((a.mask >> sb) & 1)) {
      atomicMax(&W.khi[sb], 63 + e + kBpBiasDev);
      atomicMin(&W.klo[sb], stop + e + kBpBiasDev);
    }
  }
  __syncthreads();
  // stream regions of the first kWalkCount keys of each subband: the bytes
  // this query reads, and the LDS offsets of the first kWalkSlots
  {
    uint32_t base = 0;
    for (int i0 = 0; i0 < NSB * kWalkCount; i0 += TH) {
      const int i = i0 + tid, sb = i / kWalkCount, kk = i % kWalkCount;
      uint32_t sz = 0;
      u64 off = 0;
      if (sb < k.nsb && W.khi[sb] >= W.klo[sb]) {
        const int key = W.klo[sb] + kk;
        if (key <= W.khi[sb] && key >= a.dir_klo && key < a.dir_klo + a.dir_nk) {
          const u64 d = a.dir[((size_t)sb * a.dir_nk + (key - a.dir_klo)) * a.span + lin];
          if (d) { sz = dir_size(d); off = dir_offset(d); }
        }
      }
      const uint32_t reg = sz ? ((uint32_t)(off & 3) + sz + 3) & ~3u : 0u;
      const uint32_t incl = wave_scan_add(reg, lane);
      if (lane == 63) W.wsum[w] = incl;
      __syncthreads();
      uint32_t wb = base;
      for (int j = 0; j < w; ++j) wb += W.wsum[j];
      if (kk < kWalkSlots) W.sstart[sb][kk] = sz ? wb + incl - reg + (uint32_t)(off & 3) : ~0u;
      for (int j = 0; j < TH / 64; ++j) base += W.wsum[j];
      __syncthreads();
    }
    if (tid == 0) {
      bool wide = false;
      for (int s = 0; s < k.nsb; ++s) wide |= W.khi[s] - W.klo[s] + 1 > kWalkSlots;
      W.total = (int)base;
      W.staged = !wide && base <= (uint32_t)WS::sbytes;
      atomicAdd(&a.counters[1], (unsigned long long)base);  // stream bytes this query reads
      // segments, widest first, first fit into 64-lane bins (the bins' fill
      // levels in W.wsum, which the scan above is done with)
      int nb = 0;
      uint32_t placed = 0;
      for (int i = 0; i < NSB; ++i) {
        int s = -1, ws = 0;
        for (int t = 0; t < NSB; ++t) {  // the widest subband not yet placed
          const bool act = t < k.nsb && ((a.mask >> t) & 1) && W.khi[t] >= W.klo[t] && !((placed >> t) & 1);
          const int wt = act ? min(W.khi[t] - W.klo[t] + 1, 32) : 0;
          if (wt > ws) { ws = wt; s = t; }
        }
        if (s < 0) break;
        placed |= 1u << s;
        int bin = 0;
        while (bin < nb && W.wsum[bin] + ws > 64) ++bin;
        if (bin == nb) W.wsum[nb++] = 0;
        const int ng = (W.khi[s] - W.klo[s]) / ws + 1;
        W.seg_new[s] = (uint32_t)bin | W.wsum[bin] << 4 | (uint32_t)ws << 10 | (uint32_t)ng << 16;
        W.wsum[bin] += ws;
      }
      for (int t = 0; t < NSB; ++t) W.seg[t] = ((placed >> t) & 1) ? W.seg_new[t] : ~0u;
      W.nbins = nb;
    }
  }
  __syncthreads();
  const bool staged = W.staged;
  if (staged) {
    for (int i = tid; i < NSB * kWalkSlots; i += TH) {
      const int sb = i / kWalkSlots, kk = i % kWalkSlots;
      const uint32_t st = W.sstart[sb][kk];
      if (st == ~0u) continue;
      const u64 d = a.dir[((size_t)sb * a.dir_nk + (W.klo[sb] + kk - a.dir_klo)) * a.span + lin];
      const u64 off = dir_offset(d);
      const uint32_t* src = reinterpret_cast<const uint32_t*>(a.blob) + (off >> 2);
      uint32_t* dst = W.sbuf + (st >> 2);
      const uint32_t nw = ((uint32_t)(off & 3) + dir_size(d) + 3) >> 2;
      uint32_t j = 0;
      for (; j + 8 <= nw; j += 8) {
        uint32_t t8[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) t8[q] = src[j + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) dst[j + q] = t8[q];
      }
      for (; j < nw; ++j) dst[j] = src[j];
    }
    if (tid < 4) W.sbuf[((uint32_t)W.total >> 2) + tid] = 0;
  }
  __syncthreads();
  // ---- the walks ----
  const int nbins = W.nbins;
  for (int job = w; job < nbins; job += TH / 64) {
    if (staged) walk_bin<WS>(a, W, bi, job, npl, static_cast<const uint32_t*>(W.sbuf));
    else walk_bin<WS>(a, W, bi, job, npl, gm(reinterpret_cast<const uint32_t*>(a.blob)));
  }
  __syncthreads();
  static_assert(MB % 4 == 0, "nbps rows copied in words");
  const uint32_t* nb32 = reinterpret_cast<const uint32_t*>(&W.nbps[0][0]);
  for (int i = tid; i < NSB * MB / 4; i += TH) {
    const int sb = (4 * i) / MB, b = (4 * i) % MB;
    gm(reinterpret_cast<uint32_t*>(a.nbps + (size_t)bi * 1024 + sb * 128 + b))[0] = nb32[i];
  }
}

// Step 3 for subband sb: plane words -> values in D.v.V (sample (i, j, k) at
// (k * dy + j) * dx + i), every sample of the table's blocks (the parent's
// interior is copied for subband 0 below the coarsest level).
// kPark: subband (hx, hy, 1), every sample straight to park_put (no V, no
// workgroup barrier: a wave's blocks touch nothing another wave reads).
template <int DT, bool kPark = false>
__device__ __forceinline__ void inverse_blocks(const DecArgs& a, int sb, int dx, int dy, DecLds32& D, const Pos& p,
                                               const uint32_t* s_blocks, const u64* XS, double* Z1s = nullptr,
                                               Lds* S = nullptr, int hx = 0, int hy = 0) {
  // XS: this brick's plane words, [subband] regions of xs_region(npl)
  const BrickConst& k = a.k;
  constexpr int npl = DT ? 64 : 32;
  const int lane = p.lane, q = lane >> 4, r = lane & 15;
  const int nblk = a.blk_n[sb];
  const uint32_t* blocks = s_blocks + sb * 128;
  const auto xs = gm(XS) + (size_t)sb * xs_region(npl);
  i64* T = &D.T[p.w][0];
  for (int g = p.w; g * 4 < nblk; g += kTh / 64) {
    // this lane's row-mode block
    const int b = g * 4 + q;
    const bool bin = b < nblk;
    const uint32_t bw = bin ? blocks[b] : (1u << 24);
    const int o3[3] = {int(bw & 31), int((bw >> 5) & 31), int((bw >> 10) & 31)};
    const int bd[3] = {int((bw >> 15) & 7), int((bw >> 18) & 7), int((bw >> 21) & 7)};
    const bool ux = bd[0] > 1, uy = bd[1] > 1, uz = bd[2] > 1;
    const int nd = ux + uy + uz;
    const int e = bin ? D.E[sb][b] : kSkip;
    bool live = bin && e != kSkip;
    if (live) {
      const int stop = max(64 - k.endbp, 57 + a.etol - e);
      const int np = stop > 63 ? 0 : 64 - stop;
      bool bypass = false;
      for (int d = 0; d < 3; ++d) bypass |= (o3[d] & a.spmask[sb][d]) != 0;
      live = np > 0 && !bypass;
    }
    // row mode addressing of the first-axis pass: lane (i1, i2) of its block
    // writes the row of samples vbase + i0 * s0
    const int dxy = dx * dy;
    const int s0 = ux ? 1 : (uy ? dx : dxy);
    const int s1 = ux ? (uy ? dx : dxy) : (uy ? dxy : 0);
    const int s2 = (ux && uy && uz) ? dxy : 0;
    const int i1 = r & 3, i2 = r >> 2;
    // every block of the table writes its samples (zeros when it is not
    // decoded), so V needs no zero fill (the blocks tile the subband but for
    // the subband-0 interior, which holds the parent's data)
    const bool rowok = bin && (nd == 3 || (nd == 2 && i2 == 0) || (nd <= 1 && r == 0));
    const int n0 = nd >= 1 ? 4 : 1;
    const int vbase = (o3[2] * dy + o3[1]) * dx + o3[0] + (nd >= 2 ? i1 * s1 : 0) + (nd == 3 ? i2 * s2 : 0);
    // kPark: the row's first sample (pi, pj, pk) and its axis ra (the same
    // samples as vbase + i0 * s0)
    const int ra = ux ? 0 : (uy ? 1 : 2);
    const bool a1y = ux && uy;  // second unit axis y (else z)
    const int pi = o3[0], pj = o3[1] + (nd >= 2 && a1y ? i1 : 0),
              pk = o3[2] + (nd >= 2 && !a1y ? i1 : 0) + (nd == 3 ? i2 : 0);
    auto put = [&](int i0, double v) {
      if constexpr (kPark)
        park_put(Z1s, *S, hx, hy, pi + (ra == 0 ? i0 : 0), pj + (ra == 1 ? i0 : 0), pk + (ra == 2 ? i0 : 0), v);
      else
        D.v.V[vbase + i0 * s0] = v;
    };
    if (__builtin_amdgcn_ballot_w64(live) == 0) {
      // none of the four blocks is decoded: zeros, no transposes or lifts
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0)
        if (rowok && i0 < n0) put(i0, 0.0);
      continue;
    }
    const int nd_s[4] = {__builtin_amdgcn_readlane(nd, 0), __builtin_amdgcn_readlane(nd, 16),
                         __builtin_amdgcn_readlane(nd, 32), __builtin_amdgcn_readlane(nd, 48)};
    // plane words -> coefficients (bit transpose), inverse order, negabinary;
    // f32, and f64 rounds whose blocks have at most 32 planes decoded: two
    // blocks share a transpose of their top 32 planes
    const int nbq = bin ? (int)D.nbps[sb][b] : 0;
    const bool narrow = DT == 0 || max(max(__builtin_amdgcn_readlane(nbq, 0), __builtin_amdgcn_readlane(nbq, 16)),
                                       max(__builtin_amdgcn_readlane(nbq, 32), __builtin_amdgcn_readlane(nbq, 48))) <= 32;
    if (narrow) {
      // both plane words first (no branch around the loads: an absent plane
      // reads word 0 of the region and is zeroed), so that their latencies
      // overlap
      u64 Xp[2];
      bool okp[2];
      const u64* ap[2];
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) {
        const int blk = g * 4 + 2 * pr + (lane >> 5);
        const int pp = 31 - (lane & 31);
        okp[pr] = blk < nblk && pp < D.nbps[sb][blk < nblk ? blk : 0];
        ap[pr] = XS + (size_t)sb * xs_region(npl) + (okp[pr] ? xs_at(blk, pp, npl) : 0);
      }
      // the compiler, at its register limit here, would wait for the first
      // load before issuing the second; both in one block, waited for in it
      asm volatile(
          "global_load_dwordx2 %0, %2, off\n\t"
          "global_load_dwordx2 %1, %3, off\n\t"
          "s_waitcnt vmcnt(0)"
          : "=&v"(Xp[0]), "=&v"(Xp[1])
          : "v"(ap[0]), "v"(ap[1])
          : "memory");
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) Xp[pr] = okp[pr] ? Xp[pr] : 0ull;
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) {
        const u64 R = bit_transpose(Xp[pr], lane);
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
      u64 Xq[4];
      bool okq[4];
      const u64* aq[4];
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const int blk = g * 4 + qq;
        const int pp = 63 - lane;
        okq[qq] = blk < nblk && pp < D.nbps[sb][blk < nblk ? blk : 0];
        aq[qq] = XS + (size_t)sb * xs_region(npl) + (okq[qq] ? xs_at(blk, pp, 64) : 0);
      }
      asm volatile(
          "global_load_dwordx2 %0, %4, off\n\t"
          "global_load_dwordx2 %1, %5, off\n\t"
          "global_load_dwordx2 %2, %6, off\n\t"
          "global_load_dwordx2 %3, %7, off\n\t"
          "s_waitcnt vmcnt(0)"
          : "=&v"(Xq[0]), "=&v"(Xq[1]), "=&v"(Xq[2]), "=&v"(Xq[3])
          : "v"(aq[0]), "v"(aq[1]), "v"(aq[2]), "v"(aq[3])
          : "memory");
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) Xq[qq] = okq[qq] ? Xq[qq] : 0ull;
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const u64 u = bit_transpose(Xq[qq], lane);
        const int ndq = nd_s[qq], c = lane;
        if (c < (1 << (2 * ndq))) {
          const int dst = ndq == 3 ? D.p3[c] : ndq == 2 ? D.p2[c] : c;
          T[qq * TDec::q + dst] = (i64)((u ^ IDX2G_NBMASK) - IDX2G_NBMASK);
        }
      }
    }
    wave_sync_lds();
    i64* Tq = T + q * TDec::q;
    {  // third axis (inverse lifts run z, y, x)
      const int j0 = r & 3, j1 = r >> 2;
      i64 y4[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) y4[i] = Tq[j0 + 4 * j1 + TDec::c * i];
      if (nd == 3) {
        zfp_ilift4(y4[0], y4[1], y4[2], y4[3]);
#pragma unroll
        for (int i = 0; i < 4; ++i) Tq[j0 + 4 * j1 + TDec::c * i] = y4[i];
      }
    }
    wave_sync_lds();
    {  // second axis
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
    wave_sync_lds();
    // row mode: first axis, dequantise, into V (Dequantize, Function.h:139-149)
    {
      i64 w4[4];
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0) w4[i0] = Tq[i0 + 4 * i1 + TDec::c * i2];
      if (nd >= 1) zfp_ilift4(w4[0], w4[1], w4[2], w4[3]);
      const double scale = inv_pow2d((63 - nd) - 1 - e);
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0)
        if (rowok && i0 < n0) put(i0, live ? scale * i64_to_f64(w4[i0]) : 0.0);
    }
    wave_sync_lds();
  }
}

}  // namespace b32
}  // namespace idx2g
