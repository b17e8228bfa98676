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
struct BitBuf {
  u64 b0 = 0, b1 = 0;
  u64 base = 1ull << 63;  // stream bit of b0's bit 0 (dword aligned); none yet
};

// 64 bits of the stream at bit pos (words padded by >= 16 bytes)
__device__ __forceinline__ u64 bb_peek(const uint32_t* w32, BitBuf& B, u64 pos) {
  u64 sh = pos - B.base;
  if (sh > 64) {  // outside the buffer (also pos < base): refill at pos's word
    const uint32_t* w = w32 + (pos >> 5);
    B.b0 = (u64)w[0] | ((u64)w[1] << 32);
    B.b1 = (u64)w[2] | ((u64)w[3] << 32);
    B.base = pos & ~31ull;
    sh = pos & 31;
  }
  return sh == 0 ? B.b0 : sh == 64 ? B.b1 : (B.b0 >> sh) | (B.b1 << (64 - sh));
}

// One bit plane (Decode, Zfp.h:988-1059) from 32-bit words: plane word.
// A 64-bit window is re-read only when the next group test may need bits
// beyond it.
__device__ __forceinline__ u64 decode_plane_win(const uint32_t* w32, BitBuf& B, u64& pos, int& n, int nv) {
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
    const int z = w ? __ffsll((long long)w) - 1 : 64;
    int adv;
    if (z < lim) { n += z; adv = z + 1; }
    else { n = nv - 1; adv = lim; }
    pos += adv;
    w = adv >= 64 ? 0 : w >> adv;
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

// z-odd subband (hx, hy, 1) from V into the workgroup's parking area
// (Z1[((hx + 2 hy) * 16 + k) * 256 + j * 16 + i] = column (2i + hx, 2j + hy)
// at z = 2k + 1); the columns take them after every subband is decoded.
__device__ __forceinline__ void park_subband_odd(double* Z1, Lds& S, const double* V, const Pos& p, int hx, int hy) {
  const auto dst = gm(Z1) + (hx + 2 * hy) * 16 * 256;
  for (int t = p.tid; t < 16 * 256; t += kTh) {
    const int k = t >> 8, j = (t >> 4) & 15, i = t & 15;
    dst[t] = V[(k * kVY + j) * kVX + i];
  }
  unstage_faces(S, V, p, hx, hy, 1);
}

__device__ __forceinline__ void unpark_odd(double (&c)[kN], const double* Z1, const Pos& p) {
  const auto q = gm(Z1) + ((p.x & 1) + 2 * (p.y & 1)) * 16 * 256 + (p.y >> 1) * 16 + (p.x >> 1);
#pragma unroll
  for (int k = 0; k < 16; ++k) c[2 * k + 1] = q[k * 256];
}

// Steps 1-2, one workgroup per brick (k_walk): stream staging and the
// pipelined walks.  Wave w walks subbands 2w and 2w + 1, one per 32-lane
// segment: lane l of a segment owns key g0 - l of the current group of 32
// keys.  The walk's shape is a template: 3D bricks (8 subbands of <= 128
// blocks, 256 threads) and 2D bricks (4 subbands of <= 32 blocks, 128
// threads, a small footprint so that many workgroups stay resident).
constexpr int kWalkThreads = 256;
constexpr int kWalkSlots = 64;                // staged streams per subband
// LDS stream buffer per brick.  Bricks whose streams do not fit walk them
// from HBM at nearly the same speed, and a small buffer leaves room for six
// resident workgroups per CU (measured on the 1024^3 float32 decode: 4 KB
// 24.4 ms, 20 KB 25.5 ms, 44 KB 29.3 ms).
constexpr int kWalkStreamBytes = 4 * 1024;

template <int NSB, int MB, int TH, int SBYTES, int OCC>
struct WalkShape {
  static constexpr int nsb = NSB;        // subbands
  static constexpr int mb = MB;          // blocks per subband (power of two)
  static constexpr int th = TH;          // threads: two subbands per wave
  static constexpr int sbytes = SBYTES;  // LDS stream buffer
  // waves per SIMD the register allocation is held to (the walk is VALU
  // bound and latency hidden by occupancy; measured on the 1024^3 float32
  // decode: 5 waves (92 VGPRs) 22.15 ms, 6 21.59 ms, 7 (72 VGPRs, the LDS
  // limit of the 3D shape) 21.31 ms; the 2D shape spills less at 6)
  static constexpr int occ = OCC;
  static_assert(2 * (TH / 64) == NSB, "one wave per two subbands");
};
using Walk3D = WalkShape<8, 128, kWalkThreads, kWalkStreamBytes, 7>;
using Walk2D = WalkShape<4, 32, 128, 2048, 6>;

template <class WS>
struct WalkLds {
  uint32_t sbuf[WS::sbytes / 4 + 8];
  u64 soff[WS::nsb][kWalkSlots];
  uint32_t sstart[WS::nsb][kWalkSlots];
  uint32_t ssize[WS::nsb][kWalkSlots];
  short E[WS::nsb][WS::mb];
  unsigned char nbps[WS::nsb][WS::mb];
  unsigned char nsave[WS::nsb][WS::mb];
  unsigned char ssave[WS::nsb][WS::mb];
  uint32_t wsum[WS::th / 64];
  int klo[WS::nsb], khi[WS::nsb];
  int total, staged;
};

// Exponent of block b of subband sb (EncodeBrickSubbandExponents layout),
// kSkip for blocks that are not coded.
__device__ __forceinline__ int block_exponent(const DecArgs& a, const uint32_t* s_blocks, int sb, int b, int dense,
                                              int stride = 128) {
  const uint32_t bw = s_blocks[sb * stride + b];
  if ((bw >> 24) & 1) return kSkip;
  const int ci = int(bw >> 25);
  const long long row = a.exp_off[sb] + ((long long)dense * a.ncoded[sb] + ci) * (a.k.dtype ? 2 : 1);
  return a.k.dtype ? int((uint32_t)a.exps[row] | ((uint32_t)a.exps[row + 1] << 8)) - 1023 : int(a.exps[row]) - 127;
}

// s_blocks: the level's block tables, [subband][WS::mb].  Plane words and
// counts go to the batch arrays in the 3D layout ([brick][8][128][npl] and
// [brick][8][128]) whatever the shape.
template <class WS>
__device__ __forceinline__ void walk_wg(const DecArgs& a, WalkLds<WS>& W, int bi, const uint32_t* s_blocks) {
  constexpr int NSB = WS::nsb, MB = WS::mb, TH = WS::th;
  const BrickConst& k = a.k;
  const int npl = k.dtype ? 64 : 32;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
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
    const int e = block_exponent(a, s_blocks, sb, b, dense, MB);
    W.E[sb][b] = (short)e;
    if (e == kSkip) continue;
    const int stop = max(64 - k.endbp, 57 + a.etol - e);
    if (stop <= 63 && ((a.mask >> sb) & 1)) {
      atomicMax(&W.khi[sb], 63 + e + kBpBiasDev);
      atomicMin(&W.klo[sb], stop + e + kBpBiasDev);
    }
  }
  __syncthreads();
  // directory entries of the first kWalkSlots keys of each subband, regions in LDS
  {
    uint32_t base = 0;
    for (int i0 = 0; i0 < NSB * kWalkSlots; i0 += TH) {
      const int i = i0 + tid, sb = i / kWalkSlots, kk = i % kWalkSlots;
      uint32_t sz = 0;
      u64 off = 0;
      if (sb < k.nsb && W.khi[sb] >= W.klo[sb]) {
        const int key = W.klo[sb] + kk;
        if (key <= W.khi[sb] && key >= a.dir_klo && key < a.dir_klo + a.dir_nk) {
          const u64 d = a.dir[((size_t)sb * a.dir_nk + (key - a.dir_klo)) * a.span + lin];
          if (d) { sz = dir_size(d); off = dir_offset(d); }
        }
      }
      W.ssize[sb][kk] = sz;
      W.soff[sb][kk] = off;
      const uint32_t reg = sz ? ((uint32_t)(off & 3) + sz + 3) & ~3u : 0u;
      const uint32_t incl = wave_scan_add(reg, lane);
      if (lane == 63) W.wsum[w] = incl;
      __syncthreads();
      uint32_t wb = base;
      for (int j = 0; j < w; ++j) wb += W.wsum[j];
      W.sstart[sb][kk] = wb + incl - reg + (uint32_t)(off & 3);
      for (int j = 0; j < TH / 64; ++j) base += W.wsum[j];
      __syncthreads();
    }
    if (tid == 0) {
      bool wide = false;
      for (int s = 0; s < k.nsb; ++s) wide |= W.khi[s] - W.klo[s] + 1 > kWalkSlots;
      W.total = (int)base;
      W.staged = !wide && base <= (uint32_t)WS::sbytes;
      atomicAdd(&a.counters[1], (unsigned long long)base);  // stream bytes this query reads
    }
  }
  __syncthreads();
  const bool staged = W.staged;
  if (staged) {
    for (int i = tid; i < NSB * kWalkSlots; i += TH) {
      const int sb = i / kWalkSlots, kk = i % kWalkSlots;
      const uint32_t sz = W.ssize[sb][kk];
      if (!sz) continue;
      const u64 off = W.soff[sb][kk];
      const uint32_t* src = reinterpret_cast<const uint32_t*>(a.blob) + (off >> 2);
      uint32_t* dst = W.sbuf + (W.sstart[sb][kk] >> 2);
      const uint32_t nw = ((uint32_t)(off & 3) + sz + 3) >> 2;
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
  const int seg = lane >> 5, sl = lane & 31;
  const int sb = 2 * w + seg;
  const bool sb_ok = sb < k.nsb && ((a.mask >> sb) & 1) && W.khi[sb] >= W.klo[sb];
  const int klo = sb_ok ? W.klo[sb] : 0, khi = sb_ok ? W.khi[sb] : -1;
  const int nblk = sb_ok ? a.blk_n[sb] : 0;
  const int sbi = sb < NSB ? sb : 0;
  const int ngr = sb_ok ? (khi - klo) / 32 + 1 : 0;
  const int ng = max(__builtin_amdgcn_readlane(ngr, 0), __builtin_amdgcn_readlane(ngr, 32));
  const uint32_t* blocks = s_blocks + sbi * MB;
  const auto xs = gm(a.xs) + ((size_t)bi * 8 + sbi) * 128 * npl;
  for (int gi = 0; gi < ng; ++gi) {
    const int g0 = khi - 32 * gi;
    const bool gact = sb_ok && g0 >= klo;
    const int key = g0 - sl;
    const bool kact = gact && key >= klo;
    const int kk = key - klo;
    u64 dentry = 0;
    if (staged) {
      if (kact && kk < kWalkSlots && W.ssize[sbi][kk]) dentry = 1;
    } else if (kact && key >= a.dir_klo && key < a.dir_klo + a.dir_nk) {
      dentry = gm(a.dir)[((size_t)sbi * a.dir_nk + (key - a.dir_klo)) * a.span + lin];
    }
    const bool have = dentry != 0;
    const uint32_t* w32 = staged ? W.sbuf : reinterpret_cast<const uint32_t*>(a.blob);
    u64 pos = 0;
    if (have) pos = staged ? (u64)W.sstart[sbi][kk] * 8 : dir_offset(dentry) * 8;
    BitBuf B;
    const int my_steps = gact ? nblk + min(32, g0 - klo + 1) : 0;
    const int nsteps = max(__builtin_amdgcn_readlane(my_steps, 0), __builtin_amdgcn_readlane(my_steps, 32));
    int n_out = 0, stop_out = 0;
    int e_next = (sl == 0 && nblk > 0) ? W.E[sbi][0] : kSkip;
    uint32_t bw_next = blocks[0];
    for (int t = 0; t < nsteps; ++t) {
      const int n_left = lane_prev_i(n_out);
      const int s_left = lane_prev_i(stop_out);
      const int b = t - sl;
      const bool inb = b >= 0 && b < nblk;
      const int e = e_next;
      const uint32_t bwc = bw_next;
      {
        const int bn = b + 1;
        const bool inn = bn >= 0 && bn < nblk;
        e_next = inn ? W.E[sbi][bn] : kSkip;
        bw_next = blocks[inn ? bn : 0];
      }
      if (kact && inb && e != kSkip) {
        const int n_in = sl == 0 ? (gi == 0 ? 0 : W.nsave[sbi][b]) : n_left;
        const int st_in = sl == 0 ? (gi == 0 ? 0 : W.ssave[sbi][b]) : s_left;
        const int bp = key - kBpBiasDev - e;
        const int stop = max(64 - k.endbp, 57 + a.etol - e);
        n_out = n_in;
        stop_out = st_in;
        if (bp <= 63 && bp >= stop && !st_in) {
          if (!have) {
            stop_out = 1;  // missing chunk: the block stops decoding planes
          } else {
            const int nd = (((bwc >> 15) & 7) > 1) + (((bwc >> 18) & 7) > 1) + (((bwc >> 21) & 7) > 1);
            int n = n_in;
            const u64 X = decode_plane_win(w32, B, pos, n, 1 << (2 * nd));
            n_out = n;
            xs[b * npl + (63 - bp)] = X;
            W.nbps[sbi][b] = (unsigned char)(64 - bp);
          }
        } else if (bp > 63) {
          n_out = 0;
        }
        if (sl == 31 || key == klo) {
          W.nsave[sbi][b] = (unsigned char)n_out;
          W.ssave[sbi][b] = (unsigned char)stop_out;
        }
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < NSB * MB; i += TH) gm(a.nbps)[(size_t)bi * 1024 + (i / MB) * 128 + i % MB] = (&W.nbps[0][0])[i];
}

// Step 3 for subband sb: plane words -> values in D.v.V (sample (i, j, k) at
// (k * dy + j) * dx + i), every sample of the table's blocks (the parent's
// interior is copied for subband 0 below the coarsest level).
template <int DT>
__device__ __forceinline__ void inverse_blocks(const DecArgs& a, int sb, int dx, int dy, DecLds32& D, const Pos& p,
                                               const uint32_t* s_blocks, const u64* XS) {
  // XS: this brick's plane words [subband][128][npl]
  const BrickConst& k = a.k;
  constexpr int npl = DT ? 64 : 32;
  const int lane = p.lane, q = lane >> 4, r = lane & 15;
  const int nblk = a.blk_n[sb];
  const uint32_t* blocks = s_blocks + sb * 128;
  const auto xs = gm(XS) + (size_t)sb * 128 * npl;
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
      for (int d = 0; d < 3; ++d) bypass |= (o3[d] % a.spacing[sb][d]) != 0;
      live = np > 0 && !bypass;
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
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) {
        const int slot = 2 * pr + (lane >> 5);
        const int blk = g * 4 + slot;
        const int pp = 31 - (lane & 31);
        const u64 X = (blk < nblk && pp < D.nbps[sb][blk]) ? xs[blk * (DT ? 64 : 32) + pp] : 0ull;
        const u64 R = bit_transpose(X, lane);
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
        const int blk = g * 4 + qq;
        const int pp = 63 - lane;
        const u64 X = (blk < nblk && pp < D.nbps[sb][blk]) ? xs[blk * 64 + pp] : 0ull;
        const u64 u = bit_transpose(X, lane);
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
      i64 w4[4];
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0) w4[i0] = Tq[i0 + 4 * i1 + TDec::c * i2];
      if (nd >= 1) zfp_ilift4(w4[0], w4[1], w4[2], w4[3]);
      const double scale = 1.0 / pow2d((63 - nd) - 1 - e);
      const int vbase = (o3[2] * dy + o3[1]) * dx + o3[0] + (nd >= 2 ? i1 * s1 : 0) + (nd == 3 ? i2 * s2 : 0);
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0)
        if (rowok && i0 < n0) D.v.V[vbase + i0 * s0] = live ? scale * i64_to_f64(w4[i0]) : 0.0;
    }
    wave_sync_lds();
  }
}

}  // namespace b32
}  // namespace idx2g
