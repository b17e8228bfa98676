// Decode kernels for gfx950: one workgroup per brick runs DecodeBrick
// (Source/Core/idx2Decode.cpp:238-321): parent subband-0 copy, per-subband
// bit-plane decoding (DecodeSubband, idx2Decode.cpp:85-235), inverse zfp and
// dequantisation, InverseCdf53, and the copy into the output grid
// (idx2Decode.cpp:431-445).
//
// Bit-plane streams of one brick-subband are walked by one wave: lane j owns
// the key (chunk channel) khi - j and decodes that plane of block t - j at
// step t, receiving the block's significance count N from lane j-1 (the next
// higher plane) through a lane shuffle.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "dec32.hpp"
#include "dec2d.hpp"

namespace idx2g {

// 64 bits of the blob starting at bit `pos` (blob padded by >= 16 bytes)
__device__ __forceinline__ u64 peek64(const uint8_t* blob, u64 pos) {
  const uint32_t* w = (const uint32_t*)blob + (pos >> 5);
  const int sh = int(pos & 31);
  const u64 lo = (u64)w[0] | ((u64)w[1] << 32);
  if (!sh) return lo;
  return (lo >> sh) | ((u64)w[2] << (64 - sh));
}

// One bit plane (Decode, Source/Core/Zfp.h:988-1059): returns the plane word.
__device__ __forceinline__ u64 decode_plane(const uint8_t* blob, u64& pos, int& n, int nv) {
  u64 X = 0;
  if (n > 0) {
    const u64 w = peek64(blob, pos);
    X = n >= 64 ? w : (w & ((1ull << n) - 1));
    pos += n;
  }
  while (n < nv) {
    u64 w = peek64(blob, pos);
    if (!(w & 1)) { pos += 1; break; }   // group insignificant
    pos += 1;
    w >>= 1;
    const int lim = nv - 1 - n;          // bits examined before the implicit last one
    const int z = w ? __ffsll((long long)w) - 1 : 64;
    if (z < lim) { n += z; pos += z + 1; }
    else { pos += lim; n = nv - 1; }
    X |= 1ull << n;
    ++n;
  }
  return X;
}

struct DecLds {
  u64* pw;              // [mb][npl]
  short* emax;          // [mb]
  unsigned char* nbps;  // [mb] planes decoded
  unsigned char* nsave; // [mb] N after the last key group
  unsigned char* ssave; // [mb] stopped flag after the last key group
};

__host__ __device__ inline size_t dec_lds_bytes(int npl, int mb) {
  return size_t(mb) * npl * 8 + size_t(mb) * (2 + 3) + 64;
}

// base: LDS, or the workgroup's scratch slot past the extended brick when
// the tables do not fit LDS (bricks above 32 samples per axis)
__device__ inline DecLds dcarve(unsigned char* base, int npl, int mb) {
  DecLds s;
  size_t o = 0;
  s.pw = (u64*)(base + o); o += size_t(mb) * npl * 8;
  s.emax = (short*)(base + o); o += size_t(mb) * 2;
  s.nbps = base + o; o += mb;
  s.nsave = base + o; o += mb;
  s.ssave = base + o;
  return s;
}

// Bricks of several (dataset, level) jobs per launch: table[j] describes job
// j, whose bricks are work items start[j] .. start[j + 1] - 1.
__global__ void __launch_bounds__(kThreads) k_decode_level(const DecArgs* __restrict__ table,
                                                           const int* __restrict__ start, int njobs) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int s_klo, s_khi, s_sig;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6, nwaves = blockDim.x >> 6;
  const int prec3[4] = {63, 62, 61, 60};
  const int total = start[njobs];
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int j = job_of(start, njobs, g);
    const DecArgs& a = table[j];
    const BrickConst& k = a.k;
    const int npl = k.dtype ? 64 : 32;
    DecLds S = dcarve(a.tab_off >= 0
                          ? reinterpret_cast<unsigned char*>(a.scratch + (size_t)blockIdx.x * a.wstride + a.tab_off)
                          : lds_raw,
                      npl, a.lds_blocks);
    const int bi = g - start[j];
    // the extended brick: in LDS when the launch's bricks fit there (small
    // bricks, 2D), else in the workgroup's scratch slot
    double* F = a.lds_brick >= 0 ? reinterpret_cast<double*>(lds_raw + a.lds_brick)
                                 : a.scratch + (size_t)blockIdx.x * a.wstride;
    const int b3[3] = {a.xyz[3 * bi], a.xyz[3 * bi + 1], a.xyz[3 * bi + 2]};
    const u64 lin = a.lin[bi];
    const int dense = a.dense[bi];
    for (int i = tid; i < k.nvox; i += blockDim.x) F[i] = 0;
    if (tid == 0) s_sig = 0;
    __syncthreads();
    for (int sb = 0; sb < k.nsb; ++sb) {
      if (!((a.mask >> sb) & 1)) continue;
      const DGrid& g = k.sb[sb];
      // parent's decoded interior into subband 0 (CopyExtentGrid, Volume.h:1153-1173)
      if (sb == 0 && !a.coarsest) {
        // the parent's decoded data at this brick's place in the level+1 grid
        const int n = a.sbne[0] * a.sbne[1] * a.sbne[2];
        for (int i = tid; i < n; i += blockDim.x) {
          int x = i % a.sbne[0], y = (i / a.sbne[0]) % a.sbne[1], z = i / (a.sbne[0] * a.sbne[1]);
          const size_t src = ((size_t)(b3[2] * a.sbne[2] + z) * a.pgdims[1] + (b3[1] * a.sbne[1] + y)) * a.pgdims[0] +
                             (b3[0] * a.sbne[0] + x);
          F[vox(k, g.f[0] + g.s[0] * x, g.f[1] + g.s[1] * y, g.f[2] + g.s[2] * z)] = a.pgrid[src];
        }
      }
      const int nblk = a.blk_n[sb];
      const u64* blocks = a.blocks64 + a.blk_off[sb];
      if (tid == 0) { s_klo = 0x7fffffff; s_khi = -0x7fffffff; }
      __syncthreads();
      // exponents of the coded blocks and the key range to walk
      for (int b = tid; b < nblk; b += blockDim.x) {
        int org[3], bd[3], skip, ci;
        unpack_block64(blocks[b], org, bd, &skip, &ci);
        S.nbps[b] = 0;
        S.nsave[b] = 0;
        S.ssave[b] = 0;
        for (int p = 0; p < npl; ++p) S.pw[b * npl + p] = 0;
        if (skip) { S.emax[b] = kSkip; continue; }
        const long long row = a.exp_off[sb] + ((long long)dense * a.ncoded[sb] + ci) * (k.dtype ? 2 : 1);
        const int e = k.dtype ? int((uint32_t)a.exps[row] | ((uint32_t)a.exps[row + 1] << 8)) - 1023
                              : int(a.exps[row]) - 127;
        S.emax[b] = (short)e;
        const int stop = max(64 - k.endbp, 57 + a.etol - e);
        if (stop <= 63) {
          atomicMax(&s_khi, 63 + e + kBpBiasDev);
          atomicMin(&s_klo, stop + e + kBpBiasDev);
        }
      }
      __syncthreads();
      const int klo = s_klo, khi = s_khi;
      // ---- pipelined stream walk (wave 0), key groups of 64 from the top ----
      if (wave == 0 && khi >= klo) {
        for (int g0 = khi; g0 >= klo; g0 -= 64) {
          const int key = g0 - lane;
          const bool kact = key >= klo;
          u64 pos = 0;
          bool have = false;
          if (kact && key >= a.dir_klo && key < a.dir_klo + a.dir_nk) {
            const u64 d = a.dir[((size_t)sb * a.dir_nk + (key - a.dir_klo)) * a.span + lin];
            if (d) { pos = dir_offset(d) * 8; have = true; }
          }
          int n_out = 0;
          bool stop_out = false;
          for (int t = 0; t < nblk + 64; ++t) {
            const int n_left = __shfl_up(n_out, 1, 64);
            const int s_left = __shfl_up((int)stop_out, 1, 64);
            const int b = t - lane;
            const bool inb = b >= 0 && b < nblk;
            const bool valid = kact && inb && S.emax[inb ? b : 0] != kSkip;
            if (valid) {
              int n_in = lane == 0 ? (g0 == khi ? 0 : S.nsave[b]) : n_left;
              bool st_in = lane == 0 ? (g0 == khi ? false : S.ssave[b] != 0) : (s_left != 0);
              const int e = S.emax[b];
              const int bp = key - kBpBiasDev - e;
              const int stop = max(64 - k.endbp, 57 + a.etol - e);
              n_out = n_in;
              stop_out = st_in;
              if (bp <= 63 && bp >= stop && !st_in) {
                if (!have) {
                  stop_out = true;  // missing chunk: the block stops decoding planes
                } else {
                  int org[3], bd[3], skip, ci;
                  unpack_block64(blocks[b], org, bd, &skip, &ci);
                  const int nv = 1 << (2 * ((bd[0] > 1) + (bd[1] > 1) + (bd[2] > 1)));
                  int n = n_in;
                  const u64 X = decode_plane(a.blob, pos, n, nv);
                  n_out = n;
                  S.pw[b * npl + (63 - bp)] = X;
                  S.nbps[b] = 1;  // a plane of the block was decoded
                }
              } else if (bp > 63) {
                n_out = 0;
              }
              if (lane == 63 || key == klo) {
                S.nsave[b] = (unsigned char)n_out;
                S.ssave[b] = stop_out ? 1 : 0;
              }
            }
          }
        }
      }
      __syncthreads();
      // ---- inverse block stage: one wave per block, one coefficient per lane ----
      for (int b = wave; b < nblk; b += nwaves) {
        const int e = S.emax[b];
        if (e == kSkip) continue;
        int org[3], bd[3], skip, ci;
        unpack_block64(blocks[b], org, bd, &skip, &ci);
        bool bypass = false;
        for (int d = 0; d < 3; ++d) bypass |= (org[d] & a.spmask[sb][d]) != 0;
        const int stop = max(64 - k.endbp, 57 + a.etol - e);
        const int np = stop > 63 ? 0 : 64 - stop;
        if (np == 0 || bypass) continue;
        // a significant block (DecodeSubband, idx2Decode.cpp:205-212): planes
        // decoded, on a detail subband or on the coarsest level
        if (lane == 0 && S.nbps[b] && (sb > 0 || a.coarsest)) s_sig = 1;
        const int nd = (bd[0] > 1) + (bd[1] > 1) + (bd[2] > 1);
        const int nv = 1 << (2 * nd);
        u64 u = 0;
        for (int p = 0; p < np; ++p) u |= ((S.pw[b * npl + p] >> lane) & 1ull) << (63 - p);
        // InverseShuffle (Zfp.h:580-605): coefficient j comes from position ip[j]
        int src = nd == 3 ? a.perm.ip3[lane] : nd == 2 ? a.perm.ip2[lane & 15] : lane;
        const u64 us = (u64)shfl_i64((i64)u, src);
        i64 q = (i64)((us ^ IDX2G_NBMASK) - IDX2G_NBMASK);
        q = zfp_inverse_lanes(q, lane, nd);
        // Dequantize (Function.h:139-149)
        const double scale = inv_pow2d(prec3[nd] - 1 - e);
        const double f = scale * i64_to_f64(q);
        if (lane < nv) {
          int x = lane % bd[0], y = (lane / bd[0]) % bd[1], z = lane / (bd[0] * bd[1]);
          F[vox(k, g.f[0] + g.s[0] * (org[0] + x), g.f[1] + g.s[1] * (org[1] + y), g.f[2] + g.s[2] * (org[2] + z))] = f;
        }
      }
      __syncthreads();
    }
    inverse_brick(k, F, a.winv, a.coarsest);
    if (a.pool_vol) {  // the whole extended brick into its pool slot
      double* dst = a.pool_vol + (size_t)bi * k.nvox;
      for (int i = tid; i < k.nvox; i += blockDim.x) dst[i] = F[i];
      if (tid == 0) a.pool_sig[bi] = (uint8_t)s_sig;
    }
    // data region into this level's grid (read by the next finer level)
    if (a.cgrid) {
      int ext[3];
      for (int d = 0; d < 3; ++d) ext[d] = max(0, min(a.brick[d], a.cgend[d] - b3[d] * a.brick[d]));
      const int n = ext[0] * ext[1] * ext[2];
      for (int i = tid; i < n; i += blockDim.x) {
        const int x = i % ext[0], y = (i / ext[0]) % ext[1], z = i / (ext[0] * ext[1]);
        a.cgrid[((size_t)(b3[2] * a.brick[2] + z) * a.cgdims[1] + (b3[1] * a.brick[1] + y)) * a.cgdims[0] +
                (b3[0] * a.brick[0] + x)] = F[vox(k, x, y, z)];
      }
    }
    // ---- copy into the output grid: Crop(OutGrid, BrickGrid) ----
    if (a.out) {
      int cf[3], cd[3], bf[3];
      bool any = true;
      for (int d = 0; d < 3; ++d) {
        bf[d] = b3[d] * a.B3[d];
        const int blast = bf[d] + (a.brick[d] - 1) * a.bstride;
        const int olast = a.og_f[d] + (a.og_d[d] - 1) * a.og_s[d];
        int frst = max(a.og_f[d], bf[d]);
        int last = min(olast, blast);
        frst = ((frst - a.og_f[d] + a.og_s[d] - 1) / a.og_s[d]) * a.og_s[d] + a.og_f[d];
        last = ((last - a.og_f[d]) / a.og_s[d]) * a.og_s[d] + a.og_f[d];
        cf[d] = frst;
        cd[d] = frst <= last ? (last - frst) / a.og_s[d] + 1 : 0;
        any = any && cd[d] > 0;
      }
      if (any) {
        const int n = cd[0] * cd[1] * cd[2];
        for (int i = tid; i < n; i += blockDim.x) {
          const int ix = i % cd[0], iy = (i / cd[0]) % cd[1], iz = i / (cd[0] * cd[1]);
          const int p[3] = {cf[0] + ix * a.og_s[0], cf[1] + iy * a.og_s[1], cf[2] + iz * a.og_s[2]};
          int s3[3], o3[3];
          for (int d = 0; d < 3; ++d) {
            s3[d] = (p[d] - bf[d]) / a.bstride;
            o3[d] = (p[d] - a.og_f[d]) / a.og_s[d];
          }
          const double v = F[vox(k, s3[0], s3[1], s3[2])];
          const size_t oi = ((size_t)o3[2] * a.og_d[1] + o3[1]) * (size_t)a.og_d[0] + o3[0];
          if (k.dtype) ((double*)a.out)[oi] = v;
          else ((float*)a.out)[oi] = (float)v;
        }
      }
    }
    __syncthreads();
  }
}

// Stream walks of full bricks (dec32.hpp), one workgroup per brick of shape
// WS (b32::Walk3D for 32^3 bricks, b32::Walk2D for 32x32 bricks); plane words
// and counts to the batch arrays a.xs / a.nbps.
template <class WS>
__global__ void __launch_bounds__(WS::th, WS::occ) k_walk(const DecArgs* __restrict__ table, const int* __restrict__ start,
                                                 int njobs) {
  __shared__ b32::WalkLds<WS> W;
  const int total = start[njobs];
  int cj = -1;
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int j = job_of(start, njobs, g);
    const DecArgs& a = table[j];
    if (j != cj) {
      __syncthreads();
      b32::walk_load_blocks<WS>(a, W);
      cj = j;
    }
    b32::walk_wg<WS>(a, W, g - start[j]);
    __syncthreads();
  }
}

// One subband of a brick: zero V, parent interior (subband 0), inverse block
// stage of the walked planes.
template <int DT>
__device__ __forceinline__ void decode_subband_b32(const DecArgs& a, int sb, const int (&b3)[3], b32::DecLds32& D,
                                                   const b32::Pos& p, const uint32_t* s_blocks, const u64* XS,
                                                   const double* PB = nullptr) {
  const BrickConst& k = a.k;
  const int dz = 17 - k.sb[sb].f[2];
  if (PB) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's parent rows are in LDS
  __syncthreads();  // V free (and every wave's parent rows in PB)
  if (!((a.mask >> sb) & 1)) {  // subband not decoded: zeros
    for (int i = p.tid; i < kVXY * dz; i += b32::kTh) D.v.V[i] = 0.0;
  } else {
    if (sb == 0 && !a.coarsest) {
      // parent's decoded data: subband-0 interior (CopyExtentGrid, Volume.h:1153-1173);
      // the blocks below write every other sample
      for (int i = p.tid; i < 16 * 16 * 16; i += b32::kTh) {
        const int x = i & 15, y = (i >> 4) & 15, z = i >> 8;
        D.v.V[(z * kVY + y) * kVX + x] =
            PB ? PB[i]
               : gm(a.pgrid)[((size_t)(b3[2] * 16 + z) * a.pgdims[1] + (b3[1] * 16 + y)) * a.pgdims[0] + (b3[0] * 16 + x)];
      }
    }
    b32::inverse_blocks<DT>(a, sb, kVX, kVY, D, p, s_blocks, XS);
  }
  __syncthreads();
}

// Full 32^3 bricks of a 3D dataset (dec32.hpp): one brick per iteration,
// 1024 threads, the brick on chip.  The z-odd subbands are decoded first and
// parked in the workgroup's scratch area, so that only the z-even half of the
// register columns is live while the z-even subbands are decoded.
template <int DT>
__global__ void __launch_bounds__(b32::kTh) k_decode_b32(const DecArgs* __restrict__ table,
                                                         const int* __restrict__ start, int njobs) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ uint32_t s_blocks[8 * 128];
  __shared__ double w8[8];
  __shared__ unsigned char s_perm[80];  // coefficient order: TDec positions (64), then 2D order (16)
  b32::Lds& S = *reinterpret_cast<b32::Lds*>(lds_raw);
  b32::DecLds32& D = *reinterpret_cast<b32::DecLds32*>(&S.u);
  constexpr int npl = DT ? 64 : 32;
  const int total = start[njobs];
  int cj = -1;
  IDX2G_PHASE_START();
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int j = job_of(start, njobs, g);
    const DecArgs& a = table[j];
    const BrickConst& k = a.k;
    if (j != cj) {  // block tables and inverse subband weights of the job
      __syncthreads();
      for (int i = threadIdx.x; i < 8 * 128; i += blockDim.x) {
        const int s = i >> 7, b = i & 127;
        s_blocks[i] = b < a.blk_n[s] ? a.blocks[a.blk_off[s] + b] : 0u;
      }
      if (threadIdx.x < 8) {
        const int s = threadIdx.x;
        const int idx = k.sb[s].f[0] + 2 * k.sb[s].f[1] + 4 * k.sb[s].f[2];
        w8[idx] = (s == 0 && !a.coarsest) ? 1.0 : a.winv[s];
      }
      if (threadIdx.x < 64) s_perm[threadIdx.x] = (unsigned char)TDec::pos(a.perm.p3[threadIdx.x]);
      if (threadIdx.x < 16) s_perm[64 + threadIdx.x] = a.perm.p2[threadIdx.x];
      cj = j;
      __syncthreads();
    }
    const int bi = g - start[j];
    double* Z1 = a.scratch + (size_t)blockIdx.x * a.wstride;
    const b32::Pos p = b32::pos();
    const int b3[3] = {a.xyz[3 * bi], a.xyz[3 * bi + 1], a.xyz[3 * bi + 2]};
    const int dense = a.dense[bi];
    const u64* XS = a.xs + (size_t)bi * 8 * b32::xs_region(npl);
    // the parent interior of subband 0 on its way into LDS while the z-odd
    // subbands decode (16-byte aligned rows only; otherwise read when used)
    double* PB = nullptr;
    if (!a.coarsest && (a.mask & 1) && ((uintptr_t)a.pgrid & 15) == 0 && (a.pgdims[0] & 1) == 0) {
      PB = reinterpret_cast<double*>(lds_raw + b32::dec_parent_off());
      b32::load_parent_lds(a.pgrid, a.pgdims, b3, PB, p);
    }
    // the job's coefficient orders (positions in the TDec layout), from the
    // kernel's own LDS: DecLds32 shares LDS with the transform
    if (p.tid < 64) D.p3[p.tid] = s_perm[p.tid];
    if (p.tid < 16) D.p2[p.tid] = s_perm[64 + p.tid];
    // exponents and decoded plane counts of the brick's blocks
    {
      const int sb = p.tid >> 7, b = p.tid & 127;
      D.E[sb][b] = (sb < k.nsb && b < a.blk_n[sb]) ? (short)b32::block_exponent(a, s_blocks, sb, b, dense) : kSkip;
      D.nbps[sb][b] = a.nbps[(size_t)bi * 1024 + p.tid];
    }
    __syncthreads();
    IDX2G_PHASE(16);
    // z-odd subbands straight into the parking area and the faces: no V
    // staging and no barrier between them
    for (int sb = 0; sb < 8; ++sb) {
      if (k.sb[sb].f[2] != 1) continue;
      const int hx = k.sb[sb].f[0], hy = k.sb[sb].f[1];
      double* Z1s = Z1 + (hx + 2 * hy) * 16 * 256;
      if (!((a.mask >> sb) & 1)) b32::park_zero(Z1s, S, p, hx, hy);
      else b32::inverse_blocks<DT, true>(a, sb, kVX, kVY, D, p, s_blocks, XS, Z1s, &S, hx, hy);
    }
    IDX2G_PHASE(17);
    __syncthreads();
    IDX2G_PHASE(18);
    double c[b32::kN];
    for (int sb = 0; sb < 8; ++sb) {
      if (k.sb[sb].f[2] != 0) continue;
      decode_subband_b32<DT>(a, sb, b3, D, p, s_blocks, XS, sb == 0 ? PB : nullptr);
      IDX2G_PHASE(19);
      b32::unstage_subband_even(c, S, D.v.V, p, k.sb[sb].f[0], k.sb[sb].f[1]);
      __syncthreads();
      IDX2G_PHASE(20);
    }
    b32::unpark_odd(c, Z1, p);
    b32::weigh(c, S, p, w8);
#ifdef IDX2G_PHASE_TIMERS
    __syncthreads();
#endif
    IDX2G_PHASE(21);
    b32::inverse(c, S, p);
    IDX2G_PHASE(22);
    if (a.out) {
      // Crop(OutGrid, BrickGrid) (idx2Decode.cpp:431-445), one element per (thread, z)
      int o[2];
      bool ok2 = true;
      const int xy[2] = {p.x, p.y};
      for (int d = 0; d < 2; ++d) {
        const int gp = b3[d] * a.B3[d] + xy[d] * a.bstride;
        const int olast = a.og_f[d] + (a.og_d[d] - 1) * a.og_s[d];
        ok2 = ok2 && gp >= a.og_f[d] && gp <= olast && (gp - a.og_f[d]) % a.og_s[d] == 0;
        o[d] = (gp - a.og_f[d]) / a.og_s[d];
      }
      if (ok2) {
        const int olast = a.og_f[2] + (a.og_d[2] - 1) * a.og_s[2];
        const size_t row = (size_t)o[1] * a.og_d[0] + o[0];
        const size_t plane = (size_t)a.og_d[1] * a.og_d[0];
        const int gp0 = b3[2] * a.B3[2], s2 = a.og_s[2], bs = a.bstride;
        if (bs % s2 == 0) {
          // the sample stride is a multiple of the output stride: the
          // divisibility test is the same for every z (no division per z)
          if ((gp0 - a.og_f[2]) % s2 == 0) {
            const long long o0 = (gp0 - a.og_f[2]) / s2, st = bs / s2;
#pragma unroll
            for (int z = 0; z < 32; ++z) {
              const int gp = gp0 + z * bs;
              if (gp >= a.og_f[2] && gp <= olast) {
                const size_t oi = (size_t)(o0 + z * st) * plane + row;
                if (DT) gm((double*)a.out)[oi] = c[z];
                else gm((float*)a.out)[oi] = (float)c[z];
              }
            }
          }
        } else {
#pragma unroll
          for (int z = 0; z < 32; ++z) {
            const int gp = gp0 + z * bs;
            if (gp >= a.og_f[2] && gp <= olast && (gp - a.og_f[2]) % s2 == 0) {
              const size_t oi = (size_t)((gp - a.og_f[2]) / s2) * plane + row;
              if (DT) gm((double*)a.out)[oi] = c[z];
              else gm((float*)a.out)[oi] = (float)c[z];
            }
          }
        }
      }
    } else {
      const size_t plane = (size_t)a.cgdims[1] * a.cgdims[0];
      const auto dst = gm(a.cgrid) + ((size_t)(b3[2] * 32) * a.cgdims[1] + (b3[1] * 32 + p.y)) * a.cgdims[0] + (b3[0] * 32 + p.x);
#pragma unroll
      for (int z = 0; z < 32; ++z) dst[z * plane] = c[z];
    }
    __syncthreads();
    IDX2G_PHASE(23);
  }
}

// Full 32x32 bricks of a 2D dataset (dec2d.hpp): 32 bricks per workgroup
// iteration (brick slot = half-wave), start[] in iterations; the walk
// (k_walk) has left the plane words in a.xs / a.nbps by piece-local
// brick index.
template <int DT>
__global__ void __launch_bounds__(b2d::kTh) k_decode_2d(const DecArgs* __restrict__ table,
                                                        const int* __restrict__ start, int njobs) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ uint32_t s_blocks[4 * 32];
  __shared__ double w4[4];
  __shared__ int s_b3[b2d::kSlots][2];
  __shared__ int s_poff[2];  // parking areas of the hy = 1 subbands by hx
  b2d::DecLds2& D = *reinterpret_cast<b2d::DecLds2*>(lds_raw);
  const int total = start[njobs];
  int cj = -1;
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int j = job_of(start, njobs, g);
    const DecArgs& a = table[j];
    const BrickConst& k = a.k;
    if (j != cj) {  // block tables and inverse weights of the job's subbands
      __syncthreads();
      for (int i = threadIdx.x; i < 4 * 32; i += blockDim.x) {
        const int sb = i >> 5, b = i & 31;
        s_blocks[i] = b < a.blk_n[sb] ? a.blocks[a.blk_off[sb] + b] : 0u;
      }
      if (threadIdx.x < 4) {
        const int sb = threadIdx.x;
        w4[k.sb[sb].f[0] + 2 * k.sb[sb].f[1]] = (sb == 0 && !a.coarsest) ? 1.0 : a.winv[sb];
      }
      if (threadIdx.x < 2) s_poff[threadIdx.x] = threadIdx.x ? 17 * 16 * b2d::kSlots : 0;
      cj = j;
      __syncthreads();
    }
    const int bi0 = (g - start[j]) * a.per;
    const int nvalid = min(a.per, a.nb - bi0);
    const b2d::Pos2 p = b2d::pos2();
    if (p.x == 0) {
      const int bi = bi0 + (p.slot < nvalid ? p.slot : 0);
      s_b3[p.slot][0] = a.xyz[3 * bi];
      s_b3[p.slot][1] = a.xyz[3 * bi + 1];
    }
    for (int t = p.tid; t < b2d::kSlots * 4 * 32; t += b2d::kTh) {
      // exponents (EncodeBrickSubbandExponents layout) and decoded plane counts
      const int s = t >> 7, sb = (t >> 5) & 3, b = t & 31;
      short e = kSkip;
      unsigned char nb = 0;
      if (s < nvalid && b < a.blk_n[sb]) {
        const uint32_t bw = s_blocks[sb * 32 + b];
        if (!((bw >> 24) & 1)) {
          const long long row = a.exp_off[sb] + ((long long)a.dense[bi0 + s] * a.ncoded[sb] + int(bw >> 25)) * (DT ? 2 : 1);
          e = DT ? (short)(int((uint32_t)a.exps[row] | ((uint32_t)a.exps[row + 1] << 8)) - 1023)
                 : (short)(int(a.exps[row]) - 127);
        }
        nb = a.nbps[(size_t)(bi0 + s) * 1024 + sb * 128 + b];
      }
      D.E[s][sb][b] = e;
      D.nbps[s][sb][b] = nb;
    }
    if (p.tid < 64) D.p3[p.tid] = (unsigned char)TDec::pos(a.perm.p3[p.tid]);  // positions in the TDec layout
    if (p.tid < 16) D.p2[p.tid] = a.perm.p2[p.tid];
    for (int t = p.tid; t < b2d::kSlots * b32::kN; t += b2d::kTh) D.ex[t] = 0.0;
    b32::Pos pb;
    pb.tid = p.tid;
    pb.lane = p.lane;
    pb.w = p.w;
    pb.x = p.x;
    pb.y = p.slot;
    pb.orow = p.lane >= 32;
    __syncthreads();
    // One subband of the group's bricks into D.V: zeros, the parent's
    // subband-0 interior (CopyExtentGrid, Volume.h:1153-1173), the blocks.
    auto stage = [&](int sb, int s0, int G) {
      const int hx = k.sb[sb].f[0], hy = k.sb[sb].f[1];
      const int dx = 17 - hx, dy = 17 - hy;
      if (!((a.mask >> sb) & 1)) {  // subband not decoded: zeros
        for (int t = p.tid; t < G * dx * dy; t += b2d::kTh) D.V[t] = 0.0;
        __syncthreads();
        return;
      }
      if (sb == 0 && !a.coarsest) {  // the blocks below write every other sample
        for (int t = p.tid; t < G * 256; t += b2d::kTh) {
          const int gg = t >> 8, x = t & 15, y = (t >> 4) & 15;
          D.V[(gg * dy + y) * dx + x] =
              gm(a.pgrid)[(size_t)(s_b3[s0 + gg][1] * 16 + y) * a.pgdims[0] + (s_b3[s0 + gg][0] * 16 + x)];
        }
      }
      b2d::inverse_blocks_2d<DT>(a, sb, s0, G, dx, dy, D, pb, s_blocks + sb * 32, a.blk_n[sb], bi0, D.V);
      __syncthreads();
    };
    // odd rows (hy = 1) first, parked in the workgroup's scratch slot, so
    // that only the even rows of the register columns are live meanwhile
    double* P = a.scratch + (size_t)blockIdx.x * a.wstride;
    for (int sb = 0; sb < 4; ++sb) {
      const int hx = k.sb[sb].f[0];
      if (k.sb[sb].f[1] != 1) continue;
      for (int s0 = 0; s0 < nvalid; s0 += b2d::kGroup) {
        const int G = min(b2d::kGroup, nvalid - s0);
        stage(sb, s0, G);
        b2d::park_group(D, P + s_poff[hx], p, 17 - hx, 16, s0, G);
        __syncthreads();
      }
    }
    double c[b32::kN];
#pragma unroll
    for (int y = 0; y < b32::kN; y += 2) c[y] = 0.0;
    for (int sb = 0; sb < 4; ++sb) {
      const int hx = k.sb[sb].f[0];
      if (k.sb[sb].f[1] != 0) continue;
      for (int s0 = 0; s0 < nvalid; s0 += b2d::kGroup) {
        const int G = min(b2d::kGroup, nvalid - s0);
        stage(sb, s0, G);
        b2d::take_subband<0>(c, D, p, hx, s0, G);
        __syncthreads();
      }
    }
    __threadfence_block();
    b2d::unpark_odd(c, D.ex, P, s_poff, p);
    __syncthreads();
    {  // inverse subband weights by parity (x = 32 is even)
      const double we = w4[p.x & 1], wo = w4[(p.x & 1) + 2];
#pragma unroll
      for (int y = 0; y < b32::kN; ++y) c[y] = c[y] * ((y & 1) ? wo : we);
      for (int t = p.tid; t < b2d::kSlots * b32::kN; t += b2d::kTh) {
        const int y = t % b32::kN;
        D.ex[t] = D.ex[t] * w4[2 * (y & 1)];
      }
    }
    __syncthreads();
    b2d::inverse(c, D.ex, p);
    if (p.slot < nvalid) {
      const int bx = s_b3[p.slot][0], by = s_b3[p.slot][1];
      if (a.out) {
        // Crop(OutGrid, BrickGrid) (idx2Decode.cpp:431-445); z = 0 on a 2D grid
        const int gx = bx * a.B3[0] + p.x * a.bstride;
        const int xlast = a.og_f[0] + (a.og_d[0] - 1) * a.og_s[0];
        const bool okx = gx >= a.og_f[0] && gx <= xlast && (gx - a.og_f[0]) % a.og_s[0] == 0;
        if (okx) {
          const int ox = (gx - a.og_f[0]) / a.og_s[0];
          const int ylast = a.og_f[1] + (a.og_d[1] - 1) * a.og_s[1];
#pragma unroll
          for (int y = 0; y < 32; ++y) {
            const int gy = by * a.B3[1] + y * a.bstride;
            if (gy >= a.og_f[1] && gy <= ylast && (gy - a.og_f[1]) % a.og_s[1] == 0) {
              const size_t oi = (size_t)((gy - a.og_f[1]) / a.og_s[1]) * a.og_d[0] + ox;
              if (DT) gm((double*)a.out)[oi] = c[y];
              else gm((float*)a.out)[oi] = (float)c[y];
            }
          }
        }
      } else {
        const auto dst = gm(a.cgrid) + (size_t)(by * 32) * a.cgdims[0] + (bx * 32 + p.x);
#pragma unroll
        for (int y = 0; y < 32; ++y) dst[(size_t)y * a.cgdims[0]] = c[y];
      }
    }
    __syncthreads();
  }
}

// Brick lists of the query boxes of one decode round, one thread per box
// position (work item g of job j: box position g - start[j] in x-fastest
// order).  A full brick goes to rank "full positions before it"; any other
// to nfull + "other positions before it"; both counts are closed forms of
// the position because the full bricks form the box's corner sub-box.
__global__ void k_plan_bricks(const PlanArgs* __restrict__ jobs, const int* __restrict__ start, int njobs) {
  const int total = start[njobs];
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    const int j = job_of(start, njobs, g);
    const PlanArgs& P = jobs[j];
    const int p = g - start[j];
    const int x = p % P.nbox[0], y = (p / P.nbox[0]) % P.nbox[1], z = p / (P.nbox[0] * P.nbox[1]);
    const int fx = P.nfull3[0], fy = P.nfull3[1], fz = P.nfull3[2];
    const int before = min(z, fz) * fx * fy + (z < fz ? min(y, fy) * fx + (y < fy ? min(x, fx) : 0) : 0);
    const bool full = x < fx && y < fy && z < fz;
    const int slot = full ? before : P.nfull + (p - before);
    const int b3[3] = {P.bf[0] + x, P.bf[1] + y, P.bf[2] + z};
    // linear id: the order's last position takes the lowest bit
    int c[3] = {b3[0], b3[1], b3[2]};
    u64 lin = 0;
    for (int i = P.nbits - 1; i >= 0; --i) {
      const int d = P.axis[i];
      lin |= u64(c[d] & 1) << (P.nbits - 1 - i);
      c[d] >>= 1;
    }
    int lo = 0, hi = P.nb_level - 1;  // rank among the sorted ids (every box brick exists)
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (P.sorted_lin[mid] < lin) lo = mid + 1;
      else hi = mid;
    }
    P.xyz[3 * slot] = b3[0];
    P.xyz[3 * slot + 1] = b3[1];
    P.xyz[3 * slot + 2] = b3[2];
    P.lin[slot] = lin;
    P.dense[slot] = lo;
  }
}

// Sparse-brick pool: brick b of dst is slot slots[b] of src (nvox doubles each).
__global__ void k_gather_bricks(const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ slots,
                                int nb, int nvox) {
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const double* s = src + (size_t)slots[b] * nvox;
    double* d = dst + (size_t)b * nvox;
    for (int i = threadIdx.x; i < nvox; i += blockDim.x) d[i] = s[i];
  }
}

// The pool's device hash table: open addressing with linear probing over a
// power-of-two table at most half full; a slot holds key + 1 (0: empty) and
// a value (the brick's device address).
__device__ __forceinline__ uint32_t pool_slot(u64 key, uint32_t mask) {
  u64 x = key * 0x9E3779B97F4A7C15ull;
  x ^= x >> 29;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 32;
  return (uint32_t)x & mask;
}

__global__ void k_pool_insert(const u64* __restrict__ keys, const u64* __restrict__ vals, int n, u64* tkeys, u64* tvals,
                              uint32_t mask) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const u64 k1 = keys[i] + 1;
    uint32_t h = pool_slot(keys[i], mask);
    for (uint32_t probe = 0; probe <= mask; ++probe) {  // the table has empty slots: ends within mask + 1 probes
      const u64 prev = atomicCAS((unsigned long long*)&tkeys[h], 0ull, (unsigned long long)k1);
      if (prev == 0 || prev == k1) {
        tvals[h] = vals[i];
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

// out[i] = value of q[i], 0 when the key is not in the table.
__global__ void k_pool_lookup(const u64* __restrict__ tkeys, const u64* __restrict__ tvals, uint32_t mask,
                              const u64* __restrict__ q, long long n, u64* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const u64 k1 = q[i] + 1;
    uint32_t h = pool_slot(q[i], mask);
    u64 v = 0;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
      const u64 t = tkeys[h];
      if (t == k1) {
        v = tvals[h];
        break;
      }
      if (t == 0) break;
      h = (h + 1) & mask;
    }
    out[i] = v;
  }
}

}  // namespace idx2g
