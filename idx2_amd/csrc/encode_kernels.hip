// Encode kernels for gfx950: one workgroup per brick runs the whole per-brick
// path of EncodeBrick (Source/Core/idx2Encode.cpp:276-340): gather,
// ExtrapolateCdf53, one-level CDF5/3 with subband weights, subband-0 copy to
// the parent, then per subband the zfp block stage and the embedded coder
// (idx2Encode.cpp:141-272).  Streams of one brick-subband are assembled in
// LDS and appended to a device arena; chunk assembly (idx2Write.cpp:207-290)
// runs in the kernels at the end of this file.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "brick32.hpp"
#include "code32.hpp"
#include "code2d.hpp"

namespace idx2g {

struct EncShared {
  int klo, khi, err;
  double wmn[16], wmx[16];
  unsigned long long base, ebase;
  uint32_t total;
};

// Working tables for one subband: npl planes per block, room for mb blocks
// (the launch's largest subband).  The key tables and the stream staging
// live in LDS; the per-block tables follow them in LDS when they fit (bricks
// of at most 32 samples per axis: several workgroups per CU for small
// bricks), else in the workgroup's scratch slot past the extended brick.
struct EncLds {
  u64* pw;          // [mb][npl] plane words
  uint32_t* poff;   // [mb][npl] bit offset inside the key stream
  uint16_t* plen;   // [mb][npl] emitted bits
  short* emax;      // [mb]
  unsigned char* np;// [mb]
  uint32_t* kbits;  // [kMaxKeys]
  uint32_t* kboff;  // [kMaxKeys] byte offset of the key stream in the subband
  short* kfirst;    // [kMaxKeys]
  uint32_t* stage;  // [kStageWords]
};

__host__ __device__ inline size_t enc_small_bytes() { return size_t(kStageWords) * 4 + kMaxKeys * (4 + 4 + 2) + 16; }
__host__ __device__ inline size_t enc_tab_bytes(int npl, int mb) {
  return size_t(mb) * npl * (8 + 4 + 2) + size_t(mb) * 3 + 64;
}
// both parts in LDS
__host__ __device__ inline size_t enc_lds_bytes(int npl, int mb) { return enc_small_bytes() + enc_tab_bytes(npl, mb); }

// tab: the per-block tables' base, or nullptr for LDS after the small part
__device__ inline EncLds carve(unsigned char* lds, unsigned char* tab, int npl, int mb) {
  EncLds s;
  size_t o = 0;
  s.stage = (uint32_t*)(lds + o); o += size_t(kStageWords) * 4;
  s.kbits = (uint32_t*)(lds + o); o += kMaxKeys * 4;
  s.kboff = (uint32_t*)(lds + o); o += kMaxKeys * 4;
  s.kfirst = (short*)(lds + o); o += kMaxKeys * 2;
  unsigned char* t = tab ? tab : lds + ((o + 15) & ~size_t(15));
  o = 0;
  s.pw = (u64*)(t + o); o += size_t(mb) * npl * 8;
  s.poff = (uint32_t*)(t + o); o += size_t(mb) * npl * 4;
  s.plen = (uint16_t*)(t + o); o += size_t(mb) * npl * 2;
  s.emax = (short*)(t + o); o += size_t(mb) * 2;
  s.np = (unsigned char*)(t + o);
  return s;
}


__device__ inline void flag_error(unsigned long long* counters, unsigned long long bit) {
  atomicOr(&counters[2], bit);
}

// OR `len` bits of w (LSB first) at bit position pos of a 32-bit word array
template <bool kShared>
__device__ inline void or_bits(uint32_t* words, u64 pos, const u64 w[3], int len) {
  if (len <= 0) return;
  const int sh = int(pos & 31);
  u64 wi = pos >> 5;
  // shifted view of w: 32-bit pieces
  const int nw = (sh + len + 31) >> 5;
  for (int i = 0; i < nw; ++i) {
    // bits [32*i - sh, 32*i - sh + 32) of w
    int b0 = 32 * i - sh;
    u64 v;
    if (b0 >= 0) {
      int q = b0 >> 6, r = b0 & 63;
      v = w[q] >> r;
      if (r > 32 && q + 1 < 3) v |= w[q + 1] << (64 - r);
    } else {
      v = w[0] << (-b0);
    }
    uint32_t piece = uint32_t(v);
    if (piece) {
      if (kShared) atomicOr(&words[wi + i], piece);
      else atomicOr(&words[wi + i], piece);
    }
  }
}

// Per-subband zfp block stage and bit-plane coding of one transformed brick F
// (EncodeSubbandBlocks, Source/Core/idx2Encode.cpp:141-256): stage A one wave
// per block (quantise, lift, shuffle, one ballot per plane), B emitted
// lengths, C per-key bit offsets in Morton order, D arena reservation,
// E stream bits (LDS staging).
__device__ void code_subbands(const EncArgs& a, const double* F, int dense, unsigned char* lds_raw,
                              EncShared& sh) {
  const BrickConst& k = a.k;
  const int npl = k.dtype ? 64 : 32;
  EncLds S = carve(lds_raw,
                   a.tab_off >= 0 ? reinterpret_cast<unsigned char*>(a.scratch + (size_t)blockIdx.x * a.wstride + a.tab_off)
                                  : nullptr,
                   npl, a.lds_blocks);
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6, nwaves = blockDim.x >> 6;
  const int prec3[4] = {63, 62, 61, 60};
  for (int sb = 0; sb < k.nsb; ++sb) {
    const DGrid& g = k.sb[sb];
    const int nblk = a.blk_n[sb];
    const u64* blocks = a.blocks64 + a.blk_off[sb];
    if (tid == 0) { sh.klo = 0x7fffffff; sh.khi = -0x7fffffff; sh.err = 0; }
    __syncthreads();
    // ---- stage A: one wave per block, one coefficient per lane ----
    for (int b = wave; b < nblk; b += nwaves) {
      int org[3], bd[3], skip, ci;
      unpack_block64(blocks[b], org, bd, &skip, &ci);
      if (skip) {
        if (lane == 0) { S.emax[b] = kSkip; S.np[b] = 0; }
        continue;
      }
      const int nd = (bd[0] > 1) + (bd[1] > 1) + (bd[2] > 1);
      const int nv = 1 << (2 * nd);
      const bool act = lane < nv;
      // gather order: X fastest over the block dims (idx2_BeginFor3)
      int x = lane % bd[0], y = (lane / bd[0]) % bd[1], z = lane / (bd[0] * bd[1]);
      double v = 0;
      if (act)
        v = F[vox(k, g.f[0] + g.s[0] * (org[0] + x), g.f[1] + g.s[1] * (org[1] + y),
                  g.f[2] + g.s[2] * (org[2] + z))];
      // QuantizeF32/F64 (Source/Core/Function.h:58-89)
      const double maxabs = wave_max(act ? fabs(v) : 0.0);
      const int emax = k.dtype ? exponent_f64(maxabs) : exponent_f32((float)maxabs);
      const int prec = prec3[nd];
      const double scale = pow2d(prec - 1 - emax);
      i64 q = act ? f64_to_i64(scale * v) : 0;
      q = zfp_forward_lanes(q, lane, nd);
      // ForwardShuffle + negabinary (Zfp.h:534-559)
      int src = nd == 3 ? a.perm.p3[lane] : nd == 2 ? a.perm.p2[lane & 15] : lane;
      i64 qq = shfl_i64(q, src);
      const u64 u = ((u64)qq + IDX2G_NBMASK) ^ IDX2G_NBMASK;
      int stop = max(64 - k.endbp, 57 + a.etol - emax);
      int np = stop > 63 ? 0 : 64 - stop;
      for (int p = 0; p < np; ++p) {
        u64 X = __ballot(act && ((u >> (63 - p)) & 1ull));
        if (lane == 0) S.pw[b * npl + p] = X;
      }
      if (lane == 0) {
        S.emax[b] = (short)emax;
        S.np[b] = (unsigned char)np;
        if (np > 0) {
          atomicMax(&sh.khi, 63 + emax + kBpBiasDev);
          atomicMin(&sh.klo, stop + emax + kBpBiasDev);
        }
        // exponent sub-channel (EncodeBrickSubbandExponents, idx2Encode.cpp:41-80)
        const long long row = a.exp_off[sb] + ((long long)dense * a.ncoded[sb] + ci) * (k.dtype ? 2 : 1);
        if (k.dtype) {
          uint16_t e = (uint16_t)(emax + 1023);
          a.exps[row] = (uint8_t)(e & 0xff);
          a.exps[row + 1] = (uint8_t)(e >> 8);
        } else {
          a.exps[row] = (uint8_t)(emax + 127);
        }
      }
    }
    __syncthreads();
    // ---- stage B: emitted length of every (block, plane), thread per block ----
    for (int b = tid; b < nblk; b += blockDim.x) {
      if (S.emax[b] == kSkip) continue;
      int org[3], bd[3], skip, ci;
      unpack_block64(blocks[b], org, bd, &skip, &ci);
      const int nv = 1 << (2 * ((bd[0] > 1) + (bd[1] > 1) + (bd[2] > 1)));
      int n = 0;
      for (int p = 0; p < S.np[b]; ++p) S.plen[b * npl + p] = (uint16_t)plane_len(S.pw[b * npl + p], n, nv);
    }
    __syncthreads();
    const int klo = sh.klo, khi = sh.khi;
    const int nk = khi >= klo ? khi - klo + 1 : 0;
    if (nk > kMaxKeys) {
      if (tid == 0) { flag_error(a.counters, 1); a.hdrs[(size_t)dense * k.nsb + sb] = SbHdr{0, 0, 0}; }
      __syncthreads();
      continue;
    }
    // ---- stage C: per key, bit offsets of each block's plane (Morton order) ----
    for (int kk = tid; kk < nk; kk += blockDim.x) {
      const int key = klo + kk;
      uint32_t bits = 0;
      int first = -1;
      for (int b = 0; b < nblk; ++b) {
        const int e = S.emax[b];
        if (e == kSkip) continue;
        const int p = 63 - (key - kBpBiasDev - e);
        if (p >= 0 && p < S.np[b]) {
          S.poff[b * npl + p] = bits;
          bits += S.plen[b * npl + p];
          if (first < 0) first = b;
        }
      }
      S.kbits[kk] = bits;
      S.kfirst[kk] = (short)first;
    }
    __syncthreads();
    // ---- stage D: byte layout of the key streams, arena + entry reservation ----
    if (tid == 0) {
      uint32_t tot = 0;
      for (int kk = 0; kk < nk; ++kk) {
        S.kboff[kk] = tot;
        tot += (S.kbits[kk] + 7) >> 3;
      }
      sh.total = tot;
      const unsigned long long alloc = (tot + 3ull) & ~3ull;
      sh.base = alloc ? atomicAdd(&a.counters[0], alloc) : 0;
      if (tot) atomicAdd(&a.counters[3], (unsigned long long)tot);
      sh.ebase = nk ? atomicAdd(&a.counters[1], (unsigned long long)nk) : 0;
      const bool over = sh.base + alloc > a.arena_cap || sh.ebase + nk > a.entry_cap;
      if (over) flag_error(a.counters, 2);
      SbHdr h;
      h.klo = klo;
      h.nkeys = over ? 0 : nk;
      h.ebase = sh.ebase;
      a.hdrs[(size_t)dense * k.nsb + sb] = h;
    }
    __syncthreads();
    const unsigned long long base = sh.base;
    const uint32_t total = sh.total;
    if (base + ((total + 3ull) & ~3ull) > a.arena_cap || sh.ebase + nk > a.entry_cap) continue;
    for (int kk = tid; kk < nk; kk += blockDim.x) {
      SbEntry e;
      e.off = base + S.kboff[kk];
      e.size = (S.kbits[kk] + 7) >> 3;
      e.first_block = (uint16_t)(S.kfirst[kk] < 0 ? 0xffff : S.kfirst[kk]);
      e.key = (uint16_t)(klo + kk);
      a.entries[sh.ebase + kk] = e;
    }
    // ---- stage E: write the streams (LDS staging, else global) ----
    const uint32_t nwords = (total + 3) >> 2;
    const bool staged = nwords <= (uint32_t)kStageWords;
    uint32_t* gwords = (uint32_t*)(a.arena + base);
    if (staged) {
      for (uint32_t i = tid; i < nwords; i += blockDim.x) S.stage[i] = 0;
    } else {
      for (uint32_t i = tid; i < nwords; i += blockDim.x) gwords[i] = 0;
      __threadfence();
    }
    __syncthreads();
    for (int b = tid; b < nblk; b += blockDim.x) {
      const int e = S.emax[b];
      if (e == kSkip) continue;
      int org[3], bd[3], skip, ci;
      unpack_block64(blocks[b], org, bd, &skip, &ci);
      const int nv = 1 << (2 * ((bd[0] > 1) + (bd[1] > 1) + (bd[2] > 1)));
      int n = 0;
      for (int p = 0; p < S.np[b]; ++p) {
        const u64 X = S.pw[b * npl + p];
        u64 w[3];
        const int len = plane_emit(X, n, nv, w);
        plane_len(X, n, nv);  // advance n
        const int kk = (63 - p) + e + kBpBiasDev - klo;
        const u64 pos = (u64)S.kboff[kk] * 8 + S.poff[b * npl + p];
        if (staged) or_bits<true>(S.stage, pos, w, len);
        else or_bits<false>(gwords, pos, w, len);
      }
    }
    __syncthreads();
    if (staged)
      for (uint32_t i = tid; i < nwords; i += blockDim.x) gwords[i] = S.stage[i];
    __syncthreads();
  }
}

// Gather a brick from a level grid (the input at level 0, the dense f64
// grid of subband-0 interiors above) into a zeroed extended volume
// (brick_copier::Copy, Source/Core/idx2Encode.cpp:377-386, Volume.h:1123-1149).
__device__ inline void brick_origin(const EncArgs& a, int dense, int b3[3], int from[3], int ext[3]) {
  for (int d = 0; d < 3; ++d) {
    b3[d] = a.xyz[3 * dense + d];
    from[d] = b3[d] * a.brick[d];
    ext[d] = max(0, min(a.brick[d], a.gdims[d] - from[d]));
  }
}

__device__ inline double load_sample(const EncArgs& a, size_t i) {
  return a.in_dtype ? ((const double*)a.input)[i] : (double)((const float*)a.input)[i];
}

__device__ inline void store_minmax(const EncArgs& a, EncShared& sh, int dense, double mn, double mx) {
  const int lane = lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  mn = -wave_max(-mn);
  mx = wave_max(mx);
  if (lane == 0) { sh.wmn[wave] = mn; sh.wmx[wave] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nwaves; ++w) {
      mn = sh.wmn[w] < mn ? sh.wmn[w] : mn;
      mx = sh.wmx[w] > mx ? sh.wmx[w] : mx;
    }
    a.minmax[2 * dense] = mn;
    a.minmax[2 * dense + 1] = mx;
  }
}

// Any brick (partial bricks at the volume's far faces, any brick size up to
// 32): the extended brick lives in a global scratch volume per workgroup.
// Bricks of several (dataset, level) jobs in one launch: table[j] describes
// job j, whose bricks are work items start[j] .. start[j + 1] - 1.
__global__ void __launch_bounds__(kThreads) k_encode_level(const EncArgs* __restrict__ table,
                                                           const int* __restrict__ start, int njobs) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ EncShared sh;
  const int tid = threadIdx.x;
  const int total = start[njobs];
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int j = job_of(start, njobs, g);
    const EncArgs& a = table[j];
    const BrickConst& k = a.k;
    const int bi = g - start[j];
    const int dense = a.list ? a.list[bi] : a.b0 + bi;
    int b3[3], from[3], ext[3];
    brick_origin(a, dense, b3, from, ext);
    // the extended brick: in LDS when the launch's bricks fit there (small
    // bricks, 2D), else in the workgroup's scratch slot
    double* F = a.lds_brick >= 0 ? reinterpret_cast<double*>(lds_raw + a.lds_brick)
                                 : a.scratch + (size_t)blockIdx.x * a.wstride;
    double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
    for (int i = tid; i < k.nvox; i += blockDim.x) {
      int x = i % k.N[0], y = (i / k.N[0]) % k.N[1], z = i / (k.N[0] * k.N[1]);
      double v = 0;
      if (x < ext[0] && y < ext[1] && z < ext[2]) {
        v = load_sample(a, ((size_t)(from[2] + z) * a.gdims[1] + (from[1] + y)) * (size_t)a.gdims[0] + (from[0] + x));
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
      }
      F[i] = v;
    }
    if (a.level == 0) store_minmax(a, sh, dense, mn, mx);
    __syncthreads();
    extrapolate(k, F, ext);
    forward_brick(k, F, a.wfwd, a.coarsest);
    // subband 0 (non-extended interior) into the parent level's grid
    // (EncodeBrick, idx2Encode.cpp:303-333; CopyGridExtent, Volume.h:1177-1196)
    if (!a.coarsest) {
      const DGrid& g = k.sb[0];
      const int n = a.sbne[0] * a.sbne[1] * a.sbne[2];
      for (int i = tid; i < n; i += blockDim.x) {
        int x = i % a.sbne[0], y = (i / a.sbne[0]) % a.sbne[1], z = i / (a.sbne[0] * a.sbne[1]);
        const size_t dst = ((size_t)(b3[2] * a.sbne[2] + z) * a.pdims[1] + (b3[1] * a.sbne[1] + y)) * a.pdims[0] +
                           (b3[0] * a.sbne[0] + x);
        a.parents[dst] = F[vox(k, g.f[0] + g.s[0] * x, g.f[1] + g.s[1] * y, g.f[2] + g.s[2] * z)];
      }
    }
    code_subbands(a, F, dense, lds_raw, sh);
    __syncthreads();
  }
}

// Full 32^3 bricks of a 3D grid: the transform runs on chip (brick32.hpp),
// the transformed brick is handed to the coding stages through the
// workgroup's scratch volume.
template <int DT>
__global__ void __launch_bounds__(b32::kTh) k_encode_b32(const EncArgs* __restrict__ table,
                                                         const int* __restrict__ start, int njobs) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ EncShared sh;
  __shared__ double w8[8];
  __shared__ uint32_t s_blocks[8 * 128];
  __shared__ unsigned char s_perm[80];  // coefficient order: TEnc positions (64), then 2D order (16)
  __shared__ b32::ArenaCursor cur;
  b32::Lds& S = *reinterpret_cast<b32::Lds*>(lds_raw);
  if (threadIdx.x == 0) cur = b32::ArenaCursor{0, 0, 0, 0, 0};
  const int total = start[njobs];
  int cj = -1;
  IDX2G_PHASE_START();
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int j = job_of(start, njobs, g);
    const EncArgs& a = table[j];
    const BrickConst& k = a.k;
    if (j != cj) {  // block tables and subband weights of the job
      __syncthreads();
      for (int i = threadIdx.x; i < 8 * 128; i += blockDim.x) {
        const int s = i >> 7, b = i & 127;
        s_blocks[i] = b < a.blk_n[s] ? a.blocks[a.blk_off[s] + b] : 0u;
      }
      if (threadIdx.x < 8) {
        const int s = threadIdx.x;
        const int idx = k.sb[s].f[0] + 2 * k.sb[s].f[1] + 4 * k.sb[s].f[2];
        w8[idx] = (s == 0 && !a.coarsest) ? 1.0 : a.wfwd[s];
      }
      if (threadIdx.x < 64) s_perm[threadIdx.x] = (unsigned char)TEnc::pos(a.perm.p3[threadIdx.x]);
      if (threadIdx.x < 16) s_perm[64 + threadIdx.x] = a.perm.p2[threadIdx.x];
      cj = j;
      __syncthreads();
    }
    const unsigned plane = (unsigned)a.gdims[1] * (unsigned)a.gdims[0];
    const int bi = g - start[j];
    const int dense = a.list ? a.list[bi] : a.b0 + bi;
    int b3[3], from[3], ext[3];
    brick_origin(a, dense, b3, from, ext);
    const b32::Pos p = b32::pos();
    double c[b32::kN];
    const unsigned lo = (unsigned)p.y * (unsigned)a.gdims[0] + (unsigned)p.x;
    // uniform base of the brick + 32-bit lane/z offsets
    const size_t base = ((size_t)from[2] * a.gdims[1] + from[1]) * (size_t)a.gdims[0] + from[0];
    if (a.in_dtype) {
      const auto src = gm((const double*)a.input) + base;
#pragma unroll
      for (int z = 0; z < 32; ++z) c[z] = src[lo + (unsigned)z * plane];
    } else {
      const auto src = gm((const float*)a.input) + base;
#pragma unroll
      for (int z = 0; z < 32; ++z) c[z] = (double)src[lo + (unsigned)z * plane];
    }
    c[32] = 0.0;
    if (a.level == 0) {
      double mn = c[0], mx = c[0];
#pragma unroll
      for (int z = 1; z < 32; ++z) {
        mn = c[z] < mn ? c[z] : mn;
        mx = c[z] > mx ? c[z] : mx;
      }
      store_minmax(a, sh, dense, mn, mx);
    }
    __syncthreads();
    IDX2G_PHASE(0);
    // LDS past the subgrid is free until the coding stage (it belongs to CodeLds)
    static_assert(sizeof(S.u.sub) + 8 + sizeof(double) * b32::kParkN * b32::kTh <= sizeof(b32::CodeLds),
                  "park area past the subgrid");
    b32::extrapolate(c, S, p, reinterpret_cast<double*>(&S.u) + (17 * 17 * 17 + 1));
    b32::forward(c, S, p);
    if (!a.coarsest && !(p.x & 1) && !(p.y & 1)) {
      const unsigned pplane = (unsigned)a.pdims[1] * (unsigned)a.pdims[0];
      const auto dst = gm(a.parents) + ((size_t)(b3[2] * 16) * a.pdims[1] + b3[1] * 16) * a.pdims[0] + b3[0] * 16;
      const unsigned po = (unsigned)(p.y >> 1) * (unsigned)a.pdims[0] + (unsigned)(p.x >> 1);
#pragma unroll
      for (int z = 0; z < 16; ++z) dst[po + (unsigned)z * pplane] = c[2 * z];
    }
    b32::weigh(c, S, p, w8);
    IDX2G_PHASE(11);
    b32::code_brick<DT>(a, c, S, *reinterpret_cast<b32::CodeLds*>(&S.u), p, dense,
                    reinterpret_cast<u64*>(a.scratch + (size_t)blockIdx.x * a.wstride), s_blocks, cur, s_perm);
    __syncthreads();
    IDX2G_PHASE(15);
  }
  b32::flush_cursor(cur, table[0].counters);
}

// Full 2D 32x32 bricks, 32 per workgroup iteration (one per 32-lane
// half-wave, brick2d.hpp), coded kGroup bricks per pass (code2d.hpp).  Work
// item g of job j is the batch of bricks [32 (g - start[j]), ...) of its list.
template <int DT>
__global__ void __launch_bounds__(b2d::kTh) k_encode_2d(const EncArgs* __restrict__ table,
                                                        const int* __restrict__ start, int njobs) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ double w4[4];
  __shared__ uint32_t s_blocks[4 * 32];
  __shared__ int s_dense[b2d::kSlots];
  __shared__ int s_off[4];
  __shared__ b32::ArenaCursor cur;
  b2d::Lds2& S = *reinterpret_cast<b2d::Lds2*>(lds_raw);
  b2d::CodeLds2& C = S.u.code;
  if (threadIdx.x == 0) cur = b32::ArenaCursor{0, 0, 0, 0, 0};
  const int total = start[njobs];
  int cj = -1;
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int j = job_of(start, njobs, g);
    const EncArgs& a = table[j];
    const BrickConst& k = a.k;
    if (j != cj) {  // block tables, weights and park offsets of the job's subbands
      __syncthreads();
      for (int i = threadIdx.x; i < 4 * 32; i += blockDim.x) {
        const int sb = i >> 5, b = i & 31;
        s_blocks[i] = b < a.blk_n[sb] ? a.blocks[a.blk_off[sb] + b] : 0u;
      }
      if (threadIdx.x == 0) {
        int off = 0;
        for (int sb = 0; sb < 4; ++sb) {
          const int hx = k.sb[sb].f[0], hy = k.sb[sb].f[1];
          w4[hx + 2 * hy] = (sb == 0 && !a.coarsest) ? 1.0 : a.wfwd[sb];
          s_off[hx + 2 * hy] = off;
          off += (17 - hx) * (17 - hy) * b2d::kSlots;
        }
      }
      cj = j;
      __syncthreads();
    }
    const int batch = g - start[j];
    const int nvalid = min(a.per, a.nb - batch * a.per);
    const b2d::Pos2 p = b2d::pos2();
    const bool has = p.slot < nvalid;
    const int dense = has ? a.list[batch * a.per + p.slot] : 0;
    if (p.x == 0) s_dense[p.slot] = dense;
    const int bx = a.xyz[3 * dense], by = a.xyz[3 * dense + 1];
    double c[b32::kN];
    {
      const unsigned row = (unsigned)a.gdims[0];
      const size_t base = (size_t)(by * 32) * a.gdims[0] + (size_t)bx * 32 + p.x;
      if (a.in_dtype) {
        const auto src = gm((const double*)a.input) + base;
#pragma unroll
        for (int y = 0; y < 32; ++y) c[y] = has ? src[(size_t)y * row] : 0.0;
      } else {
        const auto src = gm((const float*)a.input) + base;
#pragma unroll
        for (int y = 0; y < 32; ++y) c[y] = has ? (double)src[(size_t)y * row] : 0.0;
      }
    }
    c[32] = 0.0;
    if (a.level == 0) {
      double mn = c[0], mx = c[0];
#pragma unroll
      for (int y = 1; y < 32; ++y) {
        mn = c[y] < mn ? c[y] : mn;
        mx = c[y] > mx ? c[y] : mx;
      }
      for (int o = 16; o > 0; o >>= 1) {  // over the slot's 32 lanes
        const double omn = __shfl_xor(mn, o, 64), omx = __shfl_xor(mx, o, 64);
        mn = omn < mn ? omn : mn;
        mx = omx > mx ? omx : mx;
      }
      if (has && p.x == 0) {
        a.minmax[2 * dense] = mn;
        a.minmax[2 * dense + 1] = mx;
      }
    }
    __syncthreads();
    b2d::extrapolate(c, S.ex, S.u.sub, p, nvalid);
    b2d::forward(c, S.ex, p);
    // subband 0 interior into the next level's grid (EncodeBrick, idx2Encode.cpp:303-333)
    if (!a.coarsest && has && !(p.x & 1)) {
      const auto dst = gm(a.parents) + (size_t)(by * 16) * a.pdims[0] + (size_t)bx * 16 + (p.x >> 1);
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) dst[(size_t)jj * a.pdims[0]] = c[2 * jj];
    }
    // subband weights by parity (x = 32 is even), then park every slot's
    // subbands: P[off(hx, hy) + (slot * dy + j) * dx + i] = (2i + hx, 2j + hy)
    double* P = a.scratch + (size_t)blockIdx.x * a.wstride + b2d::kParkOff;
    {
      const double we = w4[p.x & 1], wo = w4[(p.x & 1) + 2];
#pragma unroll
      for (int y = 0; y < b32::kN; ++y) c[y] = c[y] * ((y & 1) ? wo : we);
      for (int t = p.tid; t < b2d::kSlots * b32::kN; t += b2d::kTh) {
        const int y = t % b32::kN;
        S.ex[t] = S.ex[t] * w4[2 * (y & 1)];
      }
    }
#pragma unroll
    for (int hy = 0; hy < 2; ++hy) {
      const int dy = 17 - hy;
      if ((p.x & 1) == 0) {
        double* q = P + s_off[2 * hy] + (p.slot * dy) * 17 + (p.x >> 1);
#pragma unroll
        for (int jj = 0; jj < 17 - hy; ++jj) q[jj * 17] = c[2 * jj + hy];
      } else {
        double* q = P + s_off[1 + 2 * hy] + (p.slot * dy) * 16 + (p.x >> 1);
#pragma unroll
        for (int jj = 0; jj < 17 - hy; ++jj) q[jj * 16] = c[2 * jj + hy];
      }
    }
    __syncthreads();  // EX weighted
    for (int t = p.tid; t < b2d::kSlots * 33; t += b2d::kTh) {  // x = 32 column (hx = 0)
      const int s2 = t / 33, r = t - s2 * 33, hy = r & 1, jj = r >> 1, dy = 17 - hy;
      P[s_off[2 * hy] + (s2 * dy + jj) * 17 + 16] = S.ex[s2 * b32::kN + r];
    }
    if (p.tid < 64) C.p3[p.tid] = (unsigned char)TEnc::pos(a.perm.p3[p.tid]);  // positions in the TEnc layout
    if (p.tid < 16) C.p2[p.tid] = a.perm.p2[p.tid];
    if (p.tid < 256) {
      uint32_t e = 0;
      int o = 0;
      for (int i = 0; i < 8; ++i) {
        if ((p.tid >> i) & 1) { e |= 3u << o; o += 2; }
        else o += 1;
      }
      C.e8[p.tid] = (uint16_t)e;
    }
    __syncthreads();
    b32::Pos pb;
    pb.tid = p.tid;
    pb.lane = p.lane;
    pb.w = p.w;
    pb.x = p.x;
    pb.y = p.slot;
    pb.orow = p.lane >= 32;
    u64* XG = reinterpret_cast<u64*>(a.scratch + (size_t)blockIdx.x * a.wstride);
    for (int sb = 0; sb < 4; ++sb) {
      const int hx = k.sb[sb].f[0], hy = k.sb[sb].f[1];
      const int area = (17 - hx) * (17 - hy);
      for (int s0 = 0; s0 < nvalid; s0 += b2d::kGroup) {
        const int G = min(b2d::kGroup, nvalid - s0);
        b2d::code_group<DT>(a, sb, G, s_dense + s0, P + s_off[hx + 2 * hy] + s0 * area, C, pb, XG, s_blocks + sb * 32,
                            a.blk_n[sb], cur);
      }
    }
    __syncthreads();
  }
  b32::flush_cursor(cur, table[0].counters);
}

// --------------------------------------------------------------------------
// chunk assembly (WriteChunk, Source/Core/idx2Write.cpp:207-290):
//   varbyte(NBricks) | varbyte(first brick) unary(gaps) | varbyte(sizes) | payloads
// The unary gap codes of consecutive bricks form a presence bitmap: brick i
// sets bit (brick_i - first - 1) after the first brick's varbyte.
// --------------------------------------------------------------------------
__device__ inline int varbyte_len(u64 v) {
  int n = 1;
  while (v >>= 7) ++n;
  return n;
}

// Keys of job J: [klo, klo + nk).  The graph form reads them from kr and
// takes none when they exceed the job's rows (k_job_plan flags that).
__device__ __forceinline__ void job_keys(const AsmJob& J, int* klo, int* nk) {
  if (!J.dyn) {
    *klo = J.klo;
    *nk = J.nk;
    return;
  }
  const int lo = J.kr[0], hi = J.kr[1];
  const int n = hi >= lo ? hi - lo + 1 : 0;
  *klo = lo;
  *nk = n <= J.nk ? n : 0;
}

__device__ __forceinline__ bool job_stopped(const AsmJob& J) { return J.dyn && J.guard[2] != 0; }

// Key range of each job's level from its per-brick-subband headers; start[]
// counts workgroups per job.
__global__ void k_key_range(const AsmJob* __restrict__ jobs, const int* __restrict__ start, int njobs) {
  const int j = job_of(start, njobs, blockIdx.x);
  const AsmJob& J = jobs[j];
  const int n = J.nb * J.nsb;
  const int nwg = start[j + 1] - start[j];
  int lo = 0x7fffffff, hi = -0x7fffffff;
  for (int i = (blockIdx.x - start[j]) * blockDim.x + threadIdx.x; i < n; i += nwg * blockDim.x) {
    if (J.hdrs[i].nkeys > 0) {
      lo = min(lo, J.hdrs[i].klo);
      hi = max(hi, J.hdrs[i].klo + J.hdrs[i].nkeys - 1);
    }
  }
  // one atomic pair per workgroup
  lo = wave_min_i(lo);
  hi = wave_max_i(hi);
  __shared__ int s_lo[16], s_hi[16];
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane_id() == 0) {
    s_lo[wave] = lo;
    s_hi[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      lo = min(lo, s_lo[w]);
      hi = max(hi, s_hi[w]);
    }
    if (lo != 0x7fffffff) atomicMin(&J.kr[0], lo);
    if (hi != -0x7fffffff) atomicMax(&J.kr[1], hi);
  }
}

// dense index table: idx[(sb*nk + key-klo)*span + lin] = entry + 1; start[]
// counts workgroups (256 headers each) per job.
__global__ void k_fill_index(const AsmJob* __restrict__ jobs, const int* __restrict__ start, int njobs) {
  const int j = job_of(start, njobs, blockIdx.x);
  const AsmJob& J = jobs[j];
  const int i = (blockIdx.x - start[j]) * blockDim.x + threadIdx.x;
  if (i >= J.nb * J.nsb || job_stopped(J)) return;
  int klo, nk;
  job_keys(J, &klo, &nk);
  if (nk == 0) return;
  const int dense = i / J.nsb, sb = i % J.nsb;
  const SbHdr h = J.hdrs[i];
  for (int t = 0; t < h.nkeys; ++t) {
    const int kk = h.klo + t - klo;
    J.idx[((size_t)sb * nk + kk) * J.span + J.lin[dense]] = (uint32_t)(h.ebase + t + 1);
  }
}

// Sizes of chunk r = (sb, key - klo, chunk) of job J (record index
// [sb][key][chunk] over the job's nk keys).  Every thread of the workgroup
// calls it.
__device__ __forceinline__ void chunk_size_item(const AsmJob& J, int klo, int nk, int r,
                                                const SbEntry* __restrict__ entries) {
  const int nchunks = J.nchunks;
  const int chunk = r % nchunks, kk = (r / nchunks) % nk, sb = r / (nchunks * nk);
  const u64 b0 = (u64)chunk * J.bpc, b1 = min(b0 + (u64)J.bpc, J.span);
  const uint32_t* row = J.idx + ((size_t)sb * nk + kk) * J.span;
  unsigned long long cnt = 0, sz = 0, vb = 0;
  u64 first = ~0ull, last = 0;
  __shared__ unsigned long long s_cnt, s_sz, s_vb, s_first, s_last;
  __shared__ uint32_t s_ssz[kMaxChunkSteps], s_svb[kMaxChunkSteps];  // per step of kChunkStep bricks
  if (threadIdx.x == 0) { s_cnt = s_sz = s_vb = 0; s_first = ~0ull; s_last = 0; }
  for (int i = threadIdx.x; i < J.nsteps; i += blockDim.x) s_ssz[i] = s_svb[i] = 0;
  __syncthreads();
  // blockDim.x == kChunkStep: iteration stp covers step stp of the chunk
  int stp = 0;
  for (u64 bb = b0; bb < b1; bb += blockDim.x, ++stp) {
    const u64 b = bb + threadIdx.x;
    uint32_t s = 0;
    if (b < b1) {
      const uint32_t e = row[b];
      s = e ? entries[e - 1].size : 0u;
    }
    const uint32_t v = s ? (uint32_t)varbyte_len(s) : 0u;
    if (s) {
      cnt += 1;
      sz += s;
      vb += v;
      first = min(first, b);
      last = max(last, b);
    }
    const uint32_t ws = wave_sum_u(s), wv = wave_sum_u(v);
    if (lane_id() == 0 && ws) {
      atomicAdd(&s_ssz[stp], ws);
      atomicAdd(&s_svb[stp], wv);
    }
  }
  __syncthreads();
  if (cnt) {
    atomicAdd(&s_cnt, cnt);
    atomicAdd(&s_sz, sz);
    atomicAdd(&s_vb, vb);
    atomicMin(&s_first, first);
    atomicMax(&s_last, last);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ChunkRec c;
    c.off = 0;
    c.nbricks = (uint32_t)s_cnt;
    c.sb = sb;
    c.key = klo + kk;
    c.chunk = chunk;
    c.first_brick = s_first;
    c.last_brick = s_last;
    c.first_fb = 0xffff;
    c.size = 0;
    if (s_cnt) {
      const u64 delta_bits = 8ull * varbyte_len(s_first) + (s_last - s_first);
      c.size = (uint32_t)(varbyte_len(s_cnt) + (delta_bits + 7) / 8 + s_vb + s_sz);
      c.first_fb = entries[row[s_first] - 1].first_block;
    }
    J.recs[r] = c;
    // offsets of the steps inside the sizes / payload sections, then totals
    uint32_t ov = 0, od = 0;
    uint2* so = J.steps + (size_t)r * (J.nsteps + 1);
    for (int i = 0; i < J.nsteps; ++i) {
      so[i] = make_uint2(ov, od);
      ov += s_svb[i];
      od += s_ssz[i];
    }
    so[J.nsteps] = make_uint2(ov, od);
  }
}

// One workgroup per (job, subband, key, chunk): sizes of the chunk.
__global__ void __launch_bounds__(256) k_chunk_size(const AsmJob* __restrict__ jobs, const int* __restrict__ start,
                                                    int njobs, const SbEntry* __restrict__ entries) {
  const int j = job_of(start, njobs, blockIdx.x);
  const AsmJob& J = jobs[j];
  chunk_size_item(J, J.klo, J.nk, blockIdx.x - start[j], entries);
}

// Zero the byte ranges [dst, dst + n) of the segments (one workgroup each).
__global__ void k_zero(const ScatterSeg* __restrict__ segs) {
  const ScatterSeg sg = segs[blockIdx.x];
  uint8_t* d = (uint8_t*)sg.dst;
  const u64 head = min(sg.n, (u64)((16 - ((uintptr_t)d & 15)) & 15));
  for (u64 i = threadIdx.x; i < head; i += blockDim.x) d[i] = 0;
  const u64 nv = (sg.n - head) / 16;
  uint4* dv = reinterpret_cast<uint4*>(d + head);
  for (u64 i = threadIdx.x; i < nv; i += blockDim.x) dv[i] = make_uint4(0, 0, 0, 0);
  for (u64 i = head + nv * 16 + threadIdx.x; i < sg.n; i += blockDim.x) d[i] = 0;
}

// Copy segments of an uploaded slab to their destinations (one workgroup
// each; 16-byte aligned offsets and destinations).
__global__ void k_scatter(const uint8_t* __restrict__ slab, const ScatterSeg* __restrict__ segs) {
  const ScatterSeg sg = segs[blockIdx.x];
  const uint8_t* s = slab + sg.off;
  uint8_t* d = (uint8_t*)sg.dst;
  const u64 nv = sg.n / 16;
  for (u64 i = threadIdx.x; i < nv; i += blockDim.x) reinterpret_cast<uint4*>(d)[i] = reinterpret_cast<const uint4*>(s)[i];
  for (u64 i = nv * 16 + threadIdx.x; i < sg.n; i += blockDim.x) d[i] = s[i];
}

// One workgroup per (non-empty chunk in blob order, step of kChunkStep
// bricks): the first step writes the chunk's two leading varbytes, every step
// its share of the presence bitmap, the varbyte sizes and payloads of its
// bricks, and where each brick's stream starts (decode directory: blob
// offset + 1).  The steps' offsets inside the sections come from
// k_chunk_size, so the steps of a chunk run in parallel.
// Work item rel of job J (keys [klo, klo + nk), chunks in blob order in
// `order`, the job's blob at `blob`).  Every thread of the workgroup calls it.
__device__ __forceinline__ void chunk_write_item(const AsmJob& J, int klo, int nk, int rel, const int* order,
                                                 uint8_t* blob, const SbEntry* __restrict__ entries,
                                                 const uint8_t* __restrict__ arena) {
  const int bpc = J.bpc, nsteps = J.nsteps;
  const u64 span = J.span;
  const int oi = rel / nsteps, step = rel - oi * nsteps;
  const int ri = order[oi];
  const ChunkRec c = J.recs[ri];
  const u64 b0 = (u64)c.chunk * bpc, b1 = min(b0 + (u64)bpc, span);
  const u64 s0 = b0 + (u64)step * kChunkStep;
  if (s0 >= b1) return;
  const u64 s1e = min(s0 + (u64)kChunkStep, b1);
  const size_t rowi = ((size_t)c.sb * nk + (c.key - klo)) * span;
  const uint32_t* row = J.idx + rowi;
  u64* drow = J.dir + rowi;
  uint8_t* out = blob + c.off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int vn = varbyte_len(c.nbricks), vf = varbyte_len(c.first_brick);
  const u64 nbm = c.last_brick - c.first_brick;  // bitmap bits after the first varbyte
  const u64 dbytes = (8ull * vf + nbm + 7) / 8;
  auto present = [&](u64 b) -> bool {
    const uint32_t e = row[b];
    return e && entries[e - 1].size > 0;
  };
  if (step == 0 && tid == 0) {
    u64 v = c.nbricks;
    int i = 0;
    do { uint8_t byte = v & 0x7f; v >>= 7; out[i++] = byte | (v ? 0x80 : 0); } while (v);
    v = c.first_brick;
    i = 0;
    do { uint8_t byte = v & 0x7f; v >>= 7; out[vn + i++] = byte | (v ? 0x80 : 0); } while (v);
  }
  // this thread's brick of the step: its stream (size 0: absent)
  const u64 b = s0 + tid;
  uint32_t sz = 0;
  u64 src = 0;
  if (b < s1e) {
    const uint32_t e = row[b];
    if (e) { sz = entries[e - 1].size; src = entries[e - 1].off; }
  }
  // unary gaps as a presence bitmap: byte j holds the bits of bricks
  // first_brick + 1 + 8 (j - vf) + i and belongs to the step that holds the
  // brick of its first bit; presence of the step's bricks (and of the seven
  // after it, for the last byte) comes from one lookup per brick, via LDS
  __shared__ uint8_t s_pres[kChunkStep + 8];
  s_pres[tid] = sz > 0;
  if (tid < 8) {
    const u64 bx = s1e + tid;
    s_pres[kChunkStep + tid] = bx < b1 && present(bx);
  }
  __syncthreads();
  {
    const long long f1 = (long long)c.first_brick + 1;
    const long long lo = ((long long)s0 - f1 + 7) >> 3, hi = ((long long)s1e - f1 + 7) >> 3;
    const u64 j0 = (u64)vf + (u64)max(0ll, lo);
    const u64 j1 = min(dbytes, (u64)vf + (u64)max(0ll, hi));
    for (u64 j = j0 + tid; j < j1; j += blockDim.x) {
      uint8_t byte = 0;
      for (int i = 0; i < 8; ++i) {
        const u64 bit = 8 * (j - vf) + i;
        // brick f1 + bit sits at s_pres[f1 + bit - s0] (in [0, kChunkStep + 8))
        const long long at = f1 + (long long)bit - (long long)s0;
        if (bit < nbm && at >= 0 && at < s1e - s0 + 8 && s_pres[at < (long long)(s1e - s0) ? at : kChunkStep + (at - (long long)(s1e - s0))])
          byte |= uint8_t(1u << i);
      }
      out[vn + j] = byte;
    }
  }
  __shared__ uint32_t s1[4], s2[4];
  __shared__ uint32_t s_pre[kChunkStep];
  __shared__ u64 s_src[kChunkStep];
  const uint2* so = J.steps + (size_t)ri * (nsteps + 1);
  const uint2 sto = so[step];
  const u64 sizes_off = (u64)vn + dbytes;
  const u64 payload_start = sizes_off + so[nsteps].x;
  {
    const uint32_t v1 = sz ? (uint32_t)varbyte_len(sz) : 0u;
    // exclusive prefix sums over the step's bricks: wave scans + wave totals
    const uint32_t i1 = b32::wave_scan_add(v1, lane), i2 = b32::wave_scan_add(sz, lane);
    if (lane == 63) { s1[wave] = i1; s2[wave] = i2; }
    __syncthreads();
    uint32_t w1 = 0, w2 = 0, t2 = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (w < wave) { w1 += s1[w]; w2 += s2[w]; }
      t2 += s2[w];
    }
    const u64 vboff = (u64)sto.x + w1 + i1 - v1;
    const uint32_t pre = w2 + i2 - sz;  // payload offset inside the step
    const u64 pbase = c.off + payload_start + sto.y;  // the step's payload in the blob
    const u64 dst = pbase + pre;
    if (sz) {
      uint8_t* vp = out + sizes_off + vboff;
      u64 v = sz;
      int i = 0;
      do { uint8_t byte = v & 0x7f; v >>= 7; vp[i++] = byte | (v ? 0x80 : 0); } while (v);
      drow[b] = dir_entry(dst, sz);
    }
    // payloads, destination-driven: each thread fills 16-byte aligned units
    // of the step's payload range [pbase, pbase + t2).  The stream of a
    // unit's first byte comes from a binary search over the step's payload
    // offsets; a unit inside one stream is five aligned source loads and one
    // 16-byte store, the rest (stream seams, the range's two ends) go byte
    // by byte.
    s_pre[tid] = pre;
    s_src[tid] = src;
    __syncthreads();
    const uintptr_t p0 = (uintptr_t)(blob + pbase);
    for (uintptr_t u = (p0 & ~(uintptr_t)15) + 16 * (uintptr_t)tid; u < p0 + t2; u += 16 * (uintptr_t)blockDim.x) {
      const long long r0 = (long long)u - (long long)p0;  // negative for a unit that starts before the range
      const uint32_t rs = r0 < 0 ? 0u : (uint32_t)r0;
      int j = 0;
#pragma unroll
      for (int st = kChunkStep / 2; st; st >>= 1)
        if (s_pre[j + st] <= rs) j += st;
      const uint32_t end_j = j < kChunkStep - 1 ? s_pre[j + 1] : t2;
      const bool whole = r0 >= 0 && (u64)r0 + 16 <= t2;
      if (whole && (u64)r0 + 16 <= end_j) {
        const u64 sp = s_src[j] + ((u64)r0 - s_pre[j]);
        const uint32_t* sw = reinterpret_cast<const uint32_t*>(arena + (sp & ~3ull));
        const uint32_t sh = (uint32_t)(sp & 3);
        uint32_t w[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) w[i] = sw[i];
        uint4 o;
        o.x = __builtin_amdgcn_alignbyte(w[1], w[0], sh);
        o.y = __builtin_amdgcn_alignbyte(w[2], w[1], sh);
        o.z = __builtin_amdgcn_alignbyte(w[3], w[2], sh);
        o.w = __builtin_amdgcn_alignbyte(w[4], w[3], sh);
        *reinterpret_cast<uint4*>(u) = o;
      } else {
        // source address of every in-range byte first (LDS only), then all
        // sixteen loads in flight together
        u64 sa[16];
        int jj = j;
#pragma unroll
        for (int bi = 0; bi < 16; ++bi) {
          const long long r = r0 + bi;
          sa[bi] = ~0ull;
          if (r >= 0 && r < (long long)t2) {
            while (jj < kChunkStep - 1 && s_pre[jj + 1] <= (uint32_t)r) ++jj;
            sa[bi] = s_src[jj] + ((u64)r - s_pre[jj]);
          }
        }
        uint8_t v[16];
#pragma unroll
        for (int bi = 0; bi < 16; ++bi) v[bi] = sa[bi] != ~0ull ? arena[sa[bi]] : (uint8_t)0;
        if (whole) {
          uint32_t o[4];
#pragma unroll
          for (int i = 0; i < 4; ++i)
            o[i] = (uint32_t)v[4 * i] | ((uint32_t)v[4 * i + 1] << 8) | ((uint32_t)v[4 * i + 2] << 16) |
                   ((uint32_t)v[4 * i + 3] << 24);
          *reinterpret_cast<uint4*>(u) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
          for (int bi = 0; bi < 16; ++bi)
            if (sa[bi] != ~0ull) reinterpret_cast<uint8_t*>(u)[bi] = v[bi];
        }
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_chunk_write(const AsmJob* __restrict__ jobs, const int* __restrict__ start,
                                                     int njobs, const SbEntry* __restrict__ entries,
                                                     const uint8_t* __restrict__ arena) {
  const int jb = job_of(start, njobs, blockIdx.x);
  const AsmJob& J = jobs[jb];
  chunk_write_item(J, J.klo, J.nk, blockIdx.x - start[jb], J.order, J.blob, entries, arena);
}

// ---- graph form of the chunk assembly ----

// Per job (one thread, in order): its key count from kr (flags bit 2, value
// 4, in flags[2] when it exceeds the job's rows) and the first chunk-size
// work item, cs_start[j] (njobs + 1 entries).
__global__ void k_job_plan(const AsmJob* __restrict__ jobs, int njobs, int* cs_start, unsigned long long* flags) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  bool over = false;
  const bool stopped = flags[2] != 0;
  for (int j = 0; j < njobs; ++j) {
    const AsmJob& J = jobs[j];
    const int lo = J.kr[0], hi = J.kr[1];
    over = over || (hi >= lo && hi - lo + 1 > J.nk);
    int klo, nk;
    job_keys(J, &klo, &nk);
    cs_start[j] = acc;
    if (!stopped) acc += J.nsb * nk * J.nchunks;
  }
  cs_start[njobs] = acc;
  if (over) atomicOr(&flags[2], 4ull);
}

// Clear the rows of each job's stream index and directory that its keys use
// (blockIdx.y: the job).
__global__ void k_zero_rows(const AsmJob* __restrict__ jobs) {
  const AsmJob& J = jobs[blockIdx.y];
  if (job_stopped(J)) return;
  int klo, nk;
  job_keys(J, &klo, &nk);
  const size_t n = (size_t)J.nsb * nk * J.span;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    J.idx[i] = 0;
    J.dir[i] = 0;
  }
}

// The next work item of a persistent workgroup from a queue counter (chunks
// differ widely in size, so workgroups take items as they free up).
__device__ __forceinline__ int next_item(int* queue) {
  __shared__ int s_item;
  __syncthreads();
  if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
  __syncthreads();
  return s_item;
}

// Chunk sizes over the work items k_job_plan counted (persistent: items of
// similar cost, taken round robin).
__global__ void __launch_bounds__(256) k_chunk_size_dyn(const AsmJob* __restrict__ jobs, const int* __restrict__ cs_start,
                                                        int njobs, const SbEntry* __restrict__ entries) {
  const int total = cs_start[njobs];
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int j = job_of(cs_start, njobs, item);
    const AsmJob& J = jobs[j];
    if (job_stopped(J)) return;
    int klo, nk;
    job_keys(J, &klo, &nk);
    chunk_size_item(J, klo, nk, item - cs_start[j], entries);
    __syncthreads();  // the item's LDS before the next item's
  }
}

__device__ __forceinline__ u64 shfl_up_u64(u64 v, int d) {
  const int lo = __shfl_up((int)(uint32_t)v, d, 64), hi = __shfl_up((int)(uint32_t)(v >> 32), d, 64);
  return ((u64)(uint32_t)hi << 32) | (uint32_t)lo;
}

// One workgroup of 1024 threads per job: chunk offsets inside the job's blob
// (prefix sums of the chunk sizes in [subband][key][chunk] order), the
// non-empty chunks in that order (order_dyn), the blob bytes (place[1]) and
// the number of non-empty chunks (place[2]).
__global__ void __launch_bounds__(1024) k_blob_scan(const AsmJob* __restrict__ jobs) {
  __shared__ u64 s_sz[16];
  __shared__ uint32_t s_ct[16];
  const AsmJob& J = jobs[blockIdx.x];
  if (job_stopped(J)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int klo, nk;
  job_keys(J, &klo, &nk);
  const int n = J.nsb * nk * J.nchunks;
  u64 off = 0;
  int cnt = 0;
  for (int i0 = 0; i0 < n; i0 += 1024) {
    const int i = i0 + tid;
    const uint32_t sz = i < n ? J.recs[i].size : 0u;
    u64 x = sz;
    uint32_t c = sz ? 1u : 0u;
    for (int d = 1; d < 64; d <<= 1) {  // inclusive wave scans
      const u64 t = shfl_up_u64(x, d);
      const uint32_t tc = (uint32_t)__shfl_up((int)c, d, 64);
      if (lane >= d) {
        x += t;
        c += tc;
      }
    }
    if (lane == 63) {
      s_sz[wave] = x;
      s_ct[wave] = c;
    }
    __syncthreads();
    u64 wx = 0, tot = 0;
    uint32_t wc = 0, totc = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wave) {
        wx += s_sz[w];
        wc += s_ct[w];
      }
      tot += s_sz[w];
      totc += s_ct[w];
    }
    if (i < n) {
      J.recs[i].off = off + wx + x - sz;
      if (sz) J.order_dyn[cnt + (int)(wc + c - 1u)] = i;
    }
    off += tot;
    cnt += (int)totc;
    __syncthreads();
  }
  if (tid == 0) {
    J.place[1] = off;
    J.place[2] = (u64)cnt;
  }
}

// The jobs one after the other in the batch blob (one thread, 256-byte
// aligned: place[0]), the first chunk-writer work item of each job (wstart)
// and its first record in chunks_out (cstart); both arrays njobs + 1 entries.
// Flags 4 when the blob exceeds blob_cap.
__global__ void k_blob_place(const AsmJob* __restrict__ jobs, int njobs, int* wstart, int* cstart, u64 blob_cap,
                             unsigned long long* flags, int* queue) {
  if (threadIdx.x != 0) return;
  queue[0] = 0;  // k_chunk_write_dyn's next work item
  const bool stopped = flags[2] != 0;
  u64 base = 0;
  int items = 0, chunks = 0;
  for (int j = 0; j < njobs && !stopped; ++j) {
    const AsmJob& J = jobs[j];
    J.place[0] = base;
    wstart[j] = items;
    cstart[j] = chunks;
    items += (int)J.place[2] * J.nsteps;
    chunks += (int)J.place[2];
    base = (base + J.place[1] + 255) & ~255ull;
  }
  if (stopped) {
    for (int j = 0; j < njobs; ++j) wstart[j] = cstart[j] = 0;
  }
  wstart[njobs] = items;
  cstart[njobs] = chunks;
  if (!stopped && base > blob_cap) atomicOr(&flags[2], 4ull);
}

// The non-empty chunk records of every job, in order, into chunks_out from
// cstart[j] on (blockIdx.x: the job).
__global__ void k_chunks_out(const AsmJob* __restrict__ jobs, const int* __restrict__ cstart, ChunkRec* chunks_out) {
  const AsmJob& J = jobs[blockIdx.x];
  if (job_stopped(J)) return;
  const int n = (int)J.place[2];
  for (int k = threadIdx.x; k < n; k += blockDim.x) chunks_out[cstart[blockIdx.x] + k] = J.recs[J.order_dyn[k]];
}

// Copies (and zero fills) of the segments: blockIdx.y is the segment, the
// blockIdx.x-th of gridDim.x parts of it; 16-byte units where source and
// destination are 16-byte aligned, bytes elsewhere.
__global__ void __launch_bounds__(256) k_copy_segs(const CopySeg* __restrict__ segs) {
  const CopySeg sg = segs[blockIdx.y];
  const u64 stride = (u64)gridDim.x * blockDim.x;
  const u64 t0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const bool vec = (((uintptr_t)sg.src | (uintptr_t)sg.dst) & 15) == 0;
  const u64 nv = vec ? sg.n / 16 : 0;
  for (u64 i = t0; i < nv; i += stride)
    reinterpret_cast<uint4*>(sg.dst)[i] = sg.src ? reinterpret_cast<const uint4*>(sg.src)[i] : make_uint4(0, 0, 0, 0);
  for (u64 i = nv * 16 + t0; i < sg.n; i += stride) sg.dst[i] = sg.src ? sg.src[i] : (uint8_t)0;
}

// The chunk writer over the work items k_blob_place counted (persistent,
// items taken from queue[0]).
__global__ void __launch_bounds__(256) k_chunk_write_dyn(const AsmJob* __restrict__ jobs, const int* __restrict__ wstart,
                                                         int njobs, const SbEntry* __restrict__ entries,
                                                         const uint8_t* __restrict__ arena, uint8_t* blob_all,
                                                         const unsigned long long* __restrict__ flags, int* queue) {
  if (flags[2]) return;
  const int total = wstart[njobs];
  for (int item = next_item(queue); item < total; item = next_item(queue)) {
    const int jb = job_of(wstart, njobs, item);
    const AsmJob& J = jobs[jb];
    int klo, nk;
    job_keys(J, &klo, &nk);
    chunk_write_item(J, klo, nk, item - wstart[jb], J.order_dyn, blob_all + J.place[0], entries, arena);
  }
}

}  // namespace idx2g
