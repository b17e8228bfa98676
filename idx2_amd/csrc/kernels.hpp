// Kernel argument blocks shared by the host drivers and the HIP kernels.
#pragma once

#include "device.hpp"

namespace idx2g {

// Row-mode transpose buffers of the block stages: coefficient x = a + 4 b +
// 16 c of a 4x4x4 block sits at a + 4 b + C c, and a wave's four blocks lie
// Q words apart.  The padding (C > 16, Q > 64) spreads each lifting pass's
// lanes over the LDS banks: a 32-lane half of the row, second-axis and
// third-axis passes hits 24-32 distinct 8-byte bank pairs instead of 8-16.
template <int Q, int C>
struct TLayout {
  static constexpr int q = Q, c = C;
  __host__ __device__ static constexpr int pos(int x) { return (x & 15) + C * (x >> 4); }
};
using TEnc = TLayout<68, 17>;  // encode (the coder's LDS has room for 2 KB more)
using TDec = TLayout<80, 19>;

// One subband of a 32^3 brick (at most 17^3 samples) staged in LDS: sample
// (i, j, k) at (k * kVY + j) * kVX + i, rows of 17 and planes of 17 rows for
// every subband (16 or 17 samples per axis).  With rows of 16, the block
// stage's row reads of an x-high subband hit 2 of 32 bank pairs per 16
// lanes; with rows of 17 they spread over 12-16.  (Planes of 19 rows would
// spread them further but do not fit the encode kernel's LDS.)
constexpr int kVX = 17, kVY = 17, kVXY = kVX * kVY;


constexpr int kThreads = 256;    // threads per brick workgroup (4 waves)
constexpr int kMaxBlocks = 125;  // 4^d blocks per subband (17^3 subband of a 32^3 brick)
constexpr int kMaxKeys = 128;    // bit-plane keys per brick-subband
constexpr int kStageWords = 4096;  // LDS staging for one subband's streams (16 KB)
constexpr short kSkip = -32768;

// block table entry of the on-chip kernels (bricks of at most 32 samples
// per axis): org(5 bits x3) | dims(3 bits x3) << 15 | skip << 24 | coded index << 25
__host__ __device__ inline uint32_t pack_block(const int org[3], const int dims[3], int skip, int ci) {
  return uint32_t(org[0] | (org[1] << 5) | (org[2] << 10) | (dims[0] << 15) | (dims[1] << 18) |
                  (dims[2] << 21) | (skip << 24)) | (uint32_t(ci) << 25);
}
// block table entry of the generic kernels (any brick size this build
// takes): org(8 bits x3) | dims(3 bits x3) << 24 | skip << 33 | coded index << 34
constexpr int kMaxBrickExtent = 128;      // samples per axis (subbands of <= 65: 17 blocks per axis)
constexpr int kMaxGenericBlocks = 4913;   // 4^d blocks per subband (17^3)
__host__ __device__ inline u64 pack_block64(const int org[3], const int dims[3], int skip, int ci) {
  return (u64)org[0] | ((u64)org[1] << 8) | ((u64)org[2] << 16) | ((u64)dims[0] << 24) | ((u64)dims[1] << 27) |
         ((u64)dims[2] << 30) | ((u64)skip << 33) | ((u64)ci << 34);
}
__host__ __device__ inline void unpack_block64(u64 w, int org[3], int dims[3], int* skip, int* ci) {
  org[0] = int(w & 255); org[1] = int((w >> 8) & 255); org[2] = int((w >> 16) & 255);
  dims[0] = int((w >> 24) & 7); dims[1] = int((w >> 27) & 7); dims[2] = int((w >> 30) & 7);
  *skip = int((w >> 33) & 1);
  *ci = int(w >> 34);
}

// Chunk directory entry of one (subband, key, brick) stream in a level's
// chunk blob: (byte offset + 1) | size << 40; 0 = no stream.
__host__ __device__ inline u64 dir_entry(u64 off, u64 size) { return (off + 1) | (size << 40); }
__host__ __device__ inline u64 dir_offset(u64 e) { return (e & ((1ull << 40) - 1)) - 1; }
__host__ __device__ inline uint32_t dir_size(u64 e) { return (uint32_t)(e >> 40); }

struct SbEntry {  // one (brick, subband, key) stream in the arena
  u64 off;
  uint32_t size;
  uint16_t first_block;
  uint16_t key;
};

struct SbHdr {  // per (brick, subband)
  int klo, nkeys;
  u64 ebase;
};

struct EncArgs {
  BrickConst k;
  PermTables perm;
  int level, coarsest;
  // input grid of this level: the caller's grid at level 0, the dense f64
  // grid of subband-0 interiors above
  const void* input;
  int in_dtype;          // 0 f32, 1 f64
  int gdims[3], brick[3];
  // bricks of this level, dense (linear-id) order
  const int* xyz;        // [nb_level][3]
  const int* list;       // dense ids to process, or null for [b0, b0 + nb)
  int nb, b0;
  int per;               // 2D kernels: bricks per workgroup iteration (1..32)
  double* scratch;       // [gridDim.x][wstride] (generic kernel: work volume; b32: parked z-odd half)
  int wstride;           // doubles per workgroup slot of scratch
  int lds_blocks;        // generic kernel: blocks per subband its tables hold (launch-wide)
  int lds_brick;         // generic kernel: byte offset of the extended brick in LDS, -1: in scratch
  int tab_off;           // generic kernel: per-subband tables at this double offset of the workgroup's
                         // scratch slot (past the extended brick), -1: in LDS
  // next level's grid (non-coarsest): subband-0 interiors at brick * sbne
  double* parents;
  int pdims[3];
  int sbne[3];           // non-extended subband-0 dims
  // block tables
  const uint32_t* blocks;  // on-chip kernels (pack_block)
  const u64* blocks64;     // generic kernel (pack_block64)
  int blk_off[8], blk_n[8], ncoded[8];
  double wfwd[8];
  int etol;
  // outputs
  uint8_t* exps;          // per subband [dense][ncoded] of 1 (f32) or 2 (f64) bytes
  long long exp_off[8];
  uint8_t* arena;
  unsigned long long* counters;  // [0] arena bytes, [1] entries, [2] error flags
  unsigned long long arena_cap, entry_cap;
  SbEntry* entries;
  SbHdr* hdrs;          // [dense][nsb]
  double* minmax;       // [dense][2] (level 0)
};

// Multi-job launches: work item g of a launch belongs to job j with
// start[j] <= g < start[j + 1] (start[njobs] = total items).
__device__ __forceinline__ int job_of(const int* start, int njobs, int g) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Chunk assembly of one (dataset, level): key range, dense stream index,
// chunk sizes and chunk bytes (WriteChunk, Source/Core/idx2Write.cpp:207-290).
struct ChunkRec;
struct AsmJob {
  const SbHdr* hdrs;         // [nb][nsb]
  const u64* lin;            // [nb] linear brick ids
  uint32_t* idx;             // [nsb][nk][span] entry + 1
  ChunkRec* recs;            // [nsb][nk][nchunks]
  uint2* steps;              // [nsb][nk][nchunks][nsteps + 1]: (sizes, payload) byte offsets of each step, then totals
  const int* order;          // non-empty chunks in blob order
  uint8_t* blob;
  u64* dir;
  int* kr;                   // key range out: [lo, hi]
  u64 span;
  int nb, nsb, klo, nk, bpc, nchunks, nsteps;
  // Graph form (dyn = 1): the key range is read from kr on the device and
  // nk is the cap of the job's key rows (idx, dir, recs and steps are sized
  // for it); k_blob_scan writes the chunk order into order_dyn and place[1]
  // (blob bytes), place[2] (non-empty chunks), k_blob_place place[0] (offset
  // in the batch blob).  guard: the batch's error flags; every graph-form
  // kernel does nothing once one is set.
  int dyn;
  int* order_dyn;
  u64* place;
  const unsigned long long* guard;
};

constexpr int kChunkStep = 256;                  // bricks per chunk-writer workgroup
constexpr int kMaxChunkSteps = 32768 / kChunkStep;  // bricks per chunk <= 32768

// Device copies of small host arrays: bytes [off, off + n) of the uploaded
// slab go to dst.
struct ScatterSeg {
  u64 off, n;
  void* dst;
};

// Device-to-device copy of n bytes (src == nullptr: zero fill).
struct CopySeg {
  const uint8_t* src;
  uint8_t* dst;
  u64 n;
};

// per (subband, key, chunk) assembled chunk
struct ChunkRec {
  u64 off;          // byte offset in the chunk blob
  uint32_t size;    // bytes (0: no chunk)
  uint32_t nbricks;
  u64 first_brick;  // linear id of the first brick
  u64 last_brick;
  uint32_t first_fb;  // first block (Morton rank) of key in first_brick
  int sb, key, chunk;
};

// Device-side query planning of one (query, level): the bricks of the query
// box [bf, bf + nbox) in brick coordinates, the full bricks first (those in
// the box's corner [bf, bf + nfull3), which the on-chip kernels take), then
// the rest, each with its linear id (bit interleave of the level's brick
// order, GetLinearBrick, Source/Core/idx2Lookup.cpp:13-40) and dense index
// (rank among the level's sorted linear ids).
struct PlanArgs {
  int bf[3], nbox[3], nfull3[3];
  int nfull, n;
  int nbits;                 // length of the level's brick order
  signed char axis[64];      // axis of each order position, most significant first
  const u64* sorted_lin;     // the level's bricks, sorted linear ids
  int nb_level;
  int* xyz;                  // out [n][3]
  u64* lin;                  // out [n]
  int* dense;                // out [n]
};

struct DecArgs {
  BrickConst k;
  PermTables perm;
  int level, coarsest;
  const int* xyz;        // [n][3] brick coordinates to decode (batch)
  const u64* lin;        // [n] linear ids
  const int* dense;      // [n] dense ids (exponent rows)
  int nb;
  int per;               // 2D kernels: bricks per workgroup iteration (1..32)
  double* scratch;       // per workgroup slot: work volume (generic kernel), parked z-odd half (b32)
  int wstride;           // doubles per workgroup slot
  int lds_blocks;        // generic kernel: blocks per subband its tables hold (launch-wide)
  int lds_brick;         // generic kernel: byte offset of the extended brick in LDS, -1: in scratch
  int tab_off;           // generic kernel: per-subband tables at this double offset of the workgroup's
                         // scratch slot (past the extended brick), -1: in LDS
  u64* xs;               // b32: plane words of the batch [brick][subband][128 blocks][npl]
  uint8_t* nbps;         // b32: planes decoded per block [brick][subband][128]
  // Level grids hold only the query box's bricks: pgrid / cgrid are shifted
  // base pointers, so that (Z * dims[1] + Y) * dims[0] + X with the level
  // grid's global coordinates lands inside the box (dims: the box's extent).
  const double* pgrid;   // level+1 decoded data (dense f64; subband-0 interiors of this level)
  int pgdims[3];
  double* cgrid;         // this level's decoded data when it is not the output level
  int cgdims[3];
  int cgend[3];          // global end (exclusive) of the cgrid box per axis
  int sbne[3];
  const uint32_t* blocks;  // on-chip kernels (pack_block)
  const u64* blocks64;     // generic kernel (pack_block64)
  int blk_off[8], blk_n[8], ncoded[8];
  double winv[8];
  const uint8_t* blob;
  const u64* dir;        // [nsb][nkeys][span] stream byte offset + 1, 0 = absent
  int dir_klo, dir_nk;
  u64 span;              // linear id span of the level
  const uint8_t* exps;
  long long exp_off[8];
  int mask;
  int spmask[8][3];      // block spacing - 1 per subband and axis (spacings are powers of two)
  int etol;
  // output grid (only when out != nullptr)
  void* out;
  int og_f[3], og_d[3], og_s[3];
  int bstride;           // 1 << level
  int brick[3], B3[3];   // brick dims and brick extent in samples at this level
  unsigned long long* counters;  // [0] error flags
  // sparse-brick output (generic kernel): the extended brick after the
  // inverse transform, k.nvox doubles per brick, and whether any block of a
  // counted subband decoded a plane (1) or not (0); nullptr: not wanted
  double* pool_vol;
  uint8_t* pool_sig;
};

}  // namespace idx2g
