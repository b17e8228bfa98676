// idx2g driver: encode/decode orchestration on one GPU and the C ABI.
// Kernels live in encode_kernels.hip / decode_kernels.hip (same translation
// unit, so the launches below see them directly).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <sys/stat.h>
#include <tuple>
#include <vector>

#include "decode_kernels.hip"
#include "encode_kernels.hip"
#include "host.hpp"

using namespace idx2g;

namespace idx2g {

static const size_t kBlobPad = 64;

template <class T>
static int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(dev_alloc((void**)p, n * sizeof(T)));
  return IDX2G_OK;
}
// Copy a kernel argument block into the next slot of the context's ring and
// return its device address.  The ring is drained (stream synchronised)
// before a slot is reused.
static constexpr size_t kArgSlot = 2048, kArgSlots = 256;
template <class T>
static int stage_args(idx2g_ctx* ctx, const T& a, const T** out) {
  static_assert(sizeof(T) <= kArgSlot, "argument block too large");
  if (ctx->arg_slot == (int)kArgSlots) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->arg_slot = 0;
  }
  const size_t off = (size_t)ctx->arg_slot++ * kArgSlot;
  memcpy(ctx->h_args + off, &a, sizeof(T));
  HIPCHK(hipMemcpyAsync(ctx->d_args + off, ctx->h_args + off, sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  *out = (const T*)(ctx->d_args + off);
  return IDX2G_OK;
}

#define TRY(x)                  \
  do {                          \
    int rc_ = (x);              \
    if (rc_) return rc_;        \
  } while (0)

static void free_level(LevelData& lv) {
  if (lv.d_blob) dev_free(lv.d_blob);
  if (lv.d_dir) dev_free(lv.d_dir);
  if (lv.d_exps) dev_free(lv.d_exps);
  if (lv.d_xyz) dev_free(lv.d_xyz);
  lv = LevelData();
}

static int check_layout(const Layout& L) {
  for (int l = 0; l < L.nlevels; ++l)
    for (int s = 0; s < L.nsb; ++s)
      if ((int)L.blocks[l][s].size() > kMaxBlocks) {
        set_error("brick size gives %zu blocks per subband (max %d)", L.blocks[l][s].size(), kMaxBlocks);
        return IDX2G_NOT_SUPPORTED;
      }
  return IDX2G_OK;
}

static void level_lists(idx2g_dataset* ds) {
  for (int l = 0; l < ds->L.nlevels; ++l) {
    LevelData& lv = ds->lv[l];
    ds->L.level_bricks(l, &lv.lin, &lv.xyz);
    lv.nb = (int)lv.lin.size();
  }
}

static void exp_layout(idx2g_dataset* ds) {
  const Layout& L = ds->L;
  const int es = L.dtype ? 2 : 1;
  for (int l = 0; l < L.nlevels; ++l) {
    LevelData& lv = ds->lv[l];
    long long off = 0;
    for (int s = 0; s < L.nsb; ++s) {
      lv.exp_off[s] = off;
      off += (long long)lv.nb * L.ncoded[l][s] * es;
    }
    lv.exp_bytes = (uint64_t)off;
  }
}

// ------------------------------------------------------------------------
// encode
// ------------------------------------------------------------------------
struct EncodeBuffers {
  uint8_t* arena = nullptr;
  SbEntry* entries = nullptr;
  unsigned long long* counters = nullptr;
  unsigned long long arena_cap = 0, entry_cap = 0;
  SbHdr* hdrs[kMaxLevels] = {nullptr};
  void release() {
    if (arena) dev_free(arena);
    if (entries) dev_free(entries);
    if (counters) dev_free(counters);
    for (auto& h : hdrs)
      if (h) { dev_free(h); h = nullptr; }
    arena = nullptr; entries = nullptr; counters = nullptr;
  }
};

// Bricks of a level that the on-chip kernel takes: 3D, brick 32^3, data
// extent 32 on every axis.
static bool full_b32(const Layout& L, const int* b3, const int gd[3]) {
  // the kernel addresses a brick with 32-bit offsets from its origin
  if ((uint64_t)gd[0] * (uint64_t)gd[1] * 33ull >= (1ull << 32)) return false;
  for (int d = 0; d < 3; ++d)
    if (L.brick[d] != 32 || (b3[d] + 1) * 32 > gd[d]) return false;
  return true;
}

static int run_encode_levels(idx2g_ctx* ctx, idx2g_dataset* ds, const void* d_input, EncodeBuffers& B,
                             unsigned long long* flags, unsigned long long* l0_bytes, int64_t* parent_bytes) {
  const Layout& L = ds->L;
  const BrickConst kc = make_brick_const(L);
  const PermTables pt = make_perm_tables();
  hipStream_t st = ctx->stream;
  const int npl = L.dtype ? 64 : 32;
  const size_t lds = enc_lds_bytes(npl);
  const size_t lds32 = std::max(sizeof(b32::Lds), offsetof(b32::Lds, u) + sizeof(b32::CodeLds));
  HIPCHK(hipFuncSetAttribute((const void*)k_encode_level, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const void* kb32 = L.dtype ? (const void*)k_encode_b32<1> : (const void*)k_encode_b32<0>;
  HIPCHK(hipFuncSetAttribute(kb32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds32));
  HIPCHK(hipMemsetAsync(B.counters, 0, 4 * sizeof(unsigned long long), st));
  int sbne[3], gd[3];
  for (int d = 0; d < 3; ++d) {
    sbne[d] = L.sb[0].dims[d] - (L.sb[0].dims[d] > 1);
    gd[d] = L.dims[d];
  }
  const int grid_max = std::max(1, ctx->ncu);
  double* scratch = nullptr;
  TRY(dalloc(&scratch, (size_t)std::max(grid_max, std::min(ds->lv[0].nb, ctx->ncu * 4)) * kc.nvox));
  const void* level_in = d_input;
  double* level_grid = nullptr;  // owned grid of the current level (level > 0)
  *parent_bytes = 0;
  int rc = IDX2G_OK;
  for (int l = 0; l < L.nlevels && !rc; ++l) {
    LevelData& lv = ds->lv[l];
    const bool coarsest = l + 1 == L.nlevels;
    double* parents = nullptr;
    int pd[3] = {1, 1, 1};
    if (!coarsest) {
      size_t n = 1;
      for (int d = 0; d < 3; ++d) {
        pd[d] = L.nbricks[l][d] * sbne[d];
        n *= (size_t)pd[d];
      }
      TRY(dalloc(&parents, n));
      if (l == 0) *parent_bytes = (int64_t)(n * sizeof(double));
    }
    TRY(dalloc(&lv.d_xyz, 3 * lv.nb));
    HIPCHK(hipMemcpyAsync(lv.d_xyz, lv.xyz.data(), 3 * lv.nb * sizeof(int), hipMemcpyHostToDevice, st));
    TRY(dalloc(&lv.d_exps, lv.exp_bytes + 16));
    TRY(dalloc(&B.hdrs[l], (size_t)lv.nb * L.nsb));
    // split the level's bricks between the on-chip kernel and the generic one
    std::vector<int> full, rest;
    for (int i = 0; i < lv.nb; ++i) (full_b32(L, &lv.xyz[3 * i], gd) ? full : rest).push_back(i);
    int* d_lists = nullptr;
    TRY(dalloc(&d_lists, (size_t)lv.nb));
    if (!full.empty())
      HIPCHK(hipMemcpyAsync(d_lists, full.data(), full.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (!rest.empty())
      HIPCHK(hipMemcpyAsync(d_lists + full.size(), rest.data(), rest.size() * sizeof(int), hipMemcpyHostToDevice, st));
    EncArgs a;
    memset(&a, 0, sizeof a);
    a.k = kc;
    a.perm = pt;
    a.level = l;
    a.coarsest = coarsest;
    a.input = level_in;
    a.in_dtype = l == 0 ? L.dtype : 1;
    for (int d = 0; d < 3; ++d) {
      a.gdims[d] = gd[d];
      a.brick[d] = L.brick[d];
      a.sbne[d] = sbne[d];
      a.pdims[d] = pd[d];
    }
    a.xyz = lv.d_xyz;
    a.scratch = scratch;
    a.parents = parents;
    a.blocks = ds->d_blocks;
    for (int s = 0; s < L.nsb; ++s) {
      a.blk_off[s] = ds->blk_off[l][s];
      a.blk_n[s] = (int)L.blocks[l][s].size();
      a.ncoded[s] = L.ncoded[l][s];
      a.wfwd[s] = L.wfwd[l][s];
      a.exp_off[s] = lv.exp_off[s];
    }
    int etol;
    {
      int e;
      frexp(L.tolerance, &e);
      etol = L.tolerance != 0 ? std::max(e, -1022) : -1023;
    }
    a.etol = etol;
    a.exps = lv.d_exps;
    a.arena = B.arena;
    a.counters = B.counters;
    a.arena_cap = B.arena_cap;
    a.entry_cap = B.entry_cap;
    a.entries = B.entries;
    a.hdrs = B.hdrs[l];
    a.minmax = ds->d_minmax;
    if (l == 0) HIPCHK(hipEventRecord(ctx->ev[3], st));
    if (!full.empty()) {
      a.list = d_lists;
      a.nb = (int)full.size();
      const EncArgs* da;
      TRY(stage_args(ctx, a, &da));
      if (L.dtype)
        hipLaunchKernelGGL(k_encode_b32<1>, dim3(std::min(a.nb, grid_max)), dim3(b32::kTh), lds32, st, da);
      else
        hipLaunchKernelGGL(k_encode_b32<0>, dim3(std::min(a.nb, grid_max)), dim3(b32::kTh), lds32, st, da);
      HIPCHK(hipGetLastError());
    }
    if (l == 0 && !full.empty()) HIPCHK(hipEventRecord(ctx->ev[4], st));
    if (!rest.empty()) {
      a.list = d_lists + full.size();
      a.nb = (int)rest.size();
      const EncArgs* da;
      TRY(stage_args(ctx, a, &da));
      hipLaunchKernelGGL(k_encode_level, dim3(std::min(a.nb, ctx->ncu * 4)), dim3(kThreads), lds, st, da);
      HIPCHK(hipGetLastError());
    }
    if (l == 0) {
      if (full.empty()) HIPCHK(hipEventRecord(ctx->ev[4], st));
      // stream bytes written by level 0 (counter 3 excludes workgroup slack)
      HIPCHK(hipMemcpyAsync(l0_bytes, B.counters + 3, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    dev_free(d_lists);
    if (level_grid) dev_free(level_grid);
    level_grid = parents;
    level_in = parents;
    for (int d = 0; d < 3; ++d) gd[d] = pd[d];
  }
  HIPCHK(hipStreamSynchronize(st));
  if (level_grid) dev_free(level_grid);
  dev_free(scratch);
  HIPCHK(hipMemcpy(flags, B.counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  return rc;
}

static int assemble_level(idx2g_ctx* ctx, idx2g_dataset* ds, int l, EncodeBuffers& B) {
  const Layout& L = ds->L;
  LevelData& lv = ds->lv[l];
  hipStream_t st = ctx->stream;
  const uint64_t span = L.linear_span(l);
  // key range of the level
  int* d_kr;
  TRY(dalloc(&d_kr, 2));
  int init[2] = {0x7fffffff, -0x7fffffff};
  HIPCHK(hipMemcpyAsync(d_kr, init, sizeof init, hipMemcpyHostToDevice, st));
  const int nh = lv.nb * L.nsb;
  hipLaunchKernelGGL(k_key_range, dim3(std::min((nh + 255) / 256, 1024)), dim3(256), 0, st, B.hdrs[l], nh, d_kr);
  int kr[2];
  HIPCHK(hipMemcpyAsync(kr, d_kr, sizeof kr, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  dev_free(d_kr);
  if (kr[1] < kr[0]) {  // nothing coded on this level
    lv.klo = 0; lv.nk = 0;
    TRY(dalloc(&lv.d_blob, kBlobPad));
    TRY(dalloc(&lv.d_dir, 1));
    return IDX2G_OK;
  }
  lv.klo = kr[0];
  lv.nk = kr[1] - kr[0] + 1;
  const size_t rows = (size_t)L.nsb * lv.nk;
  uint32_t* d_idx;
  TRY(dalloc(&d_idx, rows * span));
  HIPCHK(hipMemsetAsync(d_idx, 0, rows * span * 4, st));
  uint64_t* d_lin;
  TRY(dalloc(&d_lin, lv.nb));
  HIPCHK(hipMemcpyAsync(d_lin, lv.lin.data(), lv.nb * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_fill_index, dim3((nh + 255) / 256), dim3(256), 0, st, B.hdrs[l], d_lin, lv.nb, L.nsb, lv.klo,
                     lv.nk, span, d_idx);
  const int bpc = L.bpc[l];
  const int nchunks = (int)((span + bpc - 1) / bpc);
  const size_t nrec = rows * nchunks;
  ChunkRec* d_recs;
  TRY(dalloc(&d_recs, nrec));
  hipLaunchKernelGGL(k_chunk_size, dim3((unsigned)nrec), dim3(256), 0, st, d_idx, B.entries, lv.nk, span, bpc, nchunks,
                     d_recs, lv.klo);
  std::vector<ChunkRec> recs(nrec);
  HIPCHK(hipMemcpyAsync(recs.data(), d_recs, nrec * sizeof(ChunkRec), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // blob layout: non-empty chunks in (subband, key, chunk) order
  std::vector<int> order;
  uint64_t off = 0;
  for (size_t i = 0; i < nrec; ++i)
    if (recs[i].size) {
      recs[i].off = off;
      off += recs[i].size;
      order.push_back((int)i);
    }
  lv.blob_bytes = off;
  TRY(dalloc(&lv.d_blob, off + kBlobPad));
  HIPCHK(hipMemsetAsync(lv.d_blob + off, 0, kBlobPad, st));
  TRY(dalloc(&lv.d_dir, rows * span));
  HIPCHK(hipMemsetAsync(lv.d_dir, 0, rows * span * 8, st));
  int* d_order;
  TRY(dalloc(&d_order, order.size()));
  HIPCHK(hipMemcpyAsync(d_recs, recs.data(), nrec * sizeof(ChunkRec), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, st));
  if (!order.empty())
    hipLaunchKernelGGL(k_chunk_write, dim3((unsigned)order.size()), dim3(256), 0, st, d_idx, B.entries, B.arena, lv.nk,
                       lv.klo, span, bpc, d_recs, d_order, lv.d_blob, lv.d_dir);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  lv.chunks.clear();
  for (int i : order) lv.chunks.push_back(recs[i]);
  dev_free(d_idx);
  dev_free(d_lin);
  dev_free(d_recs);
  dev_free(d_order);
  return IDX2G_OK;
}

static int encode_impl(idx2g_ctx* ctx, const idx2g_encode_params* p, const void* input, int input_is_device,
                       idx2g_dataset** out) {
  if (!ctx || !p || !input || !out) return IDX2G_SIZE_ZERO;
  HIPCHK(hipSetDevice(ctx->device));
  idx2g_dataset* ds = new idx2g_dataset();
  ds->device = ctx->device;
  Layout& L = ds->L;
  L.name = p->name ? p->name : "data";
  L.field = p->field ? p->field : "field";
  for (int d = 0; d < 3; ++d) {
    L.dims[d] = p->dims[d];
    L.brick[d] = p->brick[d] ? p->brick[d] : (p->dims[d] == 1 ? 1 : 32);
  }
  L.dtype = p->dtype;
  if (L.dtype != 0 && L.dtype != 1) { delete ds; return IDX2G_TYPE_NOT_SUPPORTED; }
  L.nlevels = p->nlevels;
  L.tolerance = p->tolerance > 0 ? p->tolerance
                                 : (L.dtype ? 2.220446049250313e-16 : 1.1920928955078125e-07);  // GetMachineEpsilon
  L.bpc_in = p->bricks_per_chunk ? p->bricks_per_chunk : 4096;
  L.cpf_in = p->chunks_per_file ? p->chunks_per_file : 64;
  L.files_per_dir = p->files_per_dir ? p->files_per_dir : 64;
  L.bppc = p->bit_planes_per_chunk ? p->bit_planes_per_chunk : 1;
  int rc = L.finalize();
  if (!rc) rc = check_layout(L);
  if (rc) { delete ds; return rc; }
  level_lists(ds);
  exp_layout(ds);
  rc = upload_block_tables(ds);
  if (rc) { delete ds; return rc; }
  const size_t raw = (size_t)L.dims[0] * L.dims[1] * L.dims[2] * (L.dtype ? 8 : 4);
  const void* d_in = input;
  void* d_tmp = nullptr;
  if (!input_is_device) {
    HIPCHK(dev_alloc(&d_tmp, raw));
    HIPCHK(hipMemcpy(d_tmp, input, raw, hipMemcpyHostToDevice));
    d_in = d_tmp;
  }
  HIPCHK(dev_alloc((void**)&ds->d_minmax, (size_t)ds->lv[0].nb * 2 * sizeof(double)));
  ds->have_range = true;
  int nbt = 0;
  for (int l = 0; l < L.nlevels; ++l) nbt += ds->lv[l].nb;
  EncodeBuffers B;
  // workgroups take arena/entry slices of b32::kArenaChunk / kEntryChunk:
  // leave room for one partly used slice per workgroup and level
  const unsigned long long slack = (unsigned long long)ctx->ncu * L.nlevels;
  B.arena_cap = std::max<unsigned long long>(64ull << 20, (unsigned long long)(raw * 1.5)) + slack * b32::kArenaChunk;
  B.entry_cap = (unsigned long long)nbt * L.nsb * 48 + 4096 + slack * b32::kEntryChunk;
  hipStream_t st = ctx->stream;
  idx2g_stats stats{};
  for (int attempt = 0; attempt < 4; ++attempt) {
    TRY(dalloc(&B.arena, B.arena_cap + 16));
    TRY(dalloc(&B.entries, B.entry_cap));
    TRY(dalloc(&B.counters, 4));
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    unsigned long long flags[3], l0 = 0;
    int64_t pbytes = 0;
    rc = run_encode_levels(ctx, ds, d_in, B, flags, &l0, &pbytes);
    stats.dominant_bytes = (int64_t)(raw + l0) + pbytes;
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    if (rc) break;
    if (flags[2] & 1) {
      set_error("a brick subband needs more than %d bit-plane keys", kMaxKeys);
      rc = IDX2G_NOT_SUPPORTED;
      break;
    }
    if (flags[2] & 2) {  // arena or entry table too small: grow and redo
      B.release();
      for (int l = 0; l < L.nlevels; ++l) free_level(ds->lv[l]);
      level_lists(ds);
      exp_layout(ds);
      B.arena_cap = std::max(B.arena_cap * 2, flags[0] + (1ull << 20));
      B.entry_cap = std::max(B.entry_cap * 2, flags[1] + 4096);
      continue;
    }
    stats.chunk_bytes = 0;
    for (int l = 0; l < L.nlevels && !rc; ++l) rc = assemble_level(ctx, ds, l, B);
    HIPCHK(hipEventRecord(ctx->ev[2], st));
    HIPCHK(hipEventSynchronize(ctx->ev[2]));
    break;
  }
  B.release();
  if (d_tmp) dev_free(d_tmp);
  if (rc) {
    for (int l = 0; l < L.nlevels; ++l) free_level(ds->lv[l]);
    delete ds;
    return rc;
  }
  float t01 = 0, t12 = 0, t02 = 0;
  (void)hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]);
  (void)hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]);
  (void)hipEventElapsedTime(&t02, ctx->ev[0], ctx->ev[2]);
  float td = 0;
  (void)hipEventElapsedTime(&td, ctx->ev[3], ctx->ev[4]);
  stats.dominant_ms = td;
  stats.dominant_launches = 1;
  stats.transform_ms = t01;
  stats.assemble_ms = t12;
  stats.total_ms = t02;
  stats.raw_bytes = (int64_t)raw;
  uint64_t cb = 0, eb = 0;
  for (int l = 0; l < L.nlevels; ++l) {
    cb += ds->lv[l].blob_bytes;
    eb += ds->lv[l].exp_bytes;
  }
  ds->chunk_bytes = cb;
  stats.chunk_bytes = (int64_t)cb;
  stats.exp_bytes = (int64_t)eb;
  ctx->last = stats;
  *out = ds;
  return IDX2G_OK;
}

}  // namespace idx2g

#include "files.inc"
#include "decode.inc"
#include "api.inc"
#include "server.inc"
