// idx2g driver: encode/decode orchestration on one GPU and the C ABI.
// Kernels live in encode_kernels.hip / decode_kernels.hip (same translation
// unit, so the launches below see them directly).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <fcntl.h>
#include <list>
#include <mutex>
#include <map>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <tuple>
#include <vector>

#include "decode_kernels.hip"
#include "encode_kernels.hip"
#include "host.hpp"

using namespace idx2g;

namespace idx2g {

static const size_t kBlobPad = 64;
// Chunk bytes read from files get a larger zeroed tail: a corrupted stream
// can make a walk lane read past the stream's end, by at most one plane per
// block (<= 193 bits: 64 verbatim, 64 positions, 65 group tests) for the
// <= 4913 blocks of a subband (bricks of up to 128 samples per axis) plus its
// 24-byte window: 118.6 KB.
static const size_t kFileBlobPad = size_t(128) << 10;
// Scratch slots of a generic-kernel launch whose per-subband tables live in
// the slots (bricks above 32 samples per axis) are held to this many bytes:
// fewer resident workgroups, each walking more bricks.
static const size_t kGenericScratchBytes = size_t(4) << 30;

template <class T>
static int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(dev_alloc((void**)p, n * sizeof(T)));
  return IDX2G_OK;
}
#define TRY(x)                  \
  do {                          \
    int rc_ = (x);              \
    if (rc_) return rc_;        \
  } while (0)

static inline int hip_ok(hipError_t e, const char* what, int line) {
  return e == hipSuccess ? IDX2G_OK : hip_fail(e, what, line);
}

static void free_level(LevelData& lv) {
  if (lv.d_blob) dev_free(lv.d_blob);
  if (lv.d_dir) dev_free(lv.d_dir);
  if (lv.d_exps) dev_free(lv.d_exps);
  if (lv.d_xyz) dev_free(lv.d_xyz);
  if (lv.d_lin) dev_free(lv.d_lin);
  lv = LevelData();
}

static int check_layout(const Layout& L) {
  // generic block words hold a block's origin in 8 bits per axis and the
  // generic kernels' tables are sized for 17^3 blocks per subband: brick
  // extents of at most 128 samples
  for (int d = 0; d < 3; ++d)
    if (L.brick[d] > kMaxBrickExtent) {
      set_error("brick extent %d above %d is not supported", L.brick[d], kMaxBrickExtent);
      return IDX2G_NOT_SUPPORTED;
    }
  for (int l = 0; l < L.nlevels; ++l)
    for (int s = 0; s < L.nsb; ++s)
      if ((int)L.blocks[l][s].size() > kMaxGenericBlocks) {
        set_error("brick size gives %zu blocks per subband (max %d)", L.blocks[l][s].size(), kMaxGenericBlocks);
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

// The levels' sorted brick ids on the device, for the decode's query planner.
static int upload_level_ids(idx2g_dataset* ds) {
  for (int l = 0; l < ds->L.nlevels; ++l) {
    LevelData& lv = ds->lv[l];
    if (lv.d_lin) continue;
    TRY(dalloc(&lv.d_lin, lv.lin.size()));
    if (!lv.lin.empty()) HIPCHK(hipMemcpy(lv.d_lin, lv.lin.data(), lv.lin.size() * 8, hipMemcpyHostToDevice));
  }
  return IDX2G_OK;
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

// Bricks of a level that the on-chip kernel takes: 3D, brick 32^3, data
// extent 32 on every axis.
static bool full_b32(const Layout& L, const int* b3, const int gd[3]) {
  // the kernel addresses a brick with 32-bit offsets from its origin
  if ((uint64_t)gd[0] * (uint64_t)gd[1] * 33ull >= (1ull << 32)) return false;
  for (int d = 0; d < 3; ++d)
    if (L.brick[d] != 32 || (b3[d] + 1) * 32 > gd[d]) return false;
  return true;
}

// 2D bricks the on-chip 2D kernel takes: brick 32x32x1 of a 2D grid, data
// extent 32 on both axes.
static bool full_2d32(const Layout& L, const int* b3, const int gd[3]) {
  if (L.brick[0] != 32 || L.brick[1] != 32 || L.brick[2] != 1 || gd[2] != 1) return false;
  for (int d = 0; d < 2; ++d)
    if ((b3[d] + 1) * 32 > gd[d]) return false;
  for (int s = 0; s < L.nsb; ++s)  // the group coder holds kGroup bricks' blocks (<= 125)
    for (int l = 0; l < L.nlevels; ++l)
      if ((int)L.blocks[l][s].size() * b2d::kGroup > kMaxBlocks) return false;
  return true;
}

// Per axis, the brick coordinate below which a brick of a level whose grid is
// gd is full: the bricks full_b32 / full_2d32 accept are exactly those below
// F on every axis.  All zero when the layout takes no on-chip path.
static void full_extent(const Layout& L, const int gd[3], int F[3]) {
  const int origin[3] = {0, 0, 0};
  F[0] = F[1] = F[2] = 0;
  if (full_b32(L, origin, gd)) {
    for (int d = 0; d < 3; ++d) F[d] = gd[d] / 32;
  } else if (full_2d32(L, origin, gd)) {
    F[0] = gd[0] / 32;
    F[1] = gd[1] / 32;
    F[2] = 1;
  }
}

}  // namespace idx2g

namespace idx2g {
static int noalloc_setup(int device);  // noalloc.inc
}

#include "encode.inc"
#include "files.inc"
#include "decode.inc"
#include "api.inc"
#include "server.inc"
#include "tune.inc"
#include "pool.inc"
#include "noalloc.inc"

#ifdef IDX2G_PHASE_TIMERS
// Profiling builds: the accumulated phase ticks of k_encode_b32 (10 ns each,
// summed over workgroups) since the previous call, which clears them.
extern "C" int idx2g_phase_times(unsigned long long* ticks32) {
  HIPCHK(hipMemcpyFromSymbol(ticks32, HIP_SYMBOL(idx2g::g_phase_ticks), 32 * sizeof(unsigned long long)));
  unsigned long long z[32] = {0};
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(idx2g::g_phase_ticks), z, sizeof(z)));
  return IDX2G_OK;
}
#endif
