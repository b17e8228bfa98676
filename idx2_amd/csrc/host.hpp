// Host-side objects of the idx2g library: context, resident dataset.
#pragma once

#include <atomic>
#include <map>

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/idx2g.h"
#include "kernels.hpp"
#include "layout.hpp"

namespace idx2g {

struct LevelData {
  int nb = 0;                  // existing bricks (dense, linear-id order)
  std::vector<uint64_t> lin;
  std::vector<int> xyz;
  int klo = 0, nk = 0;         // bit-plane key range of the level
  std::vector<ChunkRec> chunks;  // non-empty chunks, blob offsets filled
  uint8_t* d_blob = nullptr;   // chunk bytes (+ padding)
  uint64_t blob_bytes = 0;
  uint64_t* d_dir = nullptr;   // [nsb][nk][span] stream byte offset + 1
  uint8_t* d_exps = nullptr;   // per subband [dense][ncoded] exponent bytes
  long long exp_off[8] = {0};
  uint64_t exp_bytes = 0;
  int* d_xyz = nullptr;
  uint64_t* d_lin = nullptr;   // lin on the device (the query planner's rank lookups)
  // files of this level absent when the dataset was opened (sorted file ids):
  // a decode that needs one of their bricks fails with FileNotFound, as the
  // reference's does when it opens the file (idx2Read.cpp:269-270)
  std::vector<uint64_t> missing_files;
  std::string missing_path;  // one of them, for the message
};

struct Context;

}  // namespace idx2g

namespace idx2g {
// pinned launch-table staging a context allocates when it is created
constexpr size_t kStageInitBytes = size_t(4) << 20;
}

struct idx2g_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  idx2g_stats last{};
  int ncu = 256;
  // timing events
  hipEvent_t ev[8] = {};
  // pinned staging of batched uploads (ev[6] marks the last transfer out of it)
  uint8_t* h_stage = nullptr;
  size_t stage_cap = 0;
  bool stage_pending = false;
  // set while idx2g_decode_noalloc runs: device buffers come from the
  // device's fixed scratch array, uploads must fit the staging buffer as it
  // is, launch shapes are capped to fit the array
  bool noalloc = false;
};

namespace idx2g {
uint64_t next_dataset_uid();  // process-unique id (cache keys never confuse two datasets)
}

struct idx2g_dataset {
  idx2g::Layout L;
  uint64_t uid = idx2g::next_dataset_uid();
  int device = 0;
  idx2g::LevelData lv[idx2g::kMaxLevels];
  uint32_t* d_blocks = nullptr;  // packed block tables (pack_block: the on-chip kernels)
  uint64_t* d_blocks64 = nullptr;  // the same tables as pack_block64 words (the generic kernels)
  int blk_off[idx2g::kMaxLevels][8];
  double* d_minmax = nullptr;    // per level-0 brick min/max (encode only)
  bool have_range = false;
  uint64_t chunk_bytes = 0;
  // serving requests queued or running on this dataset; idx2g_dataset_free
  // while some are in flight only marks it, and the last one frees it
  std::atomic<int> inflight{0};
  std::atomic<bool> free_pending{false}, freed{false};
};

namespace idx2g {
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, int line);

// Launch shapes of the current device (set by the autotuner or the caller).
idx2g_tuning tuning_of(int device);

// Device buffers of the calling thread carved from a fixed device region
// instead of hipMalloc (the no-allocation decode): first fit in 256-byte
// units, blocks handed back by dev_free and merged with free neighbours;
// when nothing fits, dev_alloc returns hipErrorOutOfMemory.
struct ScratchArena {
  uint8_t* base = nullptr;
  size_t cap = 0, in_use = 0, peak = 0;
  bool exhausted = false;          // an allocation found no room
  std::map<size_t, size_t> free_;  // offset -> bytes
  std::map<size_t, size_t> used;   // offset -> bytes
  void reset() {
    free_.clear();
    used.clear();
    if (cap) free_[0] = cap;
    in_use = peak = 0;
    exhausted = false;
  }
};
// While set (non-null), dev_alloc / dev_free of this thread use the arena.
void set_thread_arena(ScratchArena* a);

// Device memory through a per-device cache of freed blocks: repeated
// encode/decode calls reuse their large buffers instead of paying
// hipMalloc/hipFree each time.  Blocks are only handed back after the
// library call that freed them has synchronised its stream.
hipError_t dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
// Hands every kept free block of `device` back to the runtime; returns the
// bytes released.
size_t dev_cache_release(int device);
// Pinned host memory through a process-wide cache of freed blocks.
hipError_t pinned_alloc(void** p, size_t bytes);
void pinned_free(void* p);
#define HIPCHK(x)                                                  \
  do {                                                             \
    hipError_t e_ = (x);                                           \
    if (e_ != hipSuccess) return ::idx2g::hip_fail(e_, #x, __LINE__); \
  } while (0)

BrickConst make_brick_const(const Layout& L);
PermTables make_perm_tables();
int upload_block_tables(idx2g_dataset* ds);
// The dataset's block tables, every level and subband (offsets into
// ds->blk_off): 32-bit words for the on-chip kernels (zeros when a brick
// extent is above 32, where those kernels never run) and 64-bit words for
// the generic kernels.
void build_block_tables(idx2g_dataset* ds, std::vector<uint32_t>* t32, std::vector<uint64_t>* t64);

// zstd through the system libzstd.so.1
int zstd_compress(const void* src, size_t n, std::vector<uint8_t>* out);
int zstd_decompress(const void* src, size_t n, std::vector<uint8_t>* out);
}  // namespace idx2g
