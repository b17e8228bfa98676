// Shared host helpers: error reporting, constant tables, zstd binding.
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "host.hpp"

namespace idx2g {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, int line) {
  set_error("HIP error %d (%s) in %s at line %d", (int)e, hipGetErrorString(e), what, line);
  return IDX2G_DEVICE_ERROR;
}

BrickConst make_brick_const(const Layout& L) {
  BrickConst k;
  int mx = 1;
  for (int d = 0; d < 3; ++d) {
    k.N[d] = L.ext[d];
    k.M[d] = L.ext[d] == 1 ? 1 : L.ext[d] - 1;
    mx = std::max(mx, L.ext[d]);
  }
  k.nsb = L.nsb;
  for (int s = 0; s < 8; ++s)
    for (int d = 0; d < 3; ++d) {
      k.sb[s].f[d] = s < L.nsb ? L.sb[s].from[d] : 0;
      k.sb[s].d[d] = s < L.nsb ? L.sb[s].dims[d] : 1;
      k.sb[s].s[d] = s < L.nsb ? L.sb[s].strd[d] : 1;
    }
  k.ext_levels = msb64((uint64_t)mx);
  k.order = L.tform_order;
  k.dtype = L.dtype;
  k.endbp = L.dtype ? 64 : 32;
  k.nvox = L.ext[0] * L.ext[1] * L.ext[2];
  return k;
}

// Coefficient order of a 4x4x4 block: by i+j+k, then i^2+j^2+k^2, ties as
// in zfp's published perm_3 table (Source/Core/Zfp.h:414-493); 4x4 order of
// Zfp.h:497-521.  Entries (i,j,k) with i along x.
static const signed char kPerm3[64][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
    {1, 1, 1}, {2, 1, 0}, {2, 0, 1}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {0, 1, 2}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3},
    {2, 1, 1}, {1, 2, 1}, {1, 1, 2}, {0, 2, 2}, {2, 0, 2}, {2, 2, 0}, {3, 1, 0}, {3, 0, 1}, {0, 3, 1}, {1, 3, 0},
    {1, 0, 3}, {0, 1, 3}, {1, 2, 2}, {2, 1, 2}, {2, 2, 1}, {3, 1, 1}, {1, 3, 1}, {1, 1, 3}, {3, 2, 0}, {3, 0, 2},
    {0, 3, 2}, {2, 3, 0}, {2, 0, 3}, {0, 2, 3}, {2, 2, 2}, {3, 2, 1}, {3, 1, 2}, {1, 3, 2}, {2, 3, 1}, {2, 1, 3},
    {1, 2, 3}, {0, 3, 3}, {3, 0, 3}, {3, 3, 0}, {3, 2, 2}, {2, 3, 2}, {2, 2, 3}, {1, 3, 3}, {3, 1, 3}, {3, 3, 1},
    {2, 3, 3}, {3, 2, 3}, {3, 3, 2}, {3, 3, 3}};
static const signed char kPerm2[16][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}, {2, 0}, {0, 2}, {2, 1}, {1, 2},
                                          {3, 0}, {0, 3}, {2, 2}, {3, 1}, {1, 3}, {3, 2}, {2, 3}, {3, 3}};

PermTables make_perm_tables() {
  PermTables t;
  for (int i = 0; i < 64; ++i) {
    int j = kPerm3[i][0] + 4 * kPerm3[i][1] + 16 * kPerm3[i][2];
    t.p3[i] = (unsigned char)j;
    t.ip3[j] = (unsigned char)i;
  }
  for (int i = 0; i < 16; ++i) {
    int j = kPerm2[i][0] + 4 * kPerm2[i][1];
    t.p2[i] = (unsigned char)j;
    t.ip2[j] = (unsigned char)i;
  }
  return t;
}

namespace {
struct DevCache {
  std::mutex mu;
  std::map<int, std::multimap<size_t, void*>> free_;      // device -> size -> block
  std::unordered_map<void*, std::pair<int, size_t>> live;  // block -> (device, size)
  std::map<int, size_t> cached;                            // device -> bytes held free
};
// Give free blocks of dev back to the runtime, largest first, until at most
// `keep` bytes stay cached (an allocation the runtime refused: only as much
// goes back as it needs, so the rest of the cache stays warm).  Caller holds
// c.mu.
void trim(DevCache& c, int dev, size_t keep) {
  auto& fl = c.free_[dev];
  size_t& held = c.cached[dev];
  while (held > keep && !fl.empty()) {
    auto it = std::prev(fl.end());
    (void)hipFree(it->second);
    held -= it->first;
    fl.erase(it);
  }
}
DevCache& dev_cache() {
  static DevCache* c = new DevCache();  // never destroyed: blocks live until the process ends
  return *c;
}
size_t round_bytes(size_t n) {
  if (n < (1u << 20)) return (n + 255) & ~size_t(255);
  return (n + (2u << 20) - 1) & ~size_t((2u << 20) - 1);
}
}  // namespace

namespace {
thread_local ScratchArena* tl_arena = nullptr;

hipError_t arena_alloc(ScratchArena& A, void** p, size_t bytes) {
  const size_t n = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
  for (auto it = A.free_.begin(); it != A.free_.end(); ++it) {
    if (it->second < n) continue;
    const size_t off = it->first, left = it->second - n;
    A.free_.erase(it);
    if (left) A.free_[off + n] = left;
    A.used[off] = n;
    A.in_use += n;
    A.peak = std::max(A.peak, A.in_use);
    *p = A.base + off;
    return hipSuccess;
  }
  *p = nullptr;
  A.exhausted = true;
  return hipErrorOutOfMemory;
}

bool arena_free(ScratchArena& A, void* p) {
  const uint8_t* q = static_cast<const uint8_t*>(p);
  if (q < A.base || q >= A.base + A.cap) return false;
  auto it = A.used.find((size_t)(q - A.base));
  if (it == A.used.end()) return true;
  size_t off = it->first, n = it->second;
  A.used.erase(it);
  A.in_use -= n;
  auto nx = A.free_.lower_bound(off);
  if (nx != A.free_.end() && nx->first == off + n) {  // merge with the next free block
    n += nx->second;
    nx = A.free_.erase(nx);
  }
  if (nx != A.free_.begin()) {  // and with the previous one
    auto pv = std::prev(nx);
    if (pv->first + pv->second == off) {
      pv->second += n;
      return true;
    }
  }
  A.free_[off] = n;
  return true;
}
}  // namespace

void set_thread_arena(ScratchArena* a) { tl_arena = a; }

uint64_t next_dataset_uid() {
  static std::atomic<uint64_t> next{1};
  return next.fetch_add(1);
}

hipError_t dev_alloc(void** p, size_t bytes) {
  if (tl_arena) return arena_alloc(*tl_arena, p, bytes);
  DevCache& c = dev_cache();
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const size_t n = round_bytes(bytes ? bytes : 1);
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto& fl = c.free_[dev];
    auto it = fl.lower_bound(n);
    if (it != fl.end() && it->first <= n + n / 2 + (2u << 20)) {
      *p = it->second;
      c.live[*p] = {dev, it->first};
      c.cached[dev] -= it->first;
      fl.erase(it);
      return hipSuccess;
    }
  }
  e = hipMalloc(p, n);
  while (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    (void)hipGetLastError();
    std::lock_guard<std::mutex> g(c.mu);  // give cached blocks back, largest first, and retry
    const size_t held = c.cached[dev];
    if (!held) break;
    trim(c, dev, held > n ? held - n : 0);
    e = hipMalloc(p, n);
  }
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(c.mu);
  c.live[*p] = {dev, n};
  return hipSuccess;
}

void dev_free(void* p) {
  if (!p) return;
  if (tl_arena && arena_free(*tl_arena, p)) return;
  DevCache& c = dev_cache();
  std::lock_guard<std::mutex> g(c.mu);
  auto it = c.live.find(p);
  if (it == c.live.end()) { (void)hipFree(p); return; }
  const int dev = it->second.first;
  c.free_[dev].emplace(it->second.second, p);
  c.cached[dev] += it->second.second;
  c.live.erase(it);
}

size_t dev_cache_release(int device) {
  DevCache& c = dev_cache();
  std::lock_guard<std::mutex> g(c.mu);
  const size_t held = c.cached[device];
  trim(c, device, 0);
  return held - c.cached[device];
}

// Pinned host blocks kept for reuse (file writing stages whole chunk blobs
// through them; pinning hundreds of MB per call would cost more than the
// copy).  Blocks are returned by pinned_free and reused by the next request
// of at most 1.5x their size.
namespace {
struct PinCache {
  std::mutex mu;
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> live;
};
PinCache& pin_cache() {
  static PinCache* c = new PinCache();
  return *c;
}
}  // namespace

hipError_t pinned_alloc(void** p, size_t bytes) {
  PinCache& c = pin_cache();
  const size_t n = round_bytes(bytes ? bytes : 1);
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.free_.lower_bound(n);
    if (it != c.free_.end() && it->first <= n + n / 2) {
      *p = it->second;
      c.live[*p] = it->first;
      c.free_.erase(it);
      return hipSuccess;
    }
  }
  hipError_t e = hipHostMalloc(p, n, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    std::lock_guard<std::mutex> g(c.mu);  // give the cached blocks back and retry
    for (auto& kv : c.free_) (void)hipHostFree(kv.second);
    c.free_.clear();
    e = hipHostMalloc(p, n, hipHostMallocDefault);
  }
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(c.mu);
  c.live[*p] = n;
  return hipSuccess;
}

void pinned_free(void* p) {
  if (!p) return;
  PinCache& c = pin_cache();
  std::lock_guard<std::mutex> g(c.mu);
  auto it = c.live.find(p);
  if (it == c.live.end()) { (void)hipHostFree(p); return; }
  c.free_.emplace(it->second, p);
  c.live.erase(it);
}

void build_block_tables(idx2g_dataset* ds, std::vector<uint32_t>* t32, std::vector<uint64_t>* t64) {
  const Layout& L = ds->L;
  const bool small = L.brick[0] <= 32 && L.brick[1] <= 32 && L.brick[2] <= 32;
  t32->clear();
  t64->clear();
  for (int l = 0; l < L.nlevels; ++l)
    for (int s = 0; s < L.nsb; ++s) {
      ds->blk_off[l][s] = (int)t64->size();
      int ci = 0;
      for (const BlockInfo& b : L.blocks[l][s]) {
        const int c = b.skip ? 0 : ci;
        t32->push_back(small ? pack_block(b.org, b.dims, b.skip, c) : 0u);
        t64->push_back(pack_block64(b.org, b.dims, b.skip, c));
        if (!b.skip) ++ci;
      }
    }
}

int upload_block_tables(idx2g_dataset* ds) {
  std::vector<uint32_t> t32;
  std::vector<uint64_t> t64;
  build_block_tables(ds, &t32, &t64);
  HIPCHK(dev_alloc((void**)&ds->d_blocks, t32.size() * 4 + 4));
  HIPCHK(hipMemcpy(ds->d_blocks, t32.data(), t32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(dev_alloc((void**)&ds->d_blocks64, t64.size() * 8 + 8));
  HIPCHK(hipMemcpy(ds->d_blocks64, t64.data(), t64.size() * 8, hipMemcpyHostToDevice));
  return IDX2G_OK;
}

// --- zstd (libzstd.so.1, public API of zstd.h) ---
typedef size_t (*zbound_t)(size_t);
typedef size_t (*zcomp_t)(void*, size_t, const void*, size_t, int);
typedef unsigned long long (*zsize_t)(const void*, size_t);
typedef size_t (*zdec_t)(void*, size_t, const void*, size_t);
typedef unsigned (*zerr_t)(size_t);
static zbound_t z_bound;
static zcomp_t z_comp;
static zsize_t z_size;
static zdec_t z_dec;
static zerr_t z_err;
static std::once_flag z_once;
static bool z_ok = false;

static void z_load() {
  void* h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
  if (!h) return;
  z_bound = (zbound_t)dlsym(h, "ZSTD_compressBound");
  z_comp = (zcomp_t)dlsym(h, "ZSTD_compress");
  z_size = (zsize_t)dlsym(h, "ZSTD_getFrameContentSize");
  z_dec = (zdec_t)dlsym(h, "ZSTD_decompress");
  z_err = (zerr_t)dlsym(h, "ZSTD_isError");
  z_ok = z_bound && z_comp && z_size && z_dec && z_err;
}

// CompressBufZstd, Source/Core/idx2Encode.cpp:23-37 (level 1)
int zstd_compress(const void* src, size_t n, std::vector<uint8_t>* out) {
  std::call_once(z_once, z_load);
  out->clear();
  if (!z_ok) { set_error("libzstd.so.1 not available"); return IDX2G_NOT_SUPPORTED; }
  if (n == 0) return IDX2G_OK;
  out->resize(z_bound(n));
  size_t c = z_comp(out->data(), out->size(), src, n, 1);
  if (z_err(c)) { set_error("zstd compression failed"); return IDX2G_UNKNOWN; }
  out->resize(c);
  return IDX2G_OK;
}

// DecompressBufZstd, Source/Core/idx2Decode.cpp:53-78
int zstd_decompress(const void* src, size_t n, std::vector<uint8_t>* out) {
  std::call_once(z_once, z_load);
  if (!z_ok) { set_error("libzstd.so.1 not available"); return IDX2G_NOT_SUPPORTED; }
  unsigned long long sz = z_size(src, n);
  if (sz == (unsigned long long)-1 || sz == (unsigned long long)-2) { set_error("bad zstd frame"); return IDX2G_PARSE_FAILED; }
  out->resize(sz);
  size_t r = z_dec(out->data(), sz, src, n);
  if (z_err(r) || r != sz) { set_error("zstd decompression failed"); return IDX2G_PARSE_FAILED; }
  return IDX2G_OK;
}

}  // namespace idx2g

extern "C" const char* idx2g_last_error_message(void) { return idx2g::g_err; }
