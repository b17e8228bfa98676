/* idx2g: MI355X (gfx950) implementation of the idx2 encode/decode path.
 *
 * C ABI: plain pointers and sizes, no C++ or torch types.  Every call returns
 * an idx2 error code; 0 is success.  No call throws.
 *
 * Error numbering: the values of the reference enum idx2_err_code
 * (Source/Core/idx2Common.h:36-56), whose first 21 members are the common
 * codes of Source/Core/ErrorCodes.h:6-12.  The reference's effective
 * idx2_Enum (Source/Core/Enum.h:65-80, which replaces the definition at
 * Enum.h:14-26) declares an enum class with __Invalid__ last, so NoError is 0
 * and every code here equals the reference's enum value (checked at compile
 * time by include/idx2g_idx2.hpp against the reference's headers).
 * IDX2G_DEVICE_ERROR (100) has no reference counterpart (a HIP runtime
 * failure; the reference runs on the CPU).
 *
 * A context owns one HIP stream and the device pools on one GPU.  Contexts are
 * not thread safe; use one per host thread.  Datasets own their device memory.
 */
#ifndef IDX2G_H
#define IDX2G_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum idx2g_err {
  IDX2G_OK = 0,                       /* NoError */
  IDX2G_UNKNOWN = 1,                  /* UnknownError */
  IDX2G_SIZE_ZERO = 2,
  IDX2G_SIZE_TOO_SMALL = 3,
  IDX2G_SIZE_MISMATCHED = 4,
  IDX2G_DIMENSION_MISMATCHED = 5,
  IDX2G_DIMENSIONS_TOO_MANY = 6,
  IDX2G_ATTRIBUTE_NOT_FOUND = 7,
  IDX2G_OPTION_NOT_SUPPORTED = 8,
  IDX2G_TYPE_NOT_SUPPORTED = 9,
  IDX2G_FILE_CREATE_FAILED = 10,
  IDX2G_FILE_READ_FAILED = 11,
  IDX2G_FILE_WRITE_FAILED = 12,
  IDX2G_FILE_OPEN_FAILED = 13,
  IDX2G_FILE_CLOSE_FAILED = 14,
  IDX2G_FILE_SEEK_FAILED = 15,
  IDX2G_FILE_TELL_FAILED = 16,
  IDX2G_PARSE_FAILED = 17,
  IDX2G_OUT_OF_MEMORY = 18,
  IDX2G_DIMENSIONS_REPEATED = 19,
  IDX2G_DIMENSION_TOO_SMALL = 20,
  IDX2G_BRICK_SIZE_NOT_POW2 = 21,
  IDX2G_BRICK_SIZE_TOO_BIG = 22,
  IDX2G_TOO_MANY_LEVELS = 23,
  IDX2G_TOO_MANY_TRANSFORM_PASSES_PER_LEVEL = 24,
  IDX2G_TOO_MANY_LEVELS_OR_TRANSFORM_PASSES = 25,
  IDX2G_TOO_MANY_BRICKS_PER_FILE = 26,
  IDX2G_TOO_MANY_FILES_PER_DIR = 27,
  IDX2G_NOT_SUPPORTED = 28,           /* NotSupportedInVersion */
  IDX2G_CANNOT_CREATE_DIRECTORY = 29,
  IDX2G_SYNTAX_ERROR = 30,
  IDX2G_TOO_MANY_BRICKS_PER_CHUNK = 31,
  IDX2G_TOO_MANY_CHUNKS_PER_FILE = 32,
  IDX2G_CHUNKS_PER_FILE_NOT_POW2 = 33,
  IDX2G_BRICKS_PER_CHUNK_NOT_POW2 = 34,
  IDX2G_CHUNK_NOT_FOUND = 35,
  IDX2G_BRICK_NOT_FOUND = 36,
  IDX2G_FILE_NOT_FOUND = 37,
  IDX2G_UNSUPPORTED_SCHEME = 38,
  IDX2G_DEVICE_ERROR = 100, /* HIP runtime failure; see idx2g_last_error_message */
};

typedef struct idx2g_ctx idx2g_ctx;
typedef struct idx2g_dataset idx2g_dataset;

/* Encode parameters: the fields of idx2::params used by encode
 * (Source/Core/idx2Common.h:74-108, parsed by Applications/idx2App.cpp:114-168).
 * Zero/NULL fields take the reference defaults (brick 32^3, 4096 bricks per
 * chunk, 64 chunks per file, 64 files per directory, 1 bit plane per chunk,
 * tolerance 0 = machine epsilon of the dtype, nlevels 0 = guessed). */
typedef struct {
  const char* name;        /* dataset name, <= 63 chars */
  const char* field;       /* field name, <= 63 chars */
  int32_t dims[3];         /* x, y, z; x fastest in memory */
  int32_t dtype;           /* 0 = float32, 1 = float64 */
  int32_t brick[3];
  int32_t nlevels;
  double tolerance;
  int32_t bricks_per_chunk;
  int32_t chunks_per_file;
  int32_t files_per_dir;
  int32_t bit_planes_per_chunk; /* only 1 is supported (as the reference's decoder) */
} idx2g_encode_params;

/* Decode query: the decode fields of idx2::params (idx2Common.h:91-97,
 * Applications/idx2App.cpp:17-57).  first/last inclusive; first[0] < 0 means
 * the whole volume. */
typedef struct {
  int32_t first[3];
  int32_t last[3];
  int32_t downsampling[3];
  double tolerance;
} idx2g_query;

/* Per-call device timings (milliseconds, HIP events on the context stream). */
typedef struct {
  double total_ms;        /* first to last device operation of the call */
  double transform_ms;    /* encode: transform+coding kernels; decode: decode kernels */
  double assemble_ms;     /* encode: chunk assembly kernels; decode: output/copy kernels */
  int64_t dominant_launches;
  double dominant_ms;     /* device time of the output level's brick launches (encode: level 0) */
  int64_t dominant_bytes; /* algorithmic bytes of those launches: encode raw input + streams written;
                             decode chunk streams read + output grid written */
  int64_t raw_bytes;      /* encode: input bytes; decode: output grid bytes */
  int64_t chunk_bytes;    /* encode: bit-plane chunk bytes produced; decode: chunk bytes touched */
  int64_t exp_bytes;      /* exponent bytes (uncompressed) produced / read */
} idx2g_stats;

/* ---- context ---- */
int idx2g_create(idx2g_ctx** ctx, int hip_device);
int idx2g_destroy(idx2g_ctx* ctx);
const char* idx2g_last_error_message(void);

/* ---- encode ----
 * Replaces idx2::Encode(idx2_file*, const params&, brick_copier&)
 * (Source/Core/idx2Encode.h:143-144, Source/idx2.h:29-30).  The input grid is
 * read from host memory (input_is_device = 0) or device memory (1).  The
 * result stays resident on the device as a dataset (chunks assembled in HBM);
 * idx2g_dataset_write writes the .idx2 file and the L*.bin files. */
int idx2g_encode(idx2g_ctx* ctx, const idx2g_encode_params* p, const void* input,
                 int input_is_device, idx2g_dataset** out);

/* Validate encode parameters without touching the GPU: the checks of
 * idx2::Finalize (Source/Core/idx2Common.cpp:358-754, e.g. TooManyLevels when
 * a level has too few bricks for the transform order) as idx2g_encode runs
 * them.  Returns 0 when idx2g_encode would accept the parameters. */
int idx2g_validate_params(const idx2g_encode_params* p);

/* Batched encode: no counterpart in the reference (idx2App encodes one field
 * per run, Applications/idx2App.cpp:114-168).  n fields of any shapes and
 * dtypes in one pass: each level's bricks of all fields run in the same
 * persistent brick-kernel launches and chunk assembly runs once for the
 * batch; with n > 1 the whole encode is captured into one hipGraph and
 * launched once (key ranges and chunk layout decided on the device; a batch
 * whose level spans more than 128 bit-plane keys is encoded stepwise; the
 * environment variable IDX2G_BATCH_GRAPH=0 selects stepwise).  out[i] is
 * byte-identical to idx2g_encode of field i alone.
 * inputs[i] is field i's grid in host (inputs_are_device = 0) or device
 * memory (1).  On error no dataset is returned. */
int idx2g_encode_batch(idx2g_ctx* ctx, int n, const idx2g_encode_params* params, const void* const* inputs,
                       int inputs_are_device, idx2g_dataset** out);

/* Write out_dir/name/field.idx2 and out_dir/name/field/L..bin with the file
 * layout of FlushChunks/FlushChunkExponents/WriteMetaFile/PrintStats
 * (Source/Core/idx2Write.cpp:30-413, idx2Common.cpp:149-179). */
int idx2g_dataset_write(idx2g_dataset* ds, const char* out_dir);

/* ---- open ----
 * Replaces idx2::Init(idx2_file*, params&) (Source/idx2.h:20-21,
 * Source/idx2.cpp:23-32): reads the .idx2 metadata and every L*.bin file of
 * the dataset and makes chunks, chunk index and exponents resident in HBM. */
int idx2g_open(idx2g_ctx* ctx, const char* idx2_path, const char* in_dir, idx2g_dataset** out);
int idx2g_dataset_free(idx2g_dataset* ds);

/* Dataset properties: dims, dtype, levels, brick; bytes of chunk data resident. */
int idx2g_dataset_info(const idx2g_dataset* ds, int32_t dims[3], int32_t* dtype, int32_t* nlevels,
                       int32_t brick[3], double* tolerance, int64_t* chunk_bytes);

/* ---- decode ----
 * Output grid of a query: replaces idx2::GetOutputGrid (Source/idx2.h:36-37,
 * Source/Core/idx2Common.cpp:770-785). */
int idx2g_output_dims(const idx2g_dataset* ds, const idx2g_query* q, int64_t dims[3], int64_t* bytes);

/* Replaces idx2::Decode(idx2_file*, params&, buffer*) (Source/idx2.h:43-44,
 * Source/Core/idx2Decode.cpp:326-492) in RegularGridMem mode: decodes into a
 * caller-allocated buffer of idx2g_output_dims bytes (host or device).  The
 * buffer is zeroed first, so samples outside every brick read 0. */
int idx2g_decode(idx2g_ctx* ctx, idx2g_dataset* ds, const idx2g_query* q, void* out,
                 int out_is_device);

/* Decode without device allocations.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * idx2g_decode of the same query, into a caller-allocated device buffer of
 * idx2g_output_dims bytes.  Every buffer the decode needs beyond the dataset
 * and the output comes from a fixed 128 MiB array per GPU in the library's
 * code object (held until the process ends) and from the context's launch-
 * table staging (4 MiB of pinned host memory made by idx2g_create); the call
 * makes no hipMalloc/hipFree/hipHostMalloc.  A query whose working set under
 * this call's launch shapes (at most 64 full bricks per walk batch, 64
 * resident workgroups) exceeds the array returns IDX2G_OUT_OF_MEMORY before
 * anything runs.  All calls on one GPU share the array and would race on it
 * when they overlap; the library runs them one at a time (a per-device lock;
 * each call returns after its kernels have finished).  scratch_bytes_used
 * (may be NULL): the peak bytes of the array this call used. */
int idx2g_decode_noalloc(idx2g_ctx* ctx, idx2g_dataset* ds, const idx2g_query* q, void* out_device,
                         int64_t* scratch_bytes_used);

/* Batched decode: no counterpart in the reference (one query per Decode
 * call).  Query i (queries NULL: whole volumes) of dataset ds[i] into out[i];
 * each level's bricks of all queries run in shared launches.  Every result
 * is bit-identical to idx2g_decode of the same query alone. */
int idx2g_decode_batch(idx2g_ctx* ctx, int n, idx2g_dataset* const* ds, const idx2g_query* queries,
                       void* const* out, int out_is_device);

/* Device timings of the last encode/decode call on ctx. */
int idx2g_last_stats(const idx2g_ctx* ctx, idx2g_stats* s);

/* Host copy of a dataset's bit-plane chunk bytes in file order: for each file
 * its path suffix, and the concatenated chunk payloads (for parity tests).
 * Call with buf == NULL to get the size. */
int idx2g_dataset_chunk_bytes(idx2g_dataset* ds, uint8_t* buf, int64_t* nbytes);

/* ---- synthetic input ----
 * Fills a device buffer with the deterministic Gaussian+sine+noise field of
 * idx2_amd/synth.py (same bytes).  params: 192 doubles (synth.field_params). */
int idx2g_synth_field(idx2g_ctx* ctx, void* dev_out, const int32_t dims[3], int32_t dtype,
                      const double* params, uint64_t noise_seed, int32_t z0, int32_t z1);

/* ---- launch-shape autotuner ----
 * No counterpart in the reference.  The brick shape is fixed by the format
 * and the on-chip kernels run one 1024-thread workgroup per CU; what is left
 * to pick per device is how many workgroups of the generic kernels share a
 * CU, how many full bricks one decode walk / on-chip launch pair takes, and
 * the decode walk's workgroup width: its subbands' key segments are packed
 * into 64-lane bins, walked by 2-wave workgroups (many bricks resident per
 * CU: throughput) or 4-wave ones (bins in parallel: latency, for launches
 * with few bricks).  Results never depend on these (the tests check that);
 * only time does. */
typedef struct {
  int32_t enc_wg_per_cu;  /* generic encode kernel: resident workgroups per CU (0: as many as its LDS allows) */
  int32_t dec_wg_per_cu;  /* generic decode kernel: same */
  int32_t walk_batch;     /* decode: full bricks per walk + on-chip launch pair (0: 32768) */
  int32_t walk_wide_max;  /* decode: 3D walk launches of at most this many bricks use 4-wave workgroups,
                             larger ones 2-wave workgroups (0: 8192) */
} idx2g_tuning;
/* Times synthetic 2D and 3D encodes/decodes on ctx's device for each
 * candidate shape, keeps the fastest for that device (process-wide) and
 * returns it.  best_ms / default_ms: the tuned and untuned times of the
 * probe workload (may be NULL). */
int idx2g_autotune(idx2g_ctx* ctx, idx2g_tuning* out, double* best_ms, double* default_ms);
int idx2g_set_tuning(int hip_device, const idx2g_tuning* t);
int idx2g_get_tuning(int hip_device, idx2g_tuning* t);

/* ---- device memory held for reuse ----
 * No counterpart in the reference (it allocates host memory per call).  The
 * library keeps the device blocks its calls free for its next calls
 * (returning them to the runtime costs a device synchronisation each, so it
 * only does so itself when an allocation fails).  A process that shares the
 * GPU with another allocator (a framework's caching allocator, say) calls
 * this when its library calls are done to hand every kept block of the
 * device back; later calls allocate again.  Returns the bytes handed back,
 * or a negative error code. */
int64_t idx2g_release_device_cache(int hip_device);

/* ---- decode serving loop ----
 * No counterpart in the reference (its decoder serves one query per call,
 * Source/Core/idx2Decode.cpp:326-492, Source/idx2.h:43-44): a queue of decode
 * requests on resident datasets.  Each of `slots` execution slots (one HIP
 * stream each) takes the requests that fit the HBM budget when it is free --
 * up to 16, decoded together in shared launches (idx2g_decode_batch) --
 * in arrival order, except that while the oldest request does not fit, up
 * to 64 later ones that fit may pass it, at most 32 times in all (a request
 * larger than the budget alone runs when nothing else holds memory).
 * Coarse levels whose whole grid is at most 512 MB are decoded whole and
 * kept in HBM (a quarter of the budget at most, least recently used first
 * out), keyed by dataset, level, tolerance exponent and decode masks, so
 * later requests with the same key start below them.  Requests run
 * concurrently;
 * each result is the same as idx2g_decode's for the same query.  Datasets
 * are only read; idx2g_dataset_free on a dataset with requests queued or
 * running returns at once and the last of those requests releases it. */
typedef struct idx2g_server idx2g_server;
int idx2g_server_create(int hip_device, int64_t hbm_budget_bytes, int slots, idx2g_server** out);
/* Queue a decode of q (NULL: whole volume) into out (idx2g_output_dims
 * bytes; device memory when out_is_device).  Returns at once. */
int idx2g_server_submit(idx2g_server* s, idx2g_dataset* ds, const idx2g_query* q, void* out,
                        int out_is_device, int64_t* request_id);
/* Wait for a request; returns its decode error code.  latency_ms: submit to
 * completion; queue_ms: submit to start.  Each id can be waited for once. */
int idx2g_server_wait(idx2g_server* s, int64_t request_id, double* latency_ms, double* queue_ms,
                      idx2g_stats* stats);
/* Peak admitted bytes, peak concurrent requests, completed requests. */
int idx2g_server_info(idx2g_server* s, int64_t* peak_bytes, int* peak_running, int64_t* completed);
typedef struct {
  int64_t completed;
  int64_t batches;           /* batched decodes run */
  int64_t batched_requests;  /* requests in them (batched_requests / batches: mean batch) */
  int64_t bypasses;          /* requests admitted ahead of a blocked older one */
  int64_t cache_hits, cache_misses, cache_inserts, cache_evictions;
  int64_t cache_bytes;       /* HBM held by cached coarse levels now */
  int64_t peak_bytes;        /* peak admitted + cached bytes */
  int32_t peak_running;      /* peak concurrent batches */
} idx2g_server_stats;
int idx2g_server_counters(idx2g_server* s, idx2g_server_stats* out);
/* Drains the queue, then frees the server. */
int idx2g_server_destroy(idx2g_server* s);
/* The HBM footprint the server charges a request (output grid + working
 * buffers of the decode). */
int idx2g_decode_footprint(const idx2g_dataset* ds, const idx2g_query* q, int64_t* bytes);

/* ---- sparse-brick output (the reference's HashMap output mode) ----------
 * Replaces Decode with params::out_mode::HashMap (`idx2App --brick_map`,
 * Source/Core/idx2Decode.cpp:423-477, idx2SparseBricks.cpp:20-243): the
 * query is decoded level by level as by idx2g_decode, and every decoded
 * brick that is significant (a block of a detail subband, or of any subband
 * on the coarsest level, decoded at least one bit plane;
 * idx2Decode.cpp:205-212) is kept in HBM with its extended brick (brick + 1
 * per axis of extent > 1) of float64 samples after the inverse transform.
 * Bricks are addressed by the reference's brick key,
 * (linear brick id << 4) + level (GetBrickKey, idx2Lookup.h:91-94); the pool
 * holds a device hash table of these keys.  No output grid is written. */
typedef struct idx2g_brick_pool idx2g_brick_pool;
int idx2g_decode_bricks(idx2g_ctx* ctx, idx2g_dataset* ds, const idx2g_query* q, idx2g_brick_pool** out);
/* Bricks kept on a level and bricks of the level's grid (PrintStatistics,
 * idx2SparseBricks.cpp:49-65); levels the query did not decode keep 0. */
int idx2g_brick_pool_count(const idx2g_brick_pool* p, int level, int64_t* kept, int64_t* total);
/* The brick key of brick3 (brick coordinates) on a level. */
int idx2g_brick_pool_key(const idx2g_brick_pool* p, int level, const int brick3[3], uint64_t* key);
/* Device address of a kept brick's samples (x fastest) and its dims;
 * IDX2G_ATTRIBUTE_NOT_FOUND when the brick is not kept. */
int idx2g_brick_pool_find(const idx2g_brick_pool* p, int level, const int brick3[3], const double** d_samples,
                          int dims[3]);
/* Copy a kept brick's samples (dims from idx2g_brick_pool_find) to host
 * (out_is_device = 0) or device memory (1). */
int idx2g_brick_pool_copy(const idx2g_brick_pool* p, int level, const int brick3[3], double* out, int out_is_device);
/* Device-side lookup of n brick keys (device array) through the pool's hash
 * table: d_samples[i] = device address of the brick's samples, or NULL. */
int idx2g_brick_pool_lookup(idx2g_ctx* ctx, const idx2g_brick_pool* p, int64_t n, const uint64_t* d_keys,
                            const double** d_samples);
/* Resolution of a finest-level brick: the finest level at which it or an
 * ancestor is kept, the coarsest level when none is
 * (ComputeBrickResolution, idx2SparseBricks.cpp:69-152). */
int idx2g_brick_pool_resolution(const idx2g_brick_pool* p, const int brick3[3], int* level);
/* The reference's brick file (WriteBricks and GetBrickVolume, idx2SparseBricks.cpp:157-243):
 * dims, brick dims and finest brick counts (3 int32 each), then for every
 * finest brick in row-major order its coordinates and sample dims (3 int32
 * each) and its samples (float64), taken from the kept brick of its
 * resolution.  IDX2G_FILE_CREATE_FAILED when the file cannot be created,
 * IDX2G_ATTRIBUTE_NOT_FOUND when a finest brick has no kept brick on its
 * path (a query that did not cover the volume). */
int idx2g_brick_pool_write(const idx2g_brick_pool* p, const char* path);
int idx2g_brick_pool_free(idx2g_brick_pool* p);

#ifdef __cplusplus
}
#endif
#endif
