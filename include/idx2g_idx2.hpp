/* idx2g_idx2.hpp: the GPU path behind the reference's own C++ types.
 *
 * Header-only adapter for hosts that already use the reference library
 * (idx2App, idx2Samples, OpenVisus): include the reference's "idx2.h" first,
 * then this file, and link with -lidx2g.  Each function takes the reference's
 * idx2::params and returns the reference's error<idx2_err_code>, so a call
 * site swaps
 *   Encode(&Idx2, P, Copier)                (Applications/idx2App.cpp:298)
 *   Init(&Idx2, P); Decode(&Idx2, P, &Buf)  (Source/idx2.h:20-44, idx2.cpp:23-47)
 * for idx2g_idx2::Encode(P, grid, ...) / idx2g_idx2::Decode(P, &Buf, ...).
 *
 * Error mapping: an idx2g code is the value of the reference's enum class
 * idx2_err_code (Source/Core/idx2Common.h:36-56, declared by the idx2_Enum of
 * Source/Core/Enum.h:65-80 with NoError = 0); the static_asserts below check
 * every code against the reference's own enum.
 */
#ifndef IDX2G_IDX2_HPP
#define IDX2G_IDX2_HPP

#include "idx2g.h"

namespace idx2g_idx2 {

#define IDX2G_SAME_CODE(ours, theirs) \
  static_assert(int(ours) == int(idx2::idx2_err_code::theirs), #theirs " numbering")
IDX2G_SAME_CODE(IDX2G_OK, NoError);
IDX2G_SAME_CODE(IDX2G_UNKNOWN, UnknownError);
IDX2G_SAME_CODE(IDX2G_SIZE_ZERO, SizeZero);
IDX2G_SAME_CODE(IDX2G_SIZE_TOO_SMALL, SizeTooSmall);
IDX2G_SAME_CODE(IDX2G_SIZE_MISMATCHED, SizeMismatched);
IDX2G_SAME_CODE(IDX2G_DIMENSION_MISMATCHED, DimensionMismatched);
IDX2G_SAME_CODE(IDX2G_DIMENSIONS_TOO_MANY, DimensionsTooMany);
IDX2G_SAME_CODE(IDX2G_ATTRIBUTE_NOT_FOUND, AttributeNotFound);
IDX2G_SAME_CODE(IDX2G_OPTION_NOT_SUPPORTED, OptionNotSupported);
IDX2G_SAME_CODE(IDX2G_TYPE_NOT_SUPPORTED, TypeNotSupported);
IDX2G_SAME_CODE(IDX2G_FILE_CREATE_FAILED, FileCreateFailed);
IDX2G_SAME_CODE(IDX2G_FILE_READ_FAILED, FileReadFailed);
IDX2G_SAME_CODE(IDX2G_FILE_WRITE_FAILED, FileWriteFailed);
IDX2G_SAME_CODE(IDX2G_FILE_OPEN_FAILED, FileOpenFailed);
IDX2G_SAME_CODE(IDX2G_FILE_CLOSE_FAILED, FileCloseFailed);
IDX2G_SAME_CODE(IDX2G_FILE_SEEK_FAILED, FileSeekFailed);
IDX2G_SAME_CODE(IDX2G_FILE_TELL_FAILED, FileTellFailed);
IDX2G_SAME_CODE(IDX2G_PARSE_FAILED, ParseFailed);
IDX2G_SAME_CODE(IDX2G_OUT_OF_MEMORY, OutOfMemory);
IDX2G_SAME_CODE(IDX2G_DIMENSIONS_REPEATED, DimensionsRepeated);
IDX2G_SAME_CODE(IDX2G_DIMENSION_TOO_SMALL, DimensionTooSmall);
IDX2G_SAME_CODE(IDX2G_BRICK_SIZE_NOT_POW2, BrickSizeNotPowerOfTwo);
IDX2G_SAME_CODE(IDX2G_BRICK_SIZE_TOO_BIG, BrickSizeTooBig);
IDX2G_SAME_CODE(IDX2G_TOO_MANY_LEVELS, TooManyLevels);
IDX2G_SAME_CODE(IDX2G_TOO_MANY_TRANSFORM_PASSES_PER_LEVEL, TooManyTransformPassesPerLevel);
IDX2G_SAME_CODE(IDX2G_TOO_MANY_LEVELS_OR_TRANSFORM_PASSES, TooManyLevelsOrTransformPasses);
IDX2G_SAME_CODE(IDX2G_TOO_MANY_BRICKS_PER_FILE, TooManyBricksPerFile);
IDX2G_SAME_CODE(IDX2G_TOO_MANY_FILES_PER_DIR, TooManyFilesPerDir);
IDX2G_SAME_CODE(IDX2G_NOT_SUPPORTED, NotSupportedInVersion);
IDX2G_SAME_CODE(IDX2G_CANNOT_CREATE_DIRECTORY, CannotCreateDirectory);
IDX2G_SAME_CODE(IDX2G_SYNTAX_ERROR, SyntaxError);
IDX2G_SAME_CODE(IDX2G_TOO_MANY_BRICKS_PER_CHUNK, TooManyBricksPerChunk);
IDX2G_SAME_CODE(IDX2G_TOO_MANY_CHUNKS_PER_FILE, TooManyChunksPerFile);
IDX2G_SAME_CODE(IDX2G_CHUNKS_PER_FILE_NOT_POW2, ChunksPerFileNotPowerOf2);
IDX2G_SAME_CODE(IDX2G_BRICKS_PER_CHUNK_NOT_POW2, BricksPerChunkNotPowerOf2);
IDX2G_SAME_CODE(IDX2G_CHUNK_NOT_FOUND, ChunkNotFound);
IDX2G_SAME_CODE(IDX2G_BRICK_NOT_FOUND, BrickNotFound);
IDX2G_SAME_CODE(IDX2G_FILE_NOT_FOUND, FileNotFound);
IDX2G_SAME_CODE(IDX2G_UNSUPPORTED_SCHEME, UnsupportedScheme);
#undef IDX2G_SAME_CODE

/* idx2g code -> the reference's error value (a device failure has no
 * reference code and maps to UnknownError). */
inline idx2::error<idx2::idx2_err_code> ToError(int code) {
  if (code == IDX2G_DEVICE_ERROR) code = IDX2G_UNKNOWN;
  return idx2::error<idx2::idx2_err_code>(idx2::idx2_err_code(code), false,
                                          code ? idx2g_last_error_message() : "");
}

/* The encode fields of idx2::params (Source/Core/idx2Common.h:74-90). */
inline idx2g_encode_params EncodeParams(const idx2::params& P) {
  idx2g_encode_params ep = {};
  ep.name = P.Meta.Name;
  ep.field = P.Meta.Field;
  for (int d = 0; d < 3; ++d) {
    ep.dims[d] = P.Meta.Dims3[d];
    ep.brick[d] = P.BrickDims3[d];
  }
  ep.dtype = P.Meta.DType == idx2::dtype::float64 ? 1 : 0;
  ep.nlevels = P.NLevels;
  ep.tolerance = P.Tolerance;
  ep.bricks_per_chunk = P.BricksPerChunk;
  ep.chunks_per_file = P.ChunksPerFile;
  ep.files_per_dir = P.FilesPerDir;
  ep.bit_planes_per_chunk = P.BitPlanesPerChunk;
  return ep;
}

/* The checks Finalize runs for an encode (idx2App's SetParams + Finalize,
 * Applications/idx2App.cpp:216-233), without touching the GPU. */
inline idx2::error<idx2::idx2_err_code> Validate(const idx2::params& P) {
  if (P.Meta.DType != idx2::dtype::float32 && P.Meta.DType != idx2::dtype::float64)
    return ToError(IDX2G_TYPE_NOT_SUPPORTED);
  const idx2g_encode_params ep = EncodeParams(P);
  return ToError(idx2g_validate_params(&ep));
}

/* Replaces Encode(&Idx2, P, Copier) + the file flush: encodes the grid (host
 * memory, or device memory when grid_is_device) and writes P.OutDir/Name/
 * Field.idx2 and the L*.bin files. */
inline idx2::error<idx2::idx2_err_code> Encode(const idx2::params& P, const void* grid, int grid_is_device = 0,
                                               int hip_device = 0) {
  if (P.Meta.DType != idx2::dtype::float32 && P.Meta.DType != idx2::dtype::float64)
    return ToError(IDX2G_TYPE_NOT_SUPPORTED);
  idx2g_ctx* ctx = nullptr;
  int e = idx2g_create(&ctx, hip_device);
  if (e) return ToError(e);
  const idx2g_encode_params ep = EncodeParams(P);
  idx2g_dataset* ds = nullptr;
  e = idx2g_encode(ctx, &ep, grid, grid_is_device, &ds);
  if (!e) e = idx2g_dataset_write(ds, P.OutDir);
  if (ds) idx2g_dataset_free(ds);
  idx2g_destroy(ctx);
  return ToError(e);
}

/* The decode query of idx2::params (DecodeExtent, DownsamplingFactor3,
 * DecodeTolerance; an empty extent is the whole volume, as in Init). */
inline idx2g_query Query(const idx2::params& P) {
  idx2g_query q = {};
  const idx2::v3i from = idx2::From(P.DecodeExtent), dims = idx2::Dims(P.DecodeExtent);
  const bool whole = dims == idx2::v3i(0);
  for (int d = 0; d < 3; ++d) {
    q.first[d] = whole ? -1 : from[d];
    q.last[d] = whole ? -1 : from[d] + dims[d] - 1;
    q.downsampling[d] = P.DownsamplingFactor3[d];
  }
  q.tolerance = P.DecodeTolerance;
  return q;
}

/* Replaces Init(&Idx2, P) + GetOutputGrid + Decode(&Idx2, P, &OutBuf) in
 * RegularGridMem mode: OutBuf is allocated (idx2::AllocBuf) to the output
 * grid's size and filled with the same samples as the reference's decoder. */
inline idx2::error<idx2::idx2_err_code> Decode(const idx2::params& P, idx2::buffer* OutBuf, int hip_device = 0) {
  idx2g_ctx* ctx = nullptr;
  int e = idx2g_create(&ctx, hip_device);
  if (e) return ToError(e);
  char dir[1024];
  const int n = P.InDir.Size < (int)sizeof(dir) - 1 ? P.InDir.Size : (int)sizeof(dir) - 1;
  for (int i = 0; i < n; ++i) dir[i] = P.InDir.ConstPtr[i];
  dir[n] = 0;
  idx2g_dataset* ds = nullptr;
  e = idx2g_open(ctx, P.InputFile, dir, &ds);
  const idx2g_query q = Query(P);
  int64_t dims[3] = {0, 0, 0}, bytes = 0;
  if (!e) e = idx2g_output_dims(ds, &q, dims, &bytes);
  if (!e) {
    idx2::AllocBuf(OutBuf, bytes);
    e = idx2g_decode(ctx, ds, &q, OutBuf->Data, 0);
  }
  if (ds) idx2g_dataset_free(ds);
  idx2g_destroy(ctx);
  return ToError(e);
}

}  // namespace idx2g_idx2

#endif
