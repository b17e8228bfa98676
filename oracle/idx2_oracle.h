/* idx2 CPU oracle: a plain-C restatement of the idx2 encode/decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker.
 * The product (idx2_amd, libidx2g) never links or calls it.
 *
 * Error codes are the values of the reference enum class idx2_err_code
 * (idx2Common.h:36-56 after ErrorCodes.h:6-12; Enum.h:65-80 puts __Invalid__
 * last, so NoError = 0).
 */
#ifndef IDX2_ORACLE_H
#define IDX2_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  IDX2O_OK = 0,
  IDX2O_SIZE_MISMATCHED = 4,
  IDX2O_BRICK_SIZE_NOT_POW2 = 21,
  IDX2O_BRICK_SIZE_TOO_BIG = 22,
  IDX2O_TOO_MANY_LEVELS = 23,
  IDX2O_TOO_MANY_BRICKS_PER_FILE = 26,
  IDX2O_TOO_MANY_FILES_PER_DIR = 27,
  IDX2O_CANNOT_CREATE_DIR = 29,
  IDX2O_SYNTAX_ERROR = 30,
  IDX2O_TOO_MANY_BRICKS_PER_CHUNK = 31,
  IDX2O_TOO_MANY_CHUNKS_PER_FILE = 32,
  IDX2O_CHUNKS_PER_FILE_NOT_POW2 = 33,
  IDX2O_BRICKS_PER_CHUNK_NOT_POW2 = 34,
  IDX2O_CHUNK_NOT_FOUND = 35,
  IDX2O_FILE_NOT_FOUND = 37,
};

/* Encode a raw grid (x fastest) exactly like `idx2App --encode`.  Writes
 * out_dir/name/field.idx2 and out_dir/name/field/L..bin.  dtype: 0 = float32,
 * 1 = float64.  Files are appended to, so out_dir/name/field should not exist. */
int idx2o_encode(const char* out_dir, const char* name, const char* field, const void* data,
                 int dtype, const int dims[3], const int brick[3], int nlevels, double tolerance,
                 int bricks_per_chunk, int chunks_per_file, int files_per_dir);

/* Output grid dims of a decode query (GetGrid).  first/last inclusive; pass
 * first[0] < 0 for the whole volume. */
int idx2o_output_dims(const char* idx2_path, const int first[3], const int last[3],
                      const int downsampling[3], int out_dims[3], int* dtype);

/* Decode like `idx2App --decode` (serial decoder) into a caller buffer of
 * out_dims (from idx2o_output_dims) samples of the dataset dtype.  The buffer
 * is zeroed first. */
int idx2o_decode(const char* idx2_path, const char* in_dir, const int first[3], const int last[3],
                 const int downsampling[3], double tolerance, void* out, int64_t out_bytes);

/* --- pieces exposed for unit tests --- */
/* Forward/inverse CDF5/3 lifting of one brick volume (dims n3, row-major, x
 * fastest) as in Wavelet.cpp: extrapolation then one forward pass + norms. */
void idx2o_extrapolate(const int data_dims[3], const int n3[3], double* vol);
void idx2o_forward_brick(const int n3[3], int level, int coarsest, double* vol);
void idx2o_inverse_brick(const int n3[3], int level, int coarsest, double* vol);
void idx2o_flift(int axis, const int from[3], const int dims[3], const int strd[3],
                 const int m3[3], const int n3[3], double* vol);
void idx2o_ilift(int axis, const int from[3], const int dims[3], const int strd[3],
                 const int m3[3], const int n3[3], double* vol);
/* zfp block stage: quantise+lift+shuffle (returns EMax) and its inverse. */
int idx2o_block_forward(const double* in, int ndims, int is_f64, uint64_t* out);
void idx2o_block_inverse(const uint64_t* in, int ndims, int emax, double* out);
/* Embedded coder for one bit plane: returns bits written; *n updated. */
int idx2o_encode_plane(const uint64_t* block, int nvals, int bp, int* n, uint8_t* out, int64_t bitpos);
int idx2o_decode_plane(uint64_t* block, int nvals, int bp, int* n, const uint8_t* in, int64_t bitpos);
void idx2o_cdf53_norms(double scal[16], double wave[16]);
/* 4-point zfp lifting (FLift/ILift, Zfp.h:144-204) on p[0..3] */
void idx2o_zfp_lift(int64_t* p, int fwd);

#ifdef __cplusplus
}
#endif
#endif
