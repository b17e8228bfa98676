/* idx2 CPU oracle -- plain-C restatement of the idx2 encode/decode path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never linked by the product.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's Source/ directory).  Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off), zstd through the system libzstd.so.1.
 */
#define _GNU_SOURCE
#include "idx2_oracle.h"

#include <errno.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

/* zstd public API (libzstd.so.1, zstd.h); the reference vendors zstd 1.4.5
 * and calls it at Core/idx2Encode.cpp:23-37 and Core/idx2Decode.cpp:53-78. */
size_t ZSTD_compressBound(size_t srcSize);
size_t ZSTD_compress(void* dst, size_t dstCapacity, const void* src, size_t srcSize, int level);
unsigned long long ZSTD_getFrameContentSize(const void* src, size_t srcSize);
size_t ZSTD_decompress(void* dst, size_t dstCapacity, const void* src, size_t srcSize);
unsigned ZSTD_isError(size_t code);

typedef int64_t i64;
typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;

#define MAXL 16
#define BP_BIAS 1024 /* BitPlaneKeyBias_, Core/idx2Lookup.h:11 */

/* ------------------------------------------------------------------------ */
/* small integer helpers (Core/Math.h:1211-1251, Core/BitOps.h)              */
/* ------------------------------------------------------------------------ */
static int msb64(u64 v) { return v ? 63 - __builtin_clzll(v) : -1; }
static int log2ceil(i64 v) { int i = msb64((u64)v); return i + (v != (1LL << i)); }
static int log2floor(i64 v) { return msb64((u64)v); }
static i64 nextpow2(i64 v) { return v == 0 ? 1 : (1LL << (msb64((u64)(v - 1)) + 1)); }
static int ispow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* bit streams: LSB-first, zero-padded bytes (Core/BitStream.h:19-343)      */
/* ------------------------------------------------------------------------ */
typedef struct { u8* buf; i64 cap; i64 nbits; } bw_t;

static void bw_reserve(bw_t* w, i64 bytes)
{
  if (bytes + 16 <= w->cap) return;
  i64 nc = w->cap ? w->cap : 256;
  while (nc < bytes + 16) nc = nc * 2;
  w->buf = (u8*)realloc(w->buf, (size_t)nc);
  memset(w->buf + w->cap, 0, (size_t)(nc - w->cap));
  w->cap = nc;
}
static i64 bw_size(const bw_t* w) { return (w->nbits + 7) >> 3; }
static void bw_rewind(bw_t* w)
{
  if (w->buf) memset(w->buf, 0, (size_t)bw_size(w) + 8);
  w->nbits = 0;
}
static void bw_free(bw_t* w) { free(w->buf); w->buf = NULL; w->cap = w->nbits = 0; }
static void bw_put(bw_t* w, u64 v, int n) /* Write/WriteLong, BitStream.h:260-283 */
{
  if (n <= 0) return;
  if (n < 64) v &= (1ULL << n) - 1;
  bw_reserve(w, (w->nbits + n + 7) / 8 + 8);
  for (int i = 0; i < n;) {
    i64 byte = w->nbits >> 3;
    int off = (int)(w->nbits & 7);
    int take = imin(8 - off, n - i);
    w->buf[byte] |= (u8)(((v >> i) & ((1u << take) - 1)) << off);
    w->nbits += take;
    i += take;
  }
}
static void bw_align(bw_t* w) { w->nbits = (w->nbits + 7) & ~(i64)7; }
/* WriteStream (BitStream.h:286-294): byte-align dst, append Size(src) bytes */
static void bw_append_stream(bw_t* dst, const bw_t* src)
{
  bw_align(dst);
  i64 n = bw_size(src);
  bw_reserve(dst, bw_size(dst) + n + 8);
  if (n) memcpy(dst->buf + (dst->nbits >> 3), src->buf, (size_t)n);
  dst->nbits += n * 8;
}
static void bw_append_bytes(bw_t* dst, const void* p, i64 n)
{
  bw_align(dst);
  bw_reserve(dst, bw_size(dst) + n + 8);
  if (n) memcpy(dst->buf + (dst->nbits >> 3), p, (size_t)n);
  dst->nbits += n * 8;
}
/* WriteVarByte, Core/VarInt.h:13-26 */
static void bw_varbyte(bw_t* w, u64 v)
{
  for (;;) {
    bw_put(w, v & 0x7F, 7);
    v >>= 7;
    bw_put(w, v != 0, 1);
    if (v == 0) break;
  }
}
/* WriteUnary, Core/VarInt.h:81-107: Val zero bits then a one */
static void bw_unary(bw_t* w, u32 v)
{
  bw_reserve(w, (w->nbits + v + 8) / 8 + 8);
  w->nbits += v; /* zero bits (buffer is zeroed) */
  bw_put(w, 1, 1);
}

typedef struct { const u8* buf; i64 nbytes; i64 pos; } br_t;
static u64 br_get(br_t* r, int n) /* Read/ReadLong, BitStream.h:178-204 */
{
  u64 v = 0;
  for (int i = 0; i < n; ++i) {
    i64 byte = r->pos >> 3;
    u64 bit = byte < r->nbytes ? (r->buf[byte] >> (r->pos & 7)) & 1u : 0;
    v |= bit << i;
    r->pos++;
  }
  return v;
}
static u64 br_varbyte(br_t* r) /* ReadVarByte, VarInt.h:29-41 */
{
  u64 v = 0;
  int shift = 0;
  for (;;) {
    v = (br_get(r, 7) << (7 * shift++)) + v;
    if (br_get(r, 1) == 0) break;
  }
  return v;
}
static u32 br_unary(br_t* r) /* ReadUnary, VarInt.h:110-117 */
{
  u32 v = 0;
  while (!br_get(r, 1)) ++v;
  return v;
}
static void br_align(br_t* r) { r->pos = (r->pos + 7) & ~(i64)7; }

/* ------------------------------------------------------------------------ */
/* grids                                                                     */
/* ------------------------------------------------------------------------ */
typedef struct { int f[3], d[3], s[3]; } grid_t;

static grid_t mkgrid(const int f[3], const int d[3], const int s[3])
{
  grid_t g;
  for (int i = 0; i < 3; ++i) { g.f[i] = f[i]; g.d[i] = d[i]; g.s[i] = s[i]; }
  return g;
}

/* ------------------------------------------------------------------------ */
/* transform-order strings (Core/Wavelet.cpp:406-490)                       */
/* ------------------------------------------------------------------------ */
static u64 encode_order(const char* s)
{
  u64 r = 0;
  for (int i = (int)strlen(s) - 1; i >= 0; --i) {
    int t = s[i] == '+' ? 3 : s[i] - 'X';
    r = (r << 2) + (u64)t;
  }
  return r;
}
/* DecodeTransformOrder(Input, N3, Output), Wavelet.cpp:435-460 */
static int decode_order_n3(u64 in, const int n3in[3], char* out)
{
  i64 n3[3];
  for (int i = 0; i < 3; ++i) n3[i] = nextpow2(n3in[i]);
  int len = 0;
  u64 saved = in;
  while (n3[0] * n3[1] * n3[2] > 1) {
    int t = (int)(in & 3);
    in >>= 2;
    if (t == 3) {
      if (in == 3) in = saved; else saved = in;
    } else {
      out[len++] = (char)('X' + t);
      n3[t] >>= 1;
    }
  }
  out[len] = 0;
  return len;
}
/* DecodeTransformOrder(Input, Passes, Output), Wavelet.cpp:463-490 */
static int decode_order_passes(u64 in, int passes, char* out)
{
  int len = 0;
  u64 saved = in;
  int p3[3] = {passes, passes, passes};
  for (;;) {
    int t = (int)(in & 3);
    in >>= 2;
    if (t == 3) {
      if (in == 3) in = saved; else saved = in;
    } else {
      if (--p3[t] < 0) break;
      out[len++] = (char)('X' + t);
    }
  }
  out[len] = 0;
  return len;
}
static void decode_order_plain(u64 in, char* out) /* Wavelet.cpp:422-433 */
{
  int len = 0;
  while (in != 0) {
    int t = (int)(in & 3);
    out[len++] = t == 3 ? '+' : (char)('X' + t);
    in >>= 2;
  }
  out[len] = 0;
}

/* ------------------------------------------------------------------------ */
/* dataset layout: idx2_file + Finalize (Core/idx2Common.h:111-184,         */
/* Core/idx2Common.cpp:358-754)                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  char name[64], field[64];
  int dims[3];
  int dtype; /* 0 f32, 1 f64 */
  int brick[3], brick_ext[3];
  int nlevels;
  double tolerance;
  int bpc_in, cpf_in, files_per_dir, bppc;
  int version[2];
  u64 tform_order;
  char tform_order_str[8];
  int tform_full_len;
  int nsb;
  grid_t sb[8];
  int sb_lh[8][3];
  int sb_level3rev[8][3];
  int group[3];
  int nbricks[MAXL][3];
  char bstr[MAXL][128];
  int blen[MAXL];
  u64 border[MAXL];
  int bpc[MAXL], cpf[MAXL];
  i64 bpf[MAXL];
  int bpc3[MAXL][3], cpf3[MAXL][3], nchunks3[MAXL][3], nfiles3[MAXL][3];
  u64 border_in_chunk[MAXL], corder_in_file[MAXL], forder[MAXL];
  int fdd[MAXL][8], fdd_len[MAXL];
  int td_n;
  grid_t td_grid[8];
  int td_axis[8];
  int ds[3];
  u8 mask[MAXL];
  int spacing[MAXL][8][3];
  double vmin, vmax;
  char dir[1024];
} ofile_t;

/* BuildSubbands(N3, NLevels=1, order) + decode masks/spacings
 * (Wavelet.cpp:502-567, idx2Common.cpp:402-460) */
static void build_subbands(ofile_t* F)
{
  for (int d = 0; d < 3; ++d) F->brick_ext[d] = F->brick[d] + (F->brick[d] > 1);
  /* One level: split every transformed axis of the extended brick into even
   * (low) and odd (high) samples.  Index = bits over transformed axes, X most
   * significant, matching the queue order then reversal of Wavelet.cpp:513-566. */
  int axes[3], na = 0;
  for (int d = 0; d < 3; ++d) if (F->brick[d] > 1) axes[na++] = d;
  F->nsb = 1 << na;
  for (int sb = 0; sb < F->nsb; ++sb) {
    int lh[3] = {0, 0, 0};
    for (int k = 0; k < na; ++k) lh[axes[k]] = (sb >> (na - 1 - k)) & 1;
    for (int d = 0; d < 3; ++d) {
      int n = F->brick_ext[d];
      F->sb_lh[sb][d] = lh[d];
      if (F->brick[d] > 1) {
        F->sb[sb].f[d] = lh[d];
        F->sb[sb].s[d] = 2;
        F->sb[sb].d[d] = lh[d] ? n / 2 : (n + 1) / 2;
        F->sb_level3rev[sb][d] = lh[d] ? 0 : 1;
      } else {
        F->sb[sb].f[d] = 0; F->sb[sb].s[d] = 1; F->sb[sb].d[d] = n;
        F->sb_level3rev[sb][d] = 0;
      }
    }
  }
  /* decode masks (idx2Common.cpp:411-452) */
  int sp3[3];
  for (int d = 0; d < 3; ++d) sp3[d] = 1 << F->ds[d];
  int prev0s[3] = {1, 1, 1};
  for (int I = 0; I < F->nlevels; ++I) {
    u8 mask = 0xFF;
    int cur0s[3] = {1, 1, 1};
    for (int sb = 0; sb < F->nsb; ++sb) {
      int f3[3], s3[3];
      for (int d = 0; d < 3; ++d) {
        f3[d] = F->sb[sb].f[d] * (I > 0 ? prev0s[d] : 1);
        s3[d] = F->sb[sb].s[d] * (I > 0 ? prev0s[d] : 1);
        F->spacing[I][sb][d] = 1;
      }
      if (sb == 0) for (int d = 0; d < 3; ++d) cur0s[d] = s3[d];
      for (int d = 0; d < 3; ++d) {
        if (f3[d] % sp3[d] == 0) {
          if (s3[d] % sp3[d] != 0) F->spacing[I][sb][d] = sp3[d] / s3[d];
        } else {
          mask = (u8)(mask & ~(1u << sb));
        }
      }
    }
    if (I + 1 < F->nlevels && mask == 1) mask = 0;
    F->mask[I] = mask;
    for (int d = 0; d < 3; ++d) prev0s[d] = cur0s[d];
  }
}

static int finalize(ofile_t* F)
{
  /* CheckBrickSize, idx2Common.cpp:358-370 */
  for (int d = 0; d < 3; ++d)
    if (!ispow2(F->brick[d])) return IDX2O_BRICK_SIZE_NOT_POW2;
  for (int d = 0; d < 3; ++d)
    if (!(F->dims[d] >= F->brick[d])) return IDX2O_BRICK_SIZE_TOO_BIG;
  /* GuessNumLevelsIfNeeded, idx2Common.cpp:666-680 */
  if (F->nlevels == 0) {
    int b[3] = {F->brick[0], F->brick[1], F->brick[2]};
    while (b[0] <= F->dims[0] && b[1] <= F->dims[1] && b[2] <= F->dims[2]) {
      for (int d = 0; d < 3; ++d) b[d] = b[d] == 1 ? 1 : b[d] * 2;
      if (b[0] <= F->dims[0] && b[1] <= F->dims[1] && b[2] <= F->dims[2]) ++F->nlevels;
    }
  }
  if (F->nlevels > MAXL) return IDX2O_TOO_MANY_LEVELS;
  /* ComputeTransformOrder, idx2Common.cpp:373-387 */
  int j = 0;
  for (int d = 0; d < 3; ++d) if (F->brick[d] > 1) F->tform_order_str[j++] = (char)('X' + d);
  F->tform_order_str[j++] = '+';
  F->tform_order_str[j++] = '+';
  F->tform_order_str[j] = 0;
  F->tform_order = encode_order(F->tform_order_str);
  char tmp[128];
  F->tform_full_len = decode_order_passes(F->tform_order, 1, tmp);
  build_subbands(F);
  /* ComputeNumBricksPerLevel, idx2Common.cpp:463-474 */
  /* GroupBrick3 = BrickDims3 / Dims(SubbandsNonExt[0]): 2 on transformed axes */
  for (int d = 0; d < 3; ++d) F->group[d] = F->brick[d] > 1 ? F->brick[d] / ((F->brick[d] + 1) / 2) : 1;
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = (F->dims[d] + F->brick[d] - 1) / F->brick[d];
  for (int I = 0; I < F->nlevels; ++I) {
    for (int d = 0; d < 3; ++d) {
      F->nbricks[I][d] = nb[d];
      nb[d] = (nb[d] + F->group[d] - 1) / F->group[d];
    }
  }
  /* ComputeGlobalBricksOrder, idx2Common.cpp:478-521 */
  for (int I = 0; I < F->nlevels; ++I) {
    int* n3 = F->nbricks[I];
    int logn[3] = {log2ceil(n3[0]), log2ceil(n3[1]), log2ceil(n3[2])};
    int minlog = imin(logn[0], imin(logn[1], logn[2]));
    int left[3];
    for (int d = 0; d < 3; ++d) left[d] = logn[d] - (F->brick[d] > 1) * minlog;
    char bo[128];
    int J = 0;
    for (int d = 0; d < 3; ++d)
      if (F->brick[d] == 1)
        while (left[d]-- > 0) bo[J++] = (char)('X' + d);
    while (!(left[0] <= 0 && left[1] <= 0 && left[2] <= 0))
      for (int d = 0; d < 3; ++d)
        if (left[d]-- > 0) bo[J++] = (char)('X' + d);
    if (J > 0) bo[J++] = '+';
    for (int k = 0; k < 8; ++k) bo[J++] = F->tform_order_str[k] ? F->tform_order_str[k] : 0;
    bo[J] = 0;
    F->border[I] = encode_order(bo);
    F->blen[I] = decode_order_n3(F->border[I], n3, F->bstr[I]);
    if (F->blen[I] < F->tform_full_len) return IDX2O_TOO_MANY_LEVELS;
  }
  /* ComputeLocalBricksChunksFilesOrders, idx2Common.cpp:524-623 */
  F->bpc[0] = F->bpc_in;
  F->cpf[0] = F->cpf_in;
  if (!(F->bpc[0] <= 32768)) return IDX2O_TOO_MANY_BRICKS_PER_CHUNK;
  if (!ispow2(F->bpc[0])) return IDX2O_BRICKS_PER_CHUNK_NOT_POW2;
  if (!(F->cpf[0] <= 4906)) return IDX2O_TOO_MANY_CHUNKS_PER_FILE;
  if (!ispow2(F->cpf[0])) return IDX2O_CHUNKS_PER_FILE_NOT_POW2;
  for (int I = 0; I < F->nlevels; ++I) {
    int len = F->blen[I];
    F->bpc[I] = 1 << imin(log2ceil(F->bpc[0]), len);
    int bic = log2ceil(F->bpc[I]);
    char s[64];
    for (int d = 0; d < 3; ++d) F->bpc3[I][d] = 1;
    for (int J = 0; J < bic; ++J) {
      char c = F->bstr[I][len - J - 1];
      F->bpc3[I][c - 'X'] *= 2;
      s[bic - J - 1] = c;
    }
    s[bic] = 0;
    F->border_in_chunk[I] = encode_order(s);
    for (int d = 0; d < 3; ++d)
      F->nchunks3[I][d] = (F->nbricks[I][d] + F->bpc3[I][d] - 1) / F->bpc3[I][d];
    F->cpf[I] = 1 << imin(log2ceil(F->cpf[0]), len - bic);
    int cif = log2ceil(F->cpf[I]);
    for (int d = 0; d < 3; ++d) F->cpf3[I][d] = 1;
    for (int J = 0; J < cif; ++J) {
      char c = F->bstr[I][len - bic - J - 1];
      F->cpf3[I][c - 'X'] *= 2;
      s[cif - J - 1] = c;
    }
    s[cif] = 0;
    F->corder_in_file[I] = encode_order(s);
    for (int d = 0; d < 3; ++d)
      F->nfiles3[I][d] = (F->nchunks3[I][d] + F->cpf3[I][d] - 1) / F->cpf3[I][d];
    int fpv = 1 << (len - bic - cif);
    int fl = log2ceil(fpv);
    for (int J = 0; J < fl; ++J) s[fl - J - 1] = F->bstr[I][len - bic - cif - J - 1];
    s[fl] = 0;
    F->forder[I] = encode_order(s);
  }
  /* ComputeFileDirDepths, idx2Common.cpp:626-651 */
  if (!(F->files_per_dir <= 4096)) return IDX2O_TOO_MANY_FILES_PER_DIR;
  for (int I = 0; I < F->nlevels; ++I) {
    F->bpf[I] = (i64)F->bpc[I] * F->cpf[I];
    F->fdd_len[I] = 0;
    int acc = F->fdd[I][F->fdd_len[I]++] = log2ceil(F->bpf[I]);
    int len = F->blen[I];
    while (acc < len) {
      int inc = imin(len - acc, log2ceil(F->files_per_dir));
      acc += (F->fdd[I][F->fdd_len[I]++] = inc);
    }
    if (F->fdd_len[I] > 4) return IDX2O_TOO_MANY_FILES_PER_DIR;
    for (int a = 0, b = F->fdd_len[I] - 1; a < b; ++a, --b) {
      int t = F->fdd[I][a]; F->fdd[I][a] = F->fdd[I][b]; F->fdd[I][b] = t;
    }
  }
  /* ComputeTransformDetails for one pass on the extended brick
   * (Wavelet.cpp:20-60, idx2Common.cpp:655-661) */
  {
    int D3[3], R3[3], S3[3] = {1, 1, 1}, zero[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) D3[d] = R3[d] = F->brick_ext[d];
    grid_t G = mkgrid(zero, F->brick_ext, S3);
    u64 order = F->tform_order, prev = order;
    int pass = 0;
    F->td_n = 0;
    while (pass < 1) {
      int D = (int)(order & 3);
      order >>= 2;
      if (D == 3) {
        if (order == 3) order = prev; else prev = order;
        for (int d = 0; d < 3; ++d) { G.s[d] = S3[d]; G.d[d] = D3[d]; R3[d] = D3[d]; }
        ++pass;
      } else {
        F->td_grid[F->td_n] = G;
        F->td_axis[F->td_n++] = D;
        R3[D] = D3[D] + (D3[D] % 2 == 0);
        for (int d = 0; d < 3; ++d) G.d[d] = R3[d];
        D3[D] = (R3[D] + 1) >> 1;
        S3[D] <<= 1;
      }
    }
  }
  return IDX2O_OK;
}

/* GetLinearBrick, Core/idx2Lookup.cpp:12-24 */
static u64 linear_brick(const ofile_t* F, int L, const int b3in[3])
{
  u64 r = 0;
  int b3[3] = {b3in[0], b3in[1], b3in[2]};
  int n = F->blen[L];
  for (int I = n - 1; I >= 0; --I) {
    int D = F->bstr[L][I] - 'X';
    r |= ((u64)(b3[D] & 1)) << (n - I - 1);
    b3[D] >>= 1;
  }
  return r;
}

/* GetAddress / GetChunkAddress / GetFileAddress, Core/idx2Lookup.h:15-59 */
static u64 get_address(u64 brick, int shift, int level, int sb, int bp)
{
  return ((u64)level << 60) + ((brick >> shift) << 18) + ((u64)(sb << 12)) + ((u64)bp & 0xFFF);
}
static u64 chunk_address(const ofile_t* F, u64 brick, int L, int sb, int bpkey)
{
  return get_address(brick, log2ceil(F->bpc[L]), L, sb, bpkey);
}
static u64 file_address(const ofile_t* F, u64 brick, int L)
{
  return get_address(brick, log2ceil(F->bpf[L]), L, 0, 0);
}
/* ConstructFilePath, Core/idx2Lookup.cpp:77-107 */
static void file_path(const ofile_t* F, u64 brick, int L, char* out, size_t cap)
{
  int n = snprintf(out, cap, "%s/%s/%s/", F->dir, F->name, F->field);
  n += snprintf(out + n, cap - (size_t)n, "/L%02x", L);
  for (int depth = 0; depth + 1 < F->fdd_len[L]; ++depth) {
    int bitlen = 64 - F->blen[L] + F->fdd[L][depth];
    int shift = 64 - bitlen;
    u64 mask = shift >= 64 ? 0 : (~0ULL << shift);
    u64 first = shift >= 64 ? 0 : ((brick & mask) >> shift);
    n += snprintf(out + n, cap - (size_t)n, "/B%llx", (unsigned long long)first);
    brick <<= F->fdd[L][depth];
  }
  snprintf(out + n, cap - (size_t)n, ".bin");
}

static void mkdirs(const char* path)
{
  char tmp[1024];
  snprintf(tmp, sizeof tmp, "%s", path);
  for (char* p = tmp + 1; *p; ++p) {
    if (*p == '/') { *p = 0; mkdir(tmp, 0755); *p = '/'; }
  }
  mkdir(tmp, 0755);
}
static void parent_dir(const char* path, char* out)
{
  strcpy(out, path);
  char* s = strrchr(out, '/');
  if (s) *s = 0; else strcpy(out, ".");
}

/* ------------------------------------------------------------------------ */
/* CDF 5/3 lifting (Core/Wavelet.h:172-322)                                  */
/* ------------------------------------------------------------------------ */
static i64 vidx(const int n3[3], int x, int y, int z) { return (i64)z * n3[0] * n3[1] + (i64)y * n3[0] + x; }

/* other-axis mapping of the lifting macros: line axis d, inner loop axis a
 * ("y" of the macro), outer loop axis b ("z"):  X -> (Y, Z), Y -> (X, Z),
 * Z -> (X, Y), Wavelet.h:317-322 */
static void lift_axes(int d, int* a, int* b)
{
  if (d == 0) { *a = 1; *b = 2; }
  else if (d == 1) { *a = 0; *b = 2; }
  else { *a = 0; *b = 1; }
}

/* FLiftCdf53X/Y/Z with lift_option::Normal, Wavelet.h:177-246 */
void idx2o_flift(int d, const int P[3], const int D[3], const int S[3], const int M[3],
                 const int N[3], double* F)
{
  if (D[d] == 1) return;
  int a, b;
  lift_axes(d, &a, &b);
  int x0 = imin(P[d] + S[d] * D[d], M[d]);
  int x1 = imin(P[d] + S[d] * (D[d] - 1), M[d]);
  int x2 = P[d] + S[d] * (D[d] - 2);
  int x3 = P[d] + S[d] * (D[d] - 3);
  int ext = (D[d] % 2) == 0;
  int c[3];
  for (int zc = P[b]; zc < P[b] + S[b] * D[b]; zc += S[b]) {
    c[b] = imin(zc, M[b]);
    for (int yc = P[a]; yc < P[a] + S[a] * D[a]; yc += S[a]) {
      c[a] = imin(yc, M[a]);
#define AT(xx) F[(c[d] = (xx), vidx(N, c[0], c[1], c[2]))]
      if (ext) {
        double A = AT(x2);
        double B = AT(x1);
        AT(x0) = 2 * B - A;
      }
      for (int x = P[d] + S[d]; x < P[d] + S[d] * (D[d] - 2); x += 2 * S[d]) {
        double l = AT(x - S[d]);
        double r = AT(x + S[d]);
        double v = AT(x);
        AT(x) = v - (l + r) / 2;
      }
      if (!ext) {
        double l = AT(x1);
        double r = AT(x3);
        double v = AT(x2);
        AT(x2) = v - (l + r) / 2;
      } else if (x1 < M[d]) {
        AT(x1) = 0;
      }
      for (int x = P[d] + S[d]; x < P[d] + S[d] * (D[d] - 2); x += 2 * S[d]) {
        double v = AT(x);
        double t = AT(x - S[d]);
        AT(x - S[d]) = t + v / 4;
        t = AT(x + S[d]);
        AT(x + S[d]) = t + v / 4;
      }
      if (!ext) {
        double v = AT(x2);
        double t = AT(x3);
        AT(x3) = t + v / 4;
        t = AT(x1);
        AT(x1) = t + v / 4;
      }
#undef AT
    }
  }
}

/* ILiftCdf53X/Y/Z with lift_option::Normal, Wavelet.h:250-310 */
void idx2o_ilift(int d, const int P[3], const int D[3], const int S[3], const int M[3],
                 const int N[3], double* F)
{
  if (D[d] == 1) return;
  int a, b;
  lift_axes(d, &a, &b);
  int x0 = imin(P[d] + S[d] * D[d], M[d]);
  int x1 = imin(P[d] + S[d] * (D[d] - 1), M[d]);
  int x2 = P[d] + S[d] * (D[d] - 2);
  int x3 = P[d] + S[d] * (D[d] - 3);
  int ext = (D[d] % 2) == 0;
  int c[3];
  for (int zc = P[b]; zc < P[b] + S[b] * D[b]; zc += S[b]) {
    c[b] = imin(zc, M[b]);
    for (int yc = P[a]; yc < P[a] + S[a] * D[a]; yc += S[a]) {
      c[a] = imin(yc, M[a]);
#define AT(xx) F[(c[d] = (xx), vidx(N, c[0], c[1], c[2]))]
      for (int x = P[d] + S[d]; x < P[d] + S[d] * (D[d] - 2); x += 2 * S[d]) {
        double v = AT(x);
        double t = AT(x - S[d]);
        AT(x - S[d]) = t - v / 4;
        t = AT(x + S[d]);
        AT(x + S[d]) = t - v / 4;
      }
      if (!ext) {
        double v = AT(x2);
        double t = AT(x3);
        AT(x3) = t - v / 4;
        t = AT(x1);
        AT(x1) = t - v / 4;
      } else {
        double A = AT(x0);
        double B = AT(x2);
        AT(x1) = (A + B) / 2;
      }
      for (int x = P[d] + S[d]; x < P[d] + S[d] * (D[d] - 2); x += 2 * S[d]) {
        double l = AT(x - S[d]);
        double r = AT(x + S[d]);
        double v = AT(x);
        AT(x) = v + (l + r) / 2;
      }
      if (!ext) {
        double l = AT(x1);
        double r = AT(x3);
        double v = AT(x2);
        AT(x2) = v + (l + r) / 2;
      }
#undef AT
    }
  }
}

/* Runtime-order multi-level forward transform without normalisation,
 * Wavelet.cpp:209-269 (used by ExtrapolateCdf53). */
static void forward_runtime(const int dims3[3], const int M3[3], int nlevels, u64 order,
                            const int N3[3], double* vol)
{
  int level = 0;
  u64 prev = order;
  int D3[3] = {dims3[0], dims3[1], dims3[2]}, R3[3], S3[3] = {1, 1, 1}, zero[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) R3[d] = D3[d];
  grid_t G = mkgrid(zero, dims3, S3);
  while (level < nlevels) {
    int D = (int)(order & 3);
    order >>= 2;
    if (D == 3) {
      if (order == 3) order = prev; else prev = order;
      for (int d = 0; d < 3; ++d) { G.s[d] = S3[d]; G.d[d] = D3[d]; R3[d] = D3[d]; }
      ++level;
    } else {
      idx2o_flift(D, G.f, G.d, G.s, M3, N3, vol);
      R3[D] = D3[D] + (D3[D] % 2 == 0);
      for (int d = 0; d < 3; ++d) G.d[d] = R3[d];
      D3[D] = (R3[D] + 1) >> 1;
      S3[D] <<= 1;
    }
  }
}

/* Runtime-order multi-level inverse transform, Wavelet.cpp:303-404. */
static void inverse_runtime(const int dims3[3], const int M3[3], int nlevels, u64 order,
                            const int N3[3], double* vol)
{
  grid_t stack[64];
  int axes[64], n = 0;
  int level = 0;
  u64 prev = order;
  int D3[3] = {dims3[0], dims3[1], dims3[2]}, R3[3], S3[3] = {1, 1, 1}, zero[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) R3[d] = D3[d];
  grid_t G = mkgrid(zero, dims3, S3);
  while (level < nlevels) {
    int D = (int)(order & 3);
    order >>= 2;
    if (D == 3) {
      if (order == 3) order = prev; else prev = order;
      for (int d = 0; d < 3; ++d) { G.s[d] = S3[d]; G.d[d] = D3[d]; R3[d] = D3[d]; }
      ++level;
    } else {
      stack[n] = G;
      axes[n++] = D;
      R3[D] = D3[D] + (D3[D] % 2 == 0);
      for (int d = 0; d < 3; ++d) G.d[d] = R3[d];
      D3[D] = (R3[D] + 1) >> 1;
      S3[D] <<= 1;
    }
  }
  while (n-- > 0) idx2o_ilift(axes[n], stack[n].f, stack[n].d, stack[n].s, M3, N3, vol);
}

static u64 order_for_dims(const int n3[3])
{
  char s[8];
  int j = 0;
  for (int d = 0; d < 3; ++d) if (n3[d] > 1) s[j++] = (char)('X' + d);
  s[j++] = '+'; s[j++] = '+'; s[j] = 0;
  return encode_order(s);
}

/* ExtrapolateCdf53, Wavelet.cpp:67-77 */
static void extrapolate(const int data_dims[3], const int N3[3], u64 order, double* vol)
{
  int M3[3];
  for (int d = 0; d < 3; ++d) M3[d] = N3[d] == 1 ? 1 : N3[d] - 1;
  int nlevels = log2floor(imax(imax(N3[0], N3[1]), N3[2]));
  forward_runtime(data_dims, M3, nlevels, order, N3, vol);
  inverse_runtime(N3, M3, nlevels, order, N3, vol);
}
void idx2o_extrapolate(const int data_dims[3], const int n3[3], double* vol)
{
  extrapolate(data_dims, n3, order_for_dims(n3), vol);
}

/* GetCdf53NormsFast<16>, Wavelet.h:19-32 */
void idx2o_cdf53_norms(double scal[16], double wave[16])
{
  double num1 = 3, num2 = 23;
  for (int i = 0; i < 16; ++i) {
    scal[i] = sqrt(num1 / (1 << (i + 1)));
    num1 = num1 * 4 - 1;
    wave[i] = sqrt(num2 / (1 << (i + 5)));
    num2 = num2 * 4 - 33;
  }
}

/* subband norm factor W = (Wx*Wy)*Wz, Wavelet.cpp:111-130 */
static double subband_weight(const ofile_t* F, int sb, int iter)
{
  double scal[16], wave[16];
  idx2o_cdf53_norms(scal, wave);
  double w[3];
  for (int d = 0; d < 3; ++d) {
    if (F->brick_ext[d] == 1) w[d] = 1;
    else w[d] = F->sb_lh[sb][d] == 0 ? scal[iter + F->sb_level3rev[sb][d] - 1]
                                     : wave[iter + F->sb_level3rev[sb][d]];
  }
  return w[0] * w[1] * w[2];
}

/* ForwardCdf53 (subband version), Wavelet.cpp:80-138 */
static void forward_brick(const ofile_t* F, int iter, int coarsest, double* vol)
{
  for (int i = 0; i < F->td_n; ++i)
    idx2o_flift(F->td_axis[i], F->td_grid[i].f, F->td_grid[i].d, F->td_grid[i].s, F->brick_ext,
                F->brick_ext, vol);
  for (int sb = 0; sb < F->nsb; ++sb) {
    if (sb == 0 && !coarsest) continue;
    double W = subband_weight(F, sb, iter);
    const grid_t* g = &F->sb[sb];
    for (int z = 0; z < g->d[2]; ++z)
      for (int y = 0; y < g->d[1]; ++y)
        for (int x = 0; x < g->d[0]; ++x) {
          i64 i = vidx(F->brick_ext, g->f[0] + g->s[0] * x, g->f[1] + g->s[1] * y, g->f[2] + g->s[2] * z);
          vol[i] = vol[i] * W;
        }
  }
}

/* InverseCdf53 (subband version), Wavelet.cpp:143-203 */
static void inverse_brick(const ofile_t* F, int iter, int coarsest, double* vol)
{
  for (int sb = 0; sb < F->nsb; ++sb) {
    if (sb == 0 && !coarsest) continue;
    double W = 1.0 / subband_weight(F, sb, iter);
    const grid_t* g = &F->sb[sb];
    for (int z = 0; z < g->d[2]; ++z)
      for (int y = 0; y < g->d[1]; ++y)
        for (int x = 0; x < g->d[0]; ++x) {
          i64 i = vidx(F->brick_ext, g->f[0] + g->s[0] * x, g->f[1] + g->s[1] * y, g->f[2] + g->s[2] * z);
          vol[i] = vol[i] * W;
        }
  }
  for (int i = F->td_n - 1; i >= 0; --i)
    idx2o_ilift(F->td_axis[i], F->td_grid[i].f, F->td_grid[i].d, F->td_grid[i].s, F->brick_ext,
                F->brick_ext, vol);
}

static void test_file_for_brick(ofile_t* F, const int n3[3])
{
  memset(F, 0, sizeof *F);
  for (int d = 0; d < 3; ++d) { F->brick[d] = n3[d] > 1 ? n3[d] - 1 : 1; F->dims[d] = F->brick[d]; }
  F->nlevels = 1; F->bpc_in = 4096; F->cpf_in = 64; F->files_per_dir = 64;
  finalize(F);
}
void idx2o_forward_brick(const int n3[3], int level, int coarsest, double* vol)
{
  ofile_t F;
  test_file_for_brick(&F, n3);
  forward_brick(&F, level, coarsest, vol);
}
void idx2o_inverse_brick(const int n3[3], int level, int coarsest, double* vol)
{
  ofile_t F;
  test_file_for_brick(&F, n3);
  inverse_brick(&F, level, coarsest, vol);
}

/* ------------------------------------------------------------------------ */
/* zfp block stage (Core/Zfp.h:135-605, Core/Function.h:97-149)             */
/* ------------------------------------------------------------------------ */
/* Coefficient order of a 4x4x4 block: by i+j+k, then i^2+j^2+k^2, ties as
 * listed (zfp's published perm_3 table, Zfp.h:414-493).  Entries are (i,j,k). */
static const signed char PERM3_IJK[64][3] = {
  {0,0,0},
  {1,0,0},{0,1,0},{0,0,1},
  {0,1,1},{1,0,1},{1,1,0},{2,0,0},{0,2,0},{0,0,2},
  {1,1,1},{2,1,0},{2,0,1},{0,2,1},{1,2,0},{1,0,2},{0,1,2},{3,0,0},{0,3,0},{0,0,3},
  {2,1,1},{1,2,1},{1,1,2},{0,2,2},{2,0,2},{2,2,0},{3,1,0},{3,0,1},{0,3,1},{1,3,0},{1,0,3},{0,1,3},
  {1,2,2},{2,1,2},{2,2,1},{3,1,1},{1,3,1},{1,1,3},{3,2,0},{3,0,2},{0,3,2},{2,3,0},{2,0,3},{0,2,3},
  {2,2,2},{3,2,1},{3,1,2},{1,3,2},{2,3,1},{2,1,3},{1,2,3},{0,3,3},{3,0,3},{3,3,0},
  {3,2,2},{2,3,2},{2,2,3},{1,3,3},{3,1,3},{3,3,1},
  {2,3,3},{3,2,3},{3,3,2},
  {3,3,3},
};
/* 4x4 order (Zfp.h:497-521), entries (i,j). */
static const signed char PERM2_IJ[16][2] = {
  {0,0},{1,0},{0,1},{1,1},{2,0},{0,2},{2,1},{1,2},{3,0},{0,3},{2,2},{3,1},{1,3},{3,2},{2,3},{3,3},
};
static int perm_index(int ndims, int i)
{
  if (ndims == 3) return PERM3_IJK[i][0] + 4 * PERM3_IJK[i][1] + 16 * PERM3_IJK[i][2];
  if (ndims == 2) return PERM2_IJ[i][0] + 4 * PERM2_IJ[i][1];
  return i;
}

#define NBMASK 0xaaaaaaaaaaaaaaaaULL

/* FLift / ILift, Zfp.h:144-204 (arithmetic shifts on i64) */
static void zfp_flift(i64* p, int s)
{
  i64 x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  x += w; x >>= 1; w -= x;
  z += y; z >>= 1; y -= z;
  x += z; x >>= 1; z -= x;
  w += y; w >>= 1; y -= w;
  w += y >> 1; y -= w >> 1;
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}
static void zfp_ilift(i64* p, int s)
{
  i64 x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  y += w >> 1; w -= y >> 1;
  y += w; w <<= 1; w -= y;
  z += x; x <<= 1; x -= z;
  y += z; z <<= 1; z -= y;
  w += x; x <<= 1; x -= w;
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}
void idx2o_zfp_lift(i64* p, int fwd)
{
  if (fwd) zfp_flift(p, 1);
  else zfp_ilift(p, 1);
}

/* ForwardZfp(P, D), Zfp.h:261-295 */
static void zfp_forward(i64* p, int D)
{
  if (D == 3) {
    for (int z = 0; z < 4; ++z) for (int y = 0; y < 4; ++y) zfp_flift(p + 4 * y + 16 * z, 1);
    for (int x = 0; x < 4; ++x) for (int z = 0; z < 4; ++z) zfp_flift(p + 16 * z + x, 4);
    for (int y = 0; y < 4; ++y) for (int x = 0; x < 4; ++x) zfp_flift(p + x + 4 * y, 16);
  } else if (D == 2) {
    for (int y = 0; y < 4; ++y) zfp_flift(p + 4 * y, 1);
    for (int x = 0; x < 4; ++x) zfp_flift(p + x, 4);
  } else if (D == 1) {
    zfp_flift(p, 1);
  }
}
/* InverseZfp(P, D), Zfp.h:311-345 */
static void zfp_inverse(i64* p, int D)
{
  if (D == 3) {
    for (int y = 0; y < 4; ++y) for (int x = 0; x < 4; ++x) zfp_ilift(p + x + 4 * y, 16);
    for (int x = 0; x < 4; ++x) for (int z = 0; z < 4; ++z) zfp_ilift(p + 16 * z + x, 4);
    for (int z = 0; z < 4; ++z) for (int y = 0; y < 4; ++y) zfp_ilift(p + 4 * y + 16 * z, 1);
  } else if (D == 2) {
    for (int x = 0; x < 4; ++x) zfp_ilift(p + x, 4);
    for (int y = 0; y < 4; ++y) zfp_ilift(p + 4 * y, 1);
  } else if (D == 1) {
    zfp_ilift(p, 1);
  }
}
/* Exponent<t>, Core/Math.h:523-534 */
static int exponent_f64(double v)
{
  if (v != 0) { int e; frexp(v, &e); return imax(e, 1 - 1023); }
  return -1023;
}
static int exponent_f32(float v)
{
  if (v != 0) { int e; frexpf(v, &e); return imax(e, 1 - 127); }
  return -127;
}

/* QuantizeF32/F64 + ForwardZfp + ForwardShuffle (idx2Encode.cpp:363-367,
 * Function.h:58-89, Zfp.h:534-559).  in: NVals doubles in gather order. */
int idx2o_block_forward(const double* in, int ndims, int is_f64, u64* out)
{
  int nvals = 1 << (2 * ndims);
  int prec = 64 - 1 - ndims;
  double maxabs = 0;
  for (int i = 0; i < nvals; ++i) { double a = fabs(in[i]); maxabs = maxabs < a ? a : maxabs; }
  int emax = is_f64 ? exponent_f64(maxabs) : exponent_f32((float)maxabs);
  double scale = ldexp(1, prec - 1 - emax);
  i64 q[64];
  for (int i = 0; i < nvals; ++i) {
    double v = scale * in[i];
    /* (i64) truncation; out-of-range values only arise for all-zero f64
     * blocks (inf*0) whose planes are never coded. */
    q[i] = (v == v && fabs(v) < 9.2e18) ? (i64)v : INT64_MIN;
  }
  zfp_forward(q, ndims);
  for (int i = 0; i < nvals; ++i) out[i] = ((u64)q[perm_index(ndims, i)] + NBMASK) ^ NBMASK;
  return emax;
}

/* InverseShuffle + InverseZfp + Dequantize (idx2Decode.cpp:213-215) */
void idx2o_block_inverse(const u64* in, int ndims, int emax, double* out)
{
  int nvals = 1 << (2 * ndims);
  int prec = 64 - 1 - ndims;
  i64 q[64];
  for (int i = 0; i < nvals; ++i) q[perm_index(ndims, i)] = (i64)((in[i] ^ NBMASK) - NBMASK);
  zfp_inverse(q, ndims);
  double scale = 1.0 / ldexp(1, prec - 1 - emax);
  for (int i = 0; i < nvals; ++i) out[i] = scale * (double)q[i];
}

/* Embedded coder, the variant in use (Zfp.h:769-813): one bit plane. */
static void encode_plane(const u64* blk, int nvals, int bp, int* Np, bw_t* w)
{
  u64 x = 0;
  int n = *Np;
  for (int i = 0; i < nvals; ++i) x += ((blk[i] >> bp) & 1u) << i;
  if (n > 0) {
    bw_put(w, x, n);
    x = n >= 64 ? 0 : x >> n;
  }
  while (n < nvals) {
    int g = x != 0;
    bw_put(w, (u64)g, 1);
    if (!g) break;
    while (n + 1 < nvals) {
      int bit = (int)(x & 1u);
      bw_put(w, (u64)bit, 1);
      if (bit) break;
      x >>= 1;
      ++n;
    }
    x >>= 1;
    ++n;
  }
  *Np = n;
}

/* Decode, the variant in use (Zfp.h:988-1059). deposit=0 is BypassDecode. */
static void decode_plane(u64* blk, int nvals, int bp, int* Np, br_t* r, int deposit)
{
  int n = *Np;
  u64 x = n > 0 ? br_get(r, n) : 0;
  while (n < nvals) {
    if (br_get(r, 1)) {
      while (n + 1 < nvals) {
        if (br_get(r, 1)) break;
        ++n;
      }
      x += 1ULL << (n++);
    } else {
      break;
    }
  }
  if (deposit)
    for (int i = 0; x; ++i, x >>= 1) blk[i] += (x & 1u) << bp;
  *Np = n;
}

int idx2o_encode_plane(const u64* block, int nvals, int bp, int* n, u8* out, i64 bitpos)
{
  bw_t w = {0, 0, 0};
  encode_plane(block, nvals, bp, n, &w);
  for (i64 i = 0; i < w.nbits; ++i) {
    int bit = (w.buf[i >> 3] >> (i & 7)) & 1;
    i64 p = bitpos + i;
    out[p >> 3] = (u8)((out[p >> 3] & ~(1u << (p & 7))) | ((unsigned)bit << (p & 7)));
  }
  int nb = (int)w.nbits;
  bw_free(&w);
  return nb;
}
int idx2o_decode_plane(u64* block, int nvals, int bp, int* n, const u8* in, i64 bitpos)
{
  br_t r = {in, INT64_MAX / 16, bitpos};
  decode_plane(block, nvals, bp, n, &r, 1);
  return (int)(r.pos - bitpos);
}

/* ------------------------------------------------------------------------ */
/* Morton block lists (Core/BitOps.h:249-285, idx2Common.h:23-30)           */
/* ------------------------------------------------------------------------ */
static u32 split_by2(u32 x)
{
  x &= 0x000003ff;
  x = (x ^ (x << 16)) & 0x030000ff;
  x = (x ^ (x << 8)) & 0x0300f00f;
  x = (x ^ (x << 4)) & 0x030c30c3;
  x = (x ^ (x << 2)) & 0x09249249;
  return x;
}
static u32 morton3(u32 x, u32 y, u32 z) { return split_by2(x) | (split_by2(y) << 1) | (split_by2(z) << 2); }

/* at most 17^3 blocks per subband: bricks of up to 128 samples per axis */
#define OMAXBLOCKS 4913
typedef struct { int n; int org[OMAXBLOCKS][3]; int dim[OMAXBLOCKS][3]; } blocklist_t;

/* Blocks of a subband in Morton order (Zfp loop of idx2Encode.cpp:333-350) */
static void block_list(const int sbdims[3], blocklist_t* L)
{
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = (sbdims[d] + 3) / 4;
  u32 last = morton3((u32)nb[0] - 1, (u32)nb[1] - 1, (u32)nb[2] - 1);
  L->n = 0;
  for (u32 m = 0; m <= last; ++m) {
    u32 x = 0, y = 0, z = 0;
    for (int b = 0; b < 10; ++b) {
      x |= ((m >> (3 * b)) & 1u) << b;
      y |= ((m >> (3 * b + 1)) & 1u) << b;
      z |= ((m >> (3 * b + 2)) & 1u) << b;
    }
    if ((int)x >= nb[0] || (int)y >= nb[1] || (int)z >= nb[2]) continue;
    int o[3] = {(int)x * 4, (int)y * 4, (int)z * 4};
    for (int d = 0; d < 3; ++d) {
      L->org[L->n][d] = o[d];
      L->dim[L->n][d] = imin(4, sbdims[d] - o[d]);
    }
    L->n++;
  }
}
static int numdims(const int d[3]) { return (d[0] > 1) + (d[1] > 1) + (d[2] > 1); }

/* ------------------------------------------------------------------------ */
/* encoder (Core/idx2Encode.cpp, Core/idx2Write.cpp)                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  bw_t deltas, sizes, bricks, block;
  u64 last_chunk, last_brick;
  int nbricks;
  int used;
} chan_t;

typedef struct {
  bw_t block_exp, brick_exp;
  u64 last_chunk, last_brick;
  int used;
} subchan_t;

typedef struct {
  u64 id;
  char path[1024];
  bw_t data;      /* bit-plane chunks in emission order */
  bw_t sizes;     /* varbyte chunk sizes */
  u64* addrs; int naddrs, caddrs;
  int has_bp;
  bw_t exp_data;  /* exponent chunks */
  bw_t exp_sizes;
  u64* eaddrs; int neaddrs, ceaddrs;
  int has_exp;
} ofile_out_t;

typedef struct { double mn, mx, s, ssq; i64 n; } ostat_t;
static void stat_add(ostat_t* s, double v)
{
  if (s->n == 0) { s->mn = DBL_MAX; s->mx = -DBL_MAX; }
  s->mn = s->mn < v ? s->mn : v;
  s->mx = s->mx > v ? s->mx : v;
  s->s += v; s->n++; s->ssq += v * v;
}
static double stat_avg(const ostat_t* s) { return s->s / s->n; }
static double stat_sd(const ostat_t* s) { double a = stat_avg(s); return sqrt(s->ssq / s->n - a * a); }

typedef struct {
  int level; u64 brick; int nchildren_max, nchildren_done;
  int ext_local[3];
  double* vol;
} brickvol_t;

#define NCHAN_SLOTS (4096 * 128)

typedef struct {
  ofile_t* F;
  chan_t** chans;      /* dense: bpkey*128 + sb*16 + level */
  subchan_t sub[MAXL][8];
  ofile_out_t* files; int nfiles, cfiles;
  brickvol_t* pool; int npool, cpool;
  int cur_level, cur_sb;
  u64 brick[MAXL];
  int bricks3[MAXL][3];
  /* per subband pass */
  struct { u32 block; int bpkey; } lastsig[256]; int nlastsig;
  short sbexps[OMAXBLOCKS]; int nsbexps;
  ostat_t st_bpchunk, st_deltas, st_bsizes, st_chunksizes, st_caddr_u, st_caddr_c;
  ostat_t st_exp_u, st_exp_c, st_expsizes, st_eaddr_u, st_eaddr_c;
  int err;
} enc_t;

static ofile_out_t* get_file(enc_t* E, u64 brick, int L)
{
  u64 id = file_address(E->F, brick, L);
  for (int i = 0; i < E->nfiles; ++i) if (E->files[i].id == id) return &E->files[i];
  if (E->nfiles == E->cfiles) {
    E->cfiles = E->cfiles ? E->cfiles * 2 : 8;
    E->files = (ofile_out_t*)realloc(E->files, sizeof(ofile_out_t) * (size_t)E->cfiles);
  }
  ofile_out_t* f = &E->files[E->nfiles++];
  memset(f, 0, sizeof *f);
  f->id = id;
  file_path(E->F, brick, L, f->path, sizeof f->path);
  return f;
}
static void push_u64(u64** a, int* n, int* c, u64 v)
{
  if (*n == *c) { *c = *c ? *c * 2 : 64; *a = (u64*)realloc(*a, sizeof(u64) * (size_t)*c); }
  (*a)[(*n)++] = v;
}

/* CompressBufZstd, idx2Encode.cpp:23-37 */
static void zstd_into(const void* src, i64 n, bw_t* out)
{
  bw_rewind(out);
  if (n == 0) return;
  size_t cap = ZSTD_compressBound((size_t)n);
  bw_reserve(out, (i64)cap + 8);
  size_t c = ZSTD_compress(out->buf, cap, src, (size_t)n, 1);
  if (ZSTD_isError(c)) { fprintf(stderr, "zstd failed\n"); exit(1); }
  out->nbits = (i64)c * 8;
}

/* WriteChunk, idx2Write.cpp:220-290 */
static void write_chunk(enc_t* E, chan_t* C, int L, int sb, int bpkey)
{
  stat_add(&E->st_deltas, (double)bw_size(&C->deltas));
  stat_add(&E->st_bsizes, (double)bw_size(&C->sizes));
  bw_t cs = {0, 0, 0};
  bw_varbyte(&cs, (u64)C->nbricks);
  bw_append_stream(&cs, &C->deltas);
  bw_append_stream(&cs, &C->sizes);
  bw_append_stream(&cs, &C->bricks);
  stat_add(&E->st_bpchunk, (double)bw_size(&cs));
  bw_rewind(&C->deltas); bw_rewind(&C->sizes); bw_rewind(&C->bricks);
  ofile_out_t* f = get_file(E, C->last_brick, L);
  f->has_bp = 1;
  bw_append_stream(&f->data, &cs);
  bw_varbyte(&f->sizes, (u64)bw_size(&cs));
  push_u64(&f->addrs, &f->naddrs, &f->caddrs, chunk_address(E->F, C->last_brick, L, sb, bpkey));
  bw_free(&cs);
}

/* WriteChunkExponents, idx2Write.cpp:37-92 */
static void write_chunk_exponents(enc_t* E, subchan_t* S, int L, int sb)
{
  stat_add(&E->st_exp_u, (double)bw_size(&S->brick_exp));
  bw_t comp = {0, 0, 0};
  zstd_into(S->brick_exp.buf, bw_size(&S->brick_exp), &comp);
  stat_add(&E->st_exp_c, (double)bw_size(&comp));
  bw_rewind(&S->brick_exp);
  ofile_out_t* f = get_file(E, S->last_brick, L);
  f->has_exp = 1;
  bw_varbyte(&f->exp_sizes, (u64)bw_size(&comp));
  bw_append_stream(&f->exp_data, &comp);
  push_u64(&f->eaddrs, &f->neaddrs, &f->ceaddrs, chunk_address(E->F, S->last_brick, L, sb, 0));
  bw_free(&comp);
}

static chan_t* get_chan(enc_t* E, int bpkey, int sb, int L)
{
  int idx = bpkey * 128 + sb * 16 + L;
  if (bpkey < 0 || bpkey >= 4096) { E->err = 1; idx = 0; }
  if (!E->chans[idx]) {
    E->chans[idx] = (chan_t*)calloc(1, sizeof(chan_t));
  }
  E->chans[idx]->used = 1;
  return E->chans[idx];
}

static double brick_at(const ofile_t* F, const double* vol, const grid_t* g, int x, int y, int z)
{
  return vol[vidx(F->brick_ext, g->f[0] + g->s[0] * x, g->f[1] + g->s[1] * y, g->f[2] + g->s[2] * z)];
}

/* EncodeSubbandBlocks + exponents + metadata (idx2Encode.cpp:141-272) */
static void encode_subband(enc_t* E, const grid_t* sbg, const double* vol)
{
  ofile_t* F = E->F;
  int L = E->cur_level, sb = E->cur_sb;
  u64 brick = E->brick[L];
  blocklist_t bl;
  block_list(sbg->d, &bl);
  E->nlastsig = 0;
  E->nsbexps = 0;
  int endbp = F->dtype ? 64 : 32;
  int etol = exponent_f64(F->tolerance);
  /* pass 1: compress blocks (idx2Encode.cpp:157-255) */
  for (int bi = 0; bi < bl.n; ++bi) {
    int* o = bl.org[bi];
    int* bd = bl.dim[bi];
    int coded_next = sb == 0 && L + 1 < F->nlevels && bd[0] == 4 && bd[1] == 4 && bd[2] == 4;
    if (coded_next) continue;
    int nd = numdims(bd);
    int nvals = 1 << (2 * nd);
    double vals[64];
    int j = 0;
    for (int z = 0; z < bd[2]; ++z)
      for (int y = 0; y < bd[1]; ++y)
        for (int x = 0; x < bd[0]; ++x) vals[j++] = brick_at(F, vol, sbg, o[0] + x, o[1] + y, o[2] + z);
    for (; j < 64; ++j) vals[j] = 0;
    u64 u[64];
    int emax = idx2o_block_forward(vals, nd, F->dtype, u);
    E->sbexps[E->nsbexps++] = (short)emax;
    int n = 0;
    u32 block_id = (u32)bi; /* rank in Morton order stands in for the code */
    for (int bp = 63; bp >= 64 - endbp; --bp) {
      int realbp = bp + emax;
      int bpkey = (realbp + BP_BIAS) / F->bppc;
      int too_high = 64 - 6 > realbp - etol + 1;
      if (too_high && (realbp + BP_BIAS) % F->bppc == 0) break;
      chan_t* C = get_chan(E, bpkey, sb, L);
      int i = 0;
      for (; i < E->nlastsig; ++i)
        if (E->lastsig[i].bpkey == bpkey) { E->lastsig[i].block = block_id; break; }
      if (i == E->nlastsig) {
        int not_empty = bw_size(&C->bricks) > 0;
        int new_chunk = brick >= (C->last_chunk + 1) * (u64)F->bpc[L];
        if (new_chunk) {
          if (not_empty) write_chunk(E, C, L, sb, bpkey);
          C->nbricks = 0;
          C->last_chunk = brick >> log2ceil(F->bpc[L]);
        }
        E->lastsig[E->nlastsig].block = block_id;
        E->lastsig[E->nlastsig].bpkey = bpkey;
        E->nlastsig++;
      }
      encode_plane(u, nvals, bp, &n, &C->block);
    }
  }
  /* EncodeBrickSubbandExponents, idx2Encode.cpp:41-80 */
  subchan_t* S = &E->sub[L][sb];
  if (!S->used) { S->used = 1; }
  if (brick >= (S->last_chunk + 1) * (u64)F->bpc[L]) {
    write_chunk_exponents(E, S, L, sb);
    S->last_chunk = brick >> log2ceil(F->bpc[L]);
  }
  for (int i = 0; i < E->nsbexps; ++i) {
    int s = E->sbexps[i] + (F->dtype ? 1023 : 127);
    bw_put(&S->block_exp, (u64)(s & 0xFFFF), F->dtype ? 16 : 8);
  }
  bw_append_stream(&S->brick_exp, &S->block_exp);
  bw_rewind(&S->block_exp);
  S->last_brick = brick;
  /* EncodeBrickSubbandMetadata, idx2Encode.cpp:84-137 */
  for (int bi = 0; bi < bl.n; ++bi) {
    int* bd = bl.dim[bi];
    int coded_next = sb == 0 && L + 1 < F->nlevels && bd[0] == 4 && bd[1] == 4 && bd[2] == 4;
    if (coded_next) continue;
    for (int i = 0; i < E->nlastsig; ++i) {
      if ((u32)bi != E->lastsig[i].block) continue;
      chan_t* C = get_chan(E, E->lastsig[i].bpkey, sb, L);
      if (C->nbricks == 0) bw_varbyte(&C->deltas, brick);
      else bw_unary(&C->deltas, (u32)(brick - C->last_brick - 1));
      i64 bs = bw_size(&C->block);
      bw_varbyte(&C->sizes, (u64)bs);
      bw_append_stream(&C->bricks, &C->block);
      bw_rewind(&C->block);
      C->nbricks++;
      C->last_brick = brick;
    }
  }
}

static brickvol_t* pool_find(enc_t* E, int L, u64 b)
{
  for (int i = 0; i < E->npool; ++i)
    if (E->pool[i].vol && E->pool[i].level == L && E->pool[i].brick == b) return &E->pool[i];
  return NULL;
}
static brickvol_t* pool_add(enc_t* E, int L, u64 b)
{
  int slot = -1;
  for (int i = 0; i < E->npool; ++i) if (!E->pool[i].vol) { slot = i; break; }
  if (slot < 0) {
    if (E->npool == E->cpool) {
      E->cpool = E->cpool ? E->cpool * 2 : 16;
      E->pool = (brickvol_t*)realloc(E->pool, sizeof(brickvol_t) * (size_t)E->cpool);
    }
    slot = E->npool++;
  }
  brickvol_t* p = &E->pool[slot];
  memset(p, 0, sizeof *p);
  p->level = L; p->brick = b;
  i64 n = (i64)E->F->brick_ext[0] * E->F->brick_ext[1] * E->F->brick_ext[2];
  p->vol = (double*)calloc((size_t)n, sizeof(double));
  return p;
}

/* EncodeBrick, idx2Encode.cpp:276-340 */
static void encode_brick(enc_t* E, int inc_level)
{
  ofile_t* F = E->F;
  int L = E->cur_level += inc_level;
  u64 brick = E->brick[L];
  brickvol_t* bv = pool_find(E, L, brick);
  double* vol = bv->vol;
  extrapolate(bv->ext_local, F->brick_ext, F->tform_order, vol);
  int coarsest = L + 1 == F->nlevels;
  forward_brick(F, L, coarsest, vol);
  for (int sb = 0; sb < F->nsb; ++sb) {
    const grid_t* S = &F->sb[sb];
    int sbne[3];
    for (int d = 0; d < 3; ++d) sbne[d] = S->d[d] - (S->d[d] > 1);
    int next = L + 1;
    if (sb == 0 && next < F->nlevels) {
      int* b3 = E->bricks3[L];
      int pb3[3];
      for (int d = 0; d < 3; ++d) pb3[d] = E->bricks3[next][d] = b3[d] / F->group[d];
      u64 pb = E->brick[next] = linear_brick(F, next, pb3);
      brickvol_t* pv = pool_find(E, next, pb);
      if (!pv) {
        pv = pool_add(E, next, pb);
        bv = pool_find(E, L, brick); /* pool may have moved */
        vol = bv->vol;
        int nch[3];
        for (int d = 0; d < 3; ++d) {
          int from = pb3[d] * F->group[d];
          int to = imin(from + F->group[d], F->nbricks[L][d]);
          nch[d] = imax(0, to - from);
        }
        pv->nchildren_max = nch[0] * nch[1] * nch[2];
        for (int d = 0; d < 3; ++d) pv->ext_local[d] = nch[d] * sbne[d];
      }
      /* CopyGridExtent<f64,f64>, Volume.h:1177-1196 */
      int lp[3];
      for (int d = 0; d < 3; ++d) lp[d] = b3[d] % F->group[d];
      for (int z = 0; z < sbne[2]; ++z)
        for (int y = 0; y < sbne[1]; ++y)
          for (int x = 0; x < sbne[0]; ++x)
            pv->vol[vidx(F->brick_ext, lp[0] * sbne[0] + x, lp[1] * sbne[1] + y, lp[2] * sbne[2] + z)] =
              brick_at(F, vol, S, x, y, z);
      int last = ++pv->nchildren_done == pv->nchildren_max;
      if (last) {
        encode_brick(E, 1);
        bv = pool_find(E, L, brick);
        vol = bv->vol;
      }
    }
    E->cur_sb = sb;
    encode_subband(E, S, vol);
  }
  bv = pool_find(E, L, brick);
  free(bv->vol);
  bv->vol = NULL;
  E->cur_level -= inc_level;
}

/* idx2_BrickTraverse over a level's bricks, idx2Common.h:415-461: visit
 * bricks of [from, from+dims) inside ext in traversal order, with the
 * position, linear address and rank among bricks inside ext2. */
typedef void (*visit_fn)(void* ctx, const int b3[3], u64 addr, int in_chunk);
static void brick_traverse(u64 order, const int from3[3], const int dims3[3], const int ext_from[3],
                           const int ext_dims[3], const int ext2_from[3], const int ext2_dims[3],
                           visit_fn fn, void* ctx)
{
  typedef struct { u64 ord, prev; int f[3], t[3]; i64 before; int in_chunk; u64 addr; } item;
  item st[256];
  int top = 0;
  st[0].ord = st[0].prev = order;
  for (int d = 0; d < 3; ++d) { st[0].f[d] = from3[d]; st[0].t[d] = from3[d] + (int)nextpow2(dims3[d]); }
  st[0].before = 0; st[0].in_chunk = 0; st[0].addr = 0;
  while (top >= 0) {
    item* T = &st[top];
    int DD = (int)(T->ord & 3);
    T->ord >>= 2;
    if (DD == 3) {
      if (T->ord == 3) T->ord = T->prev; else T->prev = T->ord;
      continue;
    }
    --top;
    if (T->t[0] - T->f[0] == 1 && T->t[1] - T->f[1] == 1 && T->t[2] - T->f[2] == 1) {
      fn(ctx, T->f, T->addr, T->in_chunk);
      continue;
    }
    item A = *T, B = *T;
    A.t[DD] = T->f[DD] + (T->t[DD] - T->f[DD]) / 2;
    B.f[DD] = T->f[DD] + (T->t[DD] - T->f[DD]) / 2;
    /* Crop(Skip, Extent2) dims product */
    i64 p2 = 1;
    for (int d = 0; d < 3; ++d) {
      int lo = imax(A.f[d], ext2_from[d]), hi = imin(A.t[d], ext2_from[d] + ext2_dims[d]);
      p2 *= imax(0, hi - lo);
    }
    B.in_chunk = A.in_chunk + (int)p2;
    A.addr = T->addr;
    i64 pa = (i64)(A.t[0] - A.f[0]) * (A.t[1] - A.f[1]) * (A.t[2] - A.f[2]);
    B.addr = T->addr + (u64)pa;
    int eto[3];
    for (int d = 0; d < 3; ++d) eto[d] = ext_from[d] + ext_dims[d];
    int bok = 1, aok = 1;
    for (int d = 0; d < 3; ++d) {
      if (!(B.f[d] < eto[d] && ext_from[d] < B.t[d])) bok = 0;
      if (!(A.f[d] < eto[d] && ext_from[d] < A.t[d])) aok = 0;
    }
    if (bok) st[++top] = B;
    if (aok) st[++top] = A;
  }
}

typedef struct { enc_t* E; const void* data; } enc_ctx_t;

static void encode_visit(void* vctx, const int b3[3], u64 addr, int in_chunk)
{
  (void)addr; (void)in_chunk;
  enc_ctx_t* c = (enc_ctx_t*)vctx;
  enc_t* E = c->E;
  ofile_t* F = E->F;
  u64 lb = linear_brick(F, 0, b3);
  brickvol_t* bv = pool_add(E, 0, lb);
  int from[3], ext[3];
  for (int d = 0; d < 3; ++d) {
    from[d] = b3[d] * F->brick[d];
    int to = imin(from[d] + F->brick[d], F->dims[d]);
    ext[d] = imax(0, to - from[d]);
    bv->ext_local[d] = ext[d];
  }
  /* CopyExtentExtentMinMax<t,f64>, Volume.h:1123-1149 */
  double mn = DBL_MAX, mx = -DBL_MAX;
  for (int z = 0; z < ext[2]; ++z)
    for (int y = 0; y < ext[1]; ++y)
      for (int x = 0; x < ext[0]; ++x) {
        i64 src = ((i64)(from[2] + z) * F->dims[1] + (from[1] + y)) * F->dims[0] + (from[0] + x);
        double v = F->dtype ? ((const double*)c->data)[src] : (double)((const float*)c->data)[src];
        bv->vol[vidx(F->brick_ext, x, y, z)] = v;
        mn = mn < v ? mn : v;
        mx = mx > v ? mx : v;
      }
  F->vmin = F->vmin < mn ? F->vmin : mn;
  F->vmax = F->vmax > mx ? F->vmax : mx;
  E->cur_level = 0;
  for (int d = 0; d < 3; ++d) E->bricks3[0][d] = b3[d];
  E->brick[0] = lb;
  encode_brick(E, 0);
}

static int write_all(const char* path, const void* p, size_t n, const char* mode)
{
  FILE* fp = fopen(path, mode);
  if (!fp) {
    char par[1024];
    parent_dir(path, par);
    mkdirs(par);
    fp = fopen(path, mode);
    if (!fp) return -1;
  }
  if (n) fwrite(p, 1, n, fp);
  fclose(fp);
  return 0;
}

static void write_meta(const ofile_t* F, const char* path)
{
  FILE* fp = fopen(path, "w");
  if (!fp) return;
  /* WriteMetaFile, idx2Common.cpp:149-179 */
  fprintf(fp, "(\n");
  fprintf(fp, "  (common\n");
  fprintf(fp, "    (type \"Simulation\")\n");
  fprintf(fp, "    (name \"%s\")\n", F->name);
  fprintf(fp, "    (field \"%s\")\n", F->field);
  fprintf(fp, "    (dimensions %d %d %d)\n", F->dims[0], F->dims[1], F->dims[2]);
  fprintf(fp, "    (data-type \"%s\")\n", F->dtype ? "float64" : "float32");
  fprintf(fp, "    (min-max %.20f %.20f)\n", F->vmin, F->vmax);
  fprintf(fp, "    (accuracy %.20f)\n", F->tolerance);
  fprintf(fp, "  )\n");
  fprintf(fp, "  (format\n");
  fprintf(fp, "    (version %d %d)\n", F->version[0], F->version[1]);
  fprintf(fp, "    (brick-size %d %d %d)\n", F->brick[0], F->brick[1], F->brick[2]);
  char tord[128];
  decode_order_plain(F->tform_order, tord);
  fprintf(fp, "    (transform-order \"%s\")\n", tord);
  fprintf(fp, "    (num-levels %d)\n", F->nlevels);
  fprintf(fp, "    (transform-passes-per-levels %d)\n", 1);
  fprintf(fp, "    (bricks-per-chunk %d)\n", F->bpc_in);
  fprintf(fp, "    (chunks-per-file %d)\n", F->cpf_in);
  fprintf(fp, "    (files-per-directory %d)\n", F->files_per_dir);
  fprintf(fp, "    (bit-planes-per-chunk %d)\n", F->bppc);
  fprintf(fp, "  )\n");
  fprintf(fp, ")\n");
  fclose(fp);
}

/* PrintStats, idx2Write.cpp:388-413 */
static void print_stats(const enc_t* E, const char* path)
{
  FILE* fp = fopen(path, "a");
  if (!fp) return;
  fprintf(fp, "\n\n---------------- Statistics ----------------\n\n");
  fprintf(fp, "num chunks = %lld (bit plane) and %lld (exponent) \n", (long long)E->st_bpchunk.n,
          (long long)E->st_exp_c.n);
  fprintf(fp, "bit plane chunk: total = %d avg = %d stddev = %d bytes\n", (int)E->st_bpchunk.s,
          (int)stat_avg(&E->st_bpchunk), (int)stat_sd(&E->st_bpchunk));
  fprintf(fp, "  (metadata) brick deltas = %d bytes\n", (int)E->st_deltas.s);
  fprintf(fp, "  (metadata) brick sizes = %d bytes\n", (int)E->st_bsizes.s);
  fprintf(fp, "  (metadata) chunk sizes = %d bytes\n", (int)E->st_chunksizes.s);
  fprintf(fp, "  (metadata) chunk addresses: uncompressed = %d, compressed = %d bytes\n",
          (int)E->st_caddr_u.s, (int)E->st_caddr_c.s);
  fprintf(fp, "exponent chunk:\n");
  fprintf(fp, "  uncompressed: total = %d bytes\n", (int)E->st_exp_u.s);
  fprintf(fp, "  compressed: total = %d avg = %d stddev = %d bytes\n", (int)E->st_exp_c.s,
          (int)stat_avg(&E->st_exp_c), (int)stat_sd(&E->st_exp_c));
  fprintf(fp, "  (metadata) chunk sizes = %d bytes\n", (int)E->st_expsizes.s);
  fprintf(fp, "  (metadata) chunk addresses: uncompressed = %d, compressed = %d bytes\n",
          (int)E->st_eaddr_u.s, (int)E->st_eaddr_c.s);
  fclose(fp);
}

int idx2o_encode(const char* out_dir, const char* name, const char* field, const void* data,
                 int dtype, const int dims[3], const int brick[3], int nlevels, double tolerance,
                 int bricks_per_chunk, int chunks_per_file, int files_per_dir)
{
  ofile_t* F = (ofile_t*)calloc(1, sizeof(ofile_t));
  snprintf(F->name, sizeof F->name, "%s", name);
  snprintf(F->field, sizeof F->field, "%s", field);
  snprintf(F->dir, sizeof F->dir, "%s", out_dir);
  for (int d = 0; d < 3; ++d) { F->dims[d] = dims[d]; F->brick[d] = brick[d]; }
  F->dtype = dtype;
  F->nlevels = nlevels;
  F->tolerance = tolerance;
  F->bpc_in = bricks_per_chunk;
  F->cpf_in = chunks_per_file;
  F->files_per_dir = files_per_dir;
  F->bppc = 1;
  F->version[0] = 1; F->version[1] = 0;
  F->vmin = DBL_MAX; F->vmax = -DBL_MAX;
  int rc = finalize(F);
  if (rc) { free(F); return rc; }
  for (int d = 0; d < 3; ++d)
    if (F->brick[d] > 1 && F->brick[d] < 8) { free(F); return IDX2O_BRICK_SIZE_NOT_POW2; }
  enc_t* E = (enc_t*)calloc(1, sizeof(enc_t));
  E->F = F;
  E->chans = (chan_t**)calloc(NCHAN_SLOTS, sizeof(chan_t*));
  enc_ctx_t ctx = {E, data};
  int zero[3] = {0, 0, 0};
  brick_traverse(F->border[0], zero, F->nbricks[0], zero, F->nbricks[0], zero, F->nbricks[0],
                 encode_visit, &ctx);
  /* FlushChunks, idx2Write.cpp:316-385 (dense index order == ascending u32 key) */
  for (int i = 0; i < NCHAN_SLOTS; ++i) {
    chan_t* C = E->chans[i];
    if (!C || !C->used) continue;
    int bpkey = i / 128, sb = (i / 16) % 8, L = i % 16;
    write_chunk(E, C, L, sb, bpkey);
  }
  char meta[1200];
  snprintf(meta, sizeof meta, "%s/%s", out_dir, name);
  mkdirs(meta);
  for (int k = 0; k < E->nfiles; ++k) {
    ofile_out_t* f = &E->files[k];
    if (!f->has_bp) continue;
    bw_t buf = {0, 0, 0};
    bw_append_stream(&buf, &f->data);
    stat_add(&E->st_chunksizes, (double)bw_size(&f->sizes));
    bw_append_stream(&buf, &f->sizes);
    int sz = (int)bw_size(&f->sizes);
    bw_append_bytes(&buf, &sz, 4);
    bw_t comp = {0, 0, 0};
    zstd_into(f->addrs, (i64)f->naddrs * 8, &comp);
    bw_append_stream(&buf, &comp);
    sz = (int)bw_size(&comp);
    bw_append_bytes(&buf, &sz, 4);
    sz = f->naddrs;
    bw_append_bytes(&buf, &sz, 4);
    stat_add(&E->st_caddr_u, (double)f->naddrs * 8);
    stat_add(&E->st_caddr_c, (double)bw_size(&comp));
    if (write_all(f->path, buf.buf, (size_t)bw_size(&buf), "ab")) E->err = 1;
    bw_free(&buf); bw_free(&comp);
  }
  /* FlushChunkExponents, idx2Write.cpp:120-204: sub-channels sorted by
   * (level desc, subband asc) */
  for (int L = F->nlevels - 1; L >= 0; --L)
    for (int sb = 0; sb < F->nsb; ++sb)
      if (E->sub[L][sb].used) write_chunk_exponents(E, &E->sub[L][sb], L, sb);
  for (int k = 0; k < E->nfiles; ++k) {
    ofile_out_t* f = &E->files[k];
    if (!f->has_exp) continue;
    bw_t buf = {0, 0, 0};
    int total = 0;
    stat_add(&E->st_expsizes, (double)bw_size(&f->exp_sizes));
    bw_append_stream(&buf, &f->exp_data);
    total += (int)bw_size(&f->exp_data);
    bw_append_stream(&buf, &f->exp_sizes);
    int sz = (int)bw_size(&f->exp_sizes);
    bw_append_bytes(&buf, &sz, 4);
    total += sz + 4;
    stat_add(&E->st_eaddr_u, (double)f->neaddrs * 8);
    bw_t comp = {0, 0, 0};
    zstd_into(f->eaddrs, (i64)f->neaddrs * 8, &comp);
    stat_add(&E->st_eaddr_c, (double)bw_size(&comp));
    bw_append_stream(&buf, &comp);
    sz = (int)bw_size(&comp);
    bw_append_bytes(&buf, &sz, 4);
    total += sz + 4;
    sz = f->neaddrs;
    bw_append_bytes(&buf, &sz, 4);
    total += 4;
    total += 4;
    bw_append_bytes(&buf, &total, 4);
    if (write_all(f->path, buf.buf, (size_t)bw_size(&buf), "ab")) E->err = 1;
    bw_free(&buf); bw_free(&comp);
  }
  snprintf(meta, sizeof meta, "%s/%s/%s.idx2", out_dir, name, field);
  write_meta(F, meta);
  print_stats(E, meta);
  rc = E->err ? IDX2O_SIZE_MISMATCHED : IDX2O_OK;
  for (int i = 0; i < NCHAN_SLOTS; ++i)
    if (E->chans[i]) {
      bw_free(&E->chans[i]->deltas); bw_free(&E->chans[i]->sizes);
      bw_free(&E->chans[i]->bricks); bw_free(&E->chans[i]->block);
      free(E->chans[i]);
    }
  free(E->chans);
  for (int L = 0; L < MAXL; ++L)
    for (int sb = 0; sb < 8; ++sb) { bw_free(&E->sub[L][sb].block_exp); bw_free(&E->sub[L][sb].brick_exp); }
  for (int k = 0; k < E->nfiles; ++k) {
    ofile_out_t* f = &E->files[k];
    bw_free(&f->data); bw_free(&f->sizes); bw_free(&f->exp_data); bw_free(&f->exp_sizes);
    free(f->addrs); free(f->eaddrs);
  }
  free(E->files);
  for (int i = 0; i < E->npool; ++i) free(E->pool[i].vol);
  free(E->pool);
  free(E);
  free(F);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* metadata reader (idx2Common.cpp:182-346): the first list of the .idx2    */
/* ------------------------------------------------------------------------ */
static int read_text(const char* path, char** out, i64* n)
{
  FILE* fp = fopen(path, "rb");
  if (!fp) return IDX2O_FILE_NOT_FOUND;
  fseek(fp, 0, SEEK_END);
  *n = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  *out = (char*)malloc((size_t)*n + 1);
  if (fread(*out, 1, (size_t)*n, fp) != (size_t)*n) { fclose(fp); free(*out); return IDX2O_FILE_NOT_FOUND; }
  (*out)[*n] = 0;
  fclose(fp);
  return 0;
}
static const char* find_key(const char* txt, const char* key)
{
  char pat[64];
  snprintf(pat, sizeof pat, "(%s ", key);
  const char* end = strstr(txt, "Statistics");
  const char* p = strstr(txt, pat);
  if (!p || (end && p > end)) return NULL;
  return p + strlen(pat);
}
static int read_meta(const char* path, ofile_t* F)
{
  char* txt;
  i64 n;
  int rc = read_text(path, &txt, &n);
  if (rc) return rc;
  const char* p;
  if (!(p = find_key(txt, "name")) || sscanf(p, "\"%63[^\"]\"", F->name) != 1) goto bad;
  if (!(p = find_key(txt, "field")) || sscanf(p, "\"%63[^\"]\"", F->field) != 1) goto bad;
  if (!(p = find_key(txt, "dimensions")) || sscanf(p, "%d %d %d", &F->dims[0], &F->dims[1], &F->dims[2]) != 3) goto bad;
  char dt[32];
  if (!(p = find_key(txt, "data-type")) || sscanf(p, "\"%31[^\"]\"", dt) != 1) goto bad;
  F->dtype = strcmp(dt, "float64") == 0 ? 1 : 0;
  if (!(p = find_key(txt, "accuracy")) || sscanf(p, "%lf", &F->tolerance) != 1) goto bad;
  if (!(p = find_key(txt, "brick-size")) || sscanf(p, "%d %d %d", &F->brick[0], &F->brick[1], &F->brick[2]) != 3) goto bad;
  if (!(p = find_key(txt, "num-levels")) || sscanf(p, "%d", &F->nlevels) != 1) goto bad;
  if (!(p = find_key(txt, "bricks-per-chunk")) || sscanf(p, "%d", &F->bpc_in) != 1) goto bad;
  if (!(p = find_key(txt, "chunks-per-file")) || sscanf(p, "%d", &F->cpf_in) != 1) goto bad;
  if (!(p = find_key(txt, "files-per-directory")) || sscanf(p, "%d", &F->files_per_dir) != 1) goto bad;
  if (!(p = find_key(txt, "bit-planes-per-chunk")) || sscanf(p, "%d", &F->bppc) != 1) goto bad;
  F->version[0] = 1;
  free(txt);
  return 0;
bad:
  free(txt);
  return IDX2O_SYNTAX_ERROR;
}

/* ------------------------------------------------------------------------ */
/* reader + decoder (idx2Read.cpp, idx2Decode.cpp)                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  u64 addr;
  i64 off, size; /* in file */
  int parsed;
  int nbricks;
  u64* bricks;
  i64* offs; /* byte offsets of brick streams in data */
  u8* data; i64 ndata;
} ochunk_t;

typedef struct {
  u64 id;
  int loaded;
  int nchunks; ochunk_t* chunks;
  int nexp; u64* eaddrs; i64* eoffs; i64* esizes; u8** edata; i64* edn;
  u8* bytes; i64 nbytes;
} ofile_in_t;

typedef struct {
  ofile_t* F;
  ofile_in_t* files; int nfiles, cfiles;
  double tol;
  char dirbuf[1024];
} dec_t;

static int load_file(dec_t* D, u64 brick, int L, ofile_in_t** out)
{
  u64 id = file_address(D->F, brick, L);
  for (int i = 0; i < D->nfiles; ++i)
    if (D->files[i].id == id) { *out = &D->files[i]; return D->files[i].loaded ? 0 : IDX2O_FILE_NOT_FOUND; }
  if (D->nfiles == D->cfiles) {
    D->cfiles = D->cfiles ? D->cfiles * 2 : 8;
    D->files = (ofile_in_t*)realloc(D->files, sizeof(ofile_in_t) * (size_t)D->cfiles);
  }
  ofile_in_t* f = &D->files[D->nfiles++];
  memset(f, 0, sizeof *f);
  f->id = id;
  *out = f;
  char path[1024];
  file_path(D->F, brick, L, path, sizeof path);
  FILE* fp = fopen(path, "rb");
  if (!fp) return IDX2O_FILE_NOT_FOUND;
  fseek(fp, 0, SEEK_END);
  f->nbytes = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  f->bytes = (u8*)malloc((size_t)f->nbytes + 16);
  if (fread(f->bytes, 1, (size_t)f->nbytes, fp) != (size_t)f->nbytes) { fclose(fp); return IDX2O_FILE_NOT_FOUND; }
  fclose(fp);
  i64 pos = f->nbytes;
#define RB(var) do { pos -= 4; memcpy(&(var), f->bytes + pos, 4); } while (0)
  /* ReadFileExponents, idx2Read.cpp:251-342 */
  int S, n, asz, ssz;
  RB(S);
  i64 expbeg = f->nbytes - S;
  RB(n);
  RB(asz);
  pos -= asz;
  u64* ea = (u64*)calloc((size_t)n + 1, 8);
  if (asz > 0) ZSTD_decompress(ea, (size_t)n * 8, f->bytes + pos, (size_t)asz);
  RB(ssz);
  pos -= ssz;
  br_t r = {f->bytes + pos, ssz, 0};
  f->nexp = n;
  f->eaddrs = ea;
  f->eoffs = (i64*)calloc((size_t)n + 1, 8);
  f->esizes = (i64*)calloc((size_t)n + 1, 8);
  f->edata = (u8**)calloc((size_t)n + 1, sizeof(u8*));
  f->edn = (i64*)calloc((size_t)n + 1, 8);
  i64 acc = 0;
  for (int i = 0; i < n; ++i) {
    i64 s = (i64)br_varbyte(&r);
    f->eoffs[i] = expbeg + acc;
    f->esizes[i] = s;
    acc += s;
  }
  /* ReadFile, idx2Read.cpp:81-159 */
  pos = f->nbytes - S;
  int nch, casz, cssz;
  RB(nch);
  RB(casz);
  pos -= casz;
  u64* ca = (u64*)calloc((size_t)nch + 1, 8);
  if (casz > 0) ZSTD_decompress(ca, (size_t)nch * 8, f->bytes + pos, (size_t)casz);
  RB(cssz);
  pos -= cssz;
  br_t rs = {f->bytes + pos, cssz, 0};
  f->nchunks = nch;
  f->chunks = (ochunk_t*)calloc((size_t)nch + 1, sizeof(ochunk_t));
  acc = 0;
  for (int i = 0; i < nch; ++i) {
    i64 s = (i64)br_varbyte(&rs);
    f->chunks[i].addr = ca[i];
    f->chunks[i].off = acc;
    f->chunks[i].size = s;
    acc += s;
  }
  free(ca);
#undef RB
  f->loaded = 1;
  return 0;
}

/* DecompressChunk, idx2Decode.cpp:506-540 */
static void parse_chunk(ofile_in_t* f, ochunk_t* c)
{
  if (c->parsed) return;
  c->parsed = 1;
  c->data = f->bytes + c->off;
  c->ndata = c->size;
  br_t r = {c->data, c->ndata, 0};
  c->nbricks = (int)br_varbyte(&r);
  c->bricks = (u64*)calloc((size_t)c->nbricks + 1, 8);
  c->offs = (i64*)calloc((size_t)c->nbricks + 1, 8);
  u64 b = br_varbyte(&r);
  c->bricks[0] = b;
  for (int i = 1; i < c->nbricks; ++i) { b += br_unary(&r) + 1; c->bricks[i] = b; }
  br_align(&r);
  i64 s = 0;
  c->offs[0] = 0;
  for (int i = 1; i < c->nbricks; ++i) c->offs[i] = (s += (i64)br_varbyte(&r));
  br_varbyte(&r);
  i64 hdr = (r.pos + 7) >> 3;
  for (int i = 0; i < c->nbricks; ++i) c->offs[i] += hdr;
}

static int find_chunk(dec_t* D, u64 brick, int L, int sb, int bpkey, ofile_in_t** fo, ochunk_t** co)
{
  ofile_in_t* f;
  int rc = load_file(D, brick, L, &f);
  if (rc) return rc;
  u64 a = chunk_address(D->F, brick, L, sb, bpkey);
  for (int i = 0; i < f->nchunks; ++i)
    if (f->chunks[i].addr == a) { parse_chunk(f, &f->chunks[i]); *fo = f; *co = &f->chunks[i]; return 0; }
  return IDX2O_CHUNK_NOT_FOUND;
}

static int find_exp_chunk(dec_t* D, u64 brick, int L, int sb, u8** data, i64* n)
{
  ofile_in_t* f;
  int rc = load_file(D, brick, L, &f);
  if (rc) return rc;
  u64 a = chunk_address(D->F, brick, L, sb, 0);
  for (int i = 0; i < f->nexp; ++i) {
    if (f->eaddrs[i] != a) continue;
    if (!f->edata[i]) {
      const u8* src = f->bytes + f->eoffs[i];
      unsigned long long out = ZSTD_getFrameContentSize(src, (size_t)f->esizes[i]);
      f->edata[i] = (u8*)calloc((size_t)out + 16, 1);
      ZSTD_decompress(f->edata[i], (size_t)out, src, (size_t)f->esizes[i]);
      f->edn[i] = (i64)out;
    }
    *data = f->edata[i];
    *n = f->edn[i];
    return 0;
  }
  return IDX2O_CHUNK_NOT_FOUND;
}

typedef struct {
  int bpkey; br_t r; int ok;
} ostream_t;

/* DecodeSubband, idx2Decode.cpp:85-235 */
static int decode_subband(dec_t* D, int L, int sb, u64 brick, int in_chunk, double* vol)
{
  ofile_t* F = D->F;
  const grid_t* g = &F->sb[sb];
  blocklist_t bl;
  block_list(g->d, &bl);
  int nb3[3];
  for (int d = 0; d < 3; ++d) nb3[d] = (g->d[d] + 3) / 4;
  int count = nb3[0] * nb3[1] * nb3[2];
  if (sb == 0 && L + 1 < F->nlevels) count -= (g->d[0] / 4) * (g->d[1] / 4) * (g->d[2] / 4);
  u8* ed;
  i64 en;
  int rc = find_exp_chunk(D, brick, L, sb, &ed, &en);
  if (rc) return rc;
  br_t er = {ed, en, (i64)in_chunk * count * (F->dtype ? 16 : 8)};
  ostream_t streams[128];
  int nstreams = 0;
  int etol = exponent_f64(D->tol);
  int endbp = F->dtype ? 64 : 32;
  for (int bi = 0; bi < bl.n; ++bi) {
    int* o = bl.org[bi];
    int* bd = bl.dim[bi];
    int bypass = 0;
    for (int d = 0; d < 3; ++d) if (o[d] % F->spacing[L][sb][d] != 0) bypass = 1;
    int coded_next = sb == 0 && L + 1 < F->nlevels && bd[0] == 4 && bd[1] == 4 && bd[2] == 4;
    if (coded_next) continue;
    int nd = numdims(bd);
    int nvals = 1 << (2 * nd);
    u64 blk[64];
    memset(blk, 0, sizeof blk);
    int emax = F->dtype ? (int)(short)br_get(&er, 16) - 1023 : (int)(short)br_get(&er, 8) - 127;
    int n = 0, nbps = 0;
    for (int bp = 63; bp >= 64 - endbp; --bp) {
      int realbp = bp + emax;
      int bpkey = (realbp + BP_BIAS) / F->bppc;
      int too_high = 64 - 6 > realbp - etol + 1;
      if (too_high && (realbp + BP_BIAS) % F->bppc == 0) break;
      ostream_t* s = NULL;
      for (int i = 0; i < nstreams; ++i) if (streams[i].bpkey == bpkey) { s = &streams[i]; break; }
      if (!s) {
        ofile_in_t* fo;
        ochunk_t* co;
        if (find_chunk(D, brick, L, sb, bpkey, &fo, &co)) break;
        int k = -1;
        for (int i = 0; i < co->nbricks; ++i) if (co->bricks[i] == brick) { k = i; break; }
        if (k < 0) break;
        s = &streams[nstreams++];
        s->bpkey = bpkey;
        s->r.buf = co->data;
        s->r.nbytes = co->ndata;
        s->r.pos = co->offs[k] * 8;
      }
      if (!too_high) ++nbps;
      decode_plane(blk, nvals, bp, &n, &s->r, !bypass);
    }
    if (nbps > 0 && !bypass) {
      double vals[64];
      idx2o_block_inverse(blk, nd, emax, vals);
      int j = 0;
      for (int z = 0; z < bd[2]; ++z)
        for (int y = 0; y < bd[1]; ++y)
          for (int x = 0; x < bd[0]; ++x)
            vol[vidx(F->brick_ext, g->f[0] + g->s[0] * (o[0] + x), g->f[1] + g->s[1] * (o[1] + y),
                     g->f[2] + g->s[2] * (o[2] + z))] = vals[j++];
    }
  }
  return 0;
}

typedef struct {
  dec_t* D;
  int L;
  /* decoded bricks of the previous (coarser) level and the current one */
  u64* pkeys; double** pvols; int np;
  u64* ckeys; double** cvols; int nc, cc;
  /* output */
  void* out; int odims[3]; grid_t og; int out_level;
  int rc;
} dec_ctx_t;

static double* find_vol(u64* keys, double** vols, int n, u64 b)
{
  for (int i = 0; i < n; ++i) if (keys[i] == b) return vols[i];
  return NULL;
}

/* Body of the brick traversal in Decode, idx2Decode.cpp:410-456 */
static void decode_visit(void* vctx, const int b3[3], u64 addr, int in_chunk)
{
  (void)addr;
  dec_ctx_t* c = (dec_ctx_t*)vctx;
  if (c->rc) return;
  dec_t* D = c->D;
  ofile_t* F = D->F;
  int L = c->L;
  i64 nvox = (i64)F->brick_ext[0] * F->brick_ext[1] * F->brick_ext[2];
  double* vol = (double*)calloc((size_t)nvox, sizeof(double));
  u64 brick = linear_brick(F, L, b3);
  /* DecodeBrick, idx2Decode.cpp:238-321 */
  for (int sb = 0; sb < F->nsb; ++sb) {
    if (!((F->mask[L] >> sb) & 1)) continue;
    const grid_t* S = &F->sb[sb];
    if (sb == 0 && L + 1 < F->nlevels) {
      int pb3[3], lp[3], sbne[3];
      for (int d = 0; d < 3; ++d) {
        pb3[d] = b3[d] / F->group[d];
        lp[d] = b3[d] % F->group[d];
        sbne[d] = S->d[d] - (S->d[d] > 1);
      }
      u64 pb = linear_brick(F, L + 1, pb3);
      double* pv = find_vol(c->pkeys, c->pvols, c->np, pb);
      if (!pv) { c->rc = IDX2O_CHUNK_NOT_FOUND; free(vol); return; }
      /* CopyExtentGrid<f64,f64>, Volume.h:1153-1173 */
      for (int z = 0; z < sbne[2]; ++z)
        for (int y = 0; y < sbne[1]; ++y)
          for (int x = 0; x < sbne[0]; ++x)
            vol[vidx(F->brick_ext, S->f[0] + S->s[0] * x, S->f[1] + S->s[1] * y, S->f[2] + S->s[2] * z)] =
              pv[vidx(F->brick_ext, lp[0] * sbne[0] + x, lp[1] * sbne[1] + y, lp[2] * sbne[2] + z)];
    }
    int rc = decode_subband(D, L, sb, brick, in_chunk, vol);
    if (rc) { c->rc = rc; free(vol); return; }
  }
  inverse_brick(F, L, L + 1 == F->nlevels, vol);
  if (L == 0 || F->mask[L - 1] == 0) {
    /* CopyGridGrid<f64, f32|f64> into the output grid, idx2Decode.cpp:431-445 */
    /* BrickGrid(Brick3 * B3, BrickDims3, 1 << Level on every axis) */
    grid_t bg;
    for (int d = 0; d < 3; ++d) {
      int B3 = F->brick[d];
      for (int k = 0; k < L; ++k) B3 *= F->group[d];
      bg.f[d] = b3[d] * B3; bg.d[d] = F->brick[d]; bg.s[d] = 1 << L;
    }
    /* Crop(OutGrid, BrickGrid), Volume.h:1358-1371 */
    grid_t og = c->og, cg = og;
    int ok = 1;
    for (int d = 0; d < 3; ++d) {
      int frst = imax(og.f[d], bg.f[d]);
      int last = imin(og.f[d] + (og.d[d] - 1) * og.s[d], bg.f[d] + (bg.d[d] - 1) * bg.s[d]);
      frst = ((frst - og.f[d] + og.s[d] - 1) / og.s[d]) * og.s[d] + og.f[d];
      last = ((last - og.f[d]) / og.s[d]) * og.s[d] + og.f[d];
      cg.f[d] = frst;
      cg.d[d] = frst <= last ? (last - frst) / og.s[d] + 1 : 0;
      if (!(frst <= last)) ok = 0;
    }
    if (ok) {
      for (int z = 0; z < cg.d[2]; ++z)
        for (int y = 0; y < cg.d[1]; ++y)
          for (int x = 0; x < cg.d[0]; ++x) {
            int p[3] = {cg.f[0] + x * cg.s[0], cg.f[1] + y * cg.s[1], cg.f[2] + z * cg.s[2]};
            int s[3], o[3];
            for (int d = 0; d < 3; ++d) {
              s[d] = (p[d] - bg.f[d]) / bg.s[d];
              o[d] = (p[d] - og.f[d]) / og.s[d];
            }
            double v = vol[vidx(F->brick_ext, s[0], s[1], s[2])];
            i64 oi = ((i64)o[2] * c->odims[1] + o[1]) * c->odims[0] + o[0];
            if (F->dtype) ((double*)c->out)[oi] = v;
            else ((float*)c->out)[oi] = (float)v;
          }
    }
    free(vol);
  } else {
    if (c->nc == c->cc) {
      c->cc = c->cc ? c->cc * 2 : 64;
      c->ckeys = (u64*)realloc(c->ckeys, sizeof(u64) * (size_t)c->cc);
      c->cvols = (double**)realloc(c->cvols, sizeof(double*) * (size_t)c->cc);
    }
    c->ckeys[c->nc] = brick;
    c->cvols[c->nc++] = vol;
  }
}

static int open_dataset(const char* idx2_path, const char* in_dir, const int ds[3], ofile_t* F)
{
  memset(F, 0, sizeof *F);
  int rc = read_meta(idx2_path, F);
  if (rc) return rc;
  for (int d = 0; d < 3; ++d) F->ds[d] = ds ? ds[d] : 0;
  snprintf(F->dir, sizeof F->dir, "%s", in_dir ? in_dir : ".");
  return finalize(F);
}

/* GetGrid, idx2Common.cpp:770-785 */
static grid_t output_grid(const ofile_t* F, const int first[3], const int last[3])
{
  int f[3], l[3];
  for (int d = 0; d < 3; ++d) {
    if (first && first[0] >= 0) { f[d] = first[d]; l[d] = last[d]; }
    else { f[d] = 0; l[d] = F->dims[d] - 1; }
    f[d] = imax(f[d], 0);
    l[d] = imin(l[d], F->dims[d] - 1);
  }
  grid_t g;
  for (int d = 0; d < 3; ++d) {
    int s = 1 << F->ds[d];
    int L3 = ((l[d] + s - 1) / s) * s;
    int F3 = (f[d] / s) * s;
    g.f[d] = F3; g.d[d] = (L3 - F3) / s + 1; g.s[d] = s;
  }
  return g;
}

int idx2o_output_dims(const char* idx2_path, const int first[3], const int last[3],
                      const int downsampling[3], int out_dims[3], int* dtype)
{
  ofile_t* F = (ofile_t*)calloc(1, sizeof(ofile_t));
  int rc = open_dataset(idx2_path, ".", downsampling, F);
  if (!rc) {
    grid_t g = output_grid(F, first, last);
    for (int d = 0; d < 3; ++d) out_dims[d] = g.d[d];
    if (dtype) *dtype = F->dtype;
  }
  free(F);
  return rc;
}

int idx2o_decode(const char* idx2_path, const char* in_dir, const int first[3], const int last[3],
                 const int downsampling[3], double tolerance, void* out, int64_t out_bytes)
{
  ofile_t* F = (ofile_t*)calloc(1, sizeof(ofile_t));
  int rc = open_dataset(idx2_path, in_dir, downsampling, F);
  if (rc) { free(F); return rc; }
  grid_t og = output_grid(F, first, last);
  i64 need = (i64)og.d[0] * og.d[1] * og.d[2] * (F->dtype ? 8 : 4);
  if (need > out_bytes) { free(F); return IDX2O_SIZE_MISMATCHED; }
  memset(out, 0, (size_t)need);
  dec_t D;
  memset(&D, 0, sizeof D);
  D.F = F;
  D.tol = F->tolerance > tolerance ? F->tolerance : tolerance;
  /* decode extent in samples: the (cropped) query box */
  int ef[3], ed[3];
  for (int d = 0; d < 3; ++d) {
    if (first && first[0] >= 0) { ef[d] = first[d]; ed[d] = last[d] - first[d] + 1; }
    else { ef[d] = 0; ed[d] = F->dims[d]; }
  }
  dec_ctx_t c;
  memset(&c, 0, sizeof c);
  c.D = &D;
  c.out = out;
  for (int d = 0; d < 3; ++d) c.odims[d] = og.d[d];
  c.og = og;
  /* level loop, idx2Decode.cpp:368-471 */
  for (int L = F->nlevels - 1; L >= 0 && !c.rc; --L) {
    if (F->mask[L] == 0) break;
    c.L = L;
    int B3[3], C3[3], F3[3];
    for (int d = 0; d < 3; ++d) {
      int g = 1;
      for (int k = 0; k < L; ++k) g *= F->group[d];
      B3[d] = F->brick[d] * g;
      C3[d] = F->bpc3[L][d] * B3[d];
      F3[d] = C3[d] * F->cpf3[L][d];
    }
    int eb_f[3], eb_d[3], ec_f[3], ec_d[3], efl_f[3], efl_d[3];
    int vb_f[3], vb_d[3], vc_f[3], vc_d[3], vf_f[3], vf_d[3];
    for (int d = 0; d < 3; ++d) {
      int lo = ef[d], hi = ef[d] + ed[d] - 1;
      eb_f[d] = lo / B3[d]; eb_d[d] = hi / B3[d] - eb_f[d] + 1;
      ec_f[d] = lo / C3[d]; ec_d[d] = hi / C3[d] - ec_f[d] + 1;
      efl_f[d] = lo / F3[d]; efl_d[d] = hi / F3[d] - efl_f[d] + 1;
      int vhi = F->dims[d] - 1;
      vb_f[d] = 0; vb_d[d] = vhi / B3[d] + 1;
      vc_f[d] = 0; vc_d[d] = vhi / C3[d] + 1;
      vf_f[d] = 0; vf_d[d] = vhi / F3[d] + 1;
    }
    /* file -> chunk -> brick traversal (idx2Common.h:317-461) */
    typedef struct { u64 ord, prev; int f[3], t[3]; } fitem;
    fitem fst[128];
    int ftop = 0;
    fst[0].ord = fst[0].prev = F->forder[L];
    for (int d = 0; d < 3; ++d) { fst[0].f[d] = 0; fst[0].t[d] = (int)nextpow2(F->nfiles3[L][d]); }
    while (ftop >= 0 && !c.rc) {
      fitem* T = &fst[ftop];
      int FD = (int)(T->ord & 3);
      T->ord >>= 2;
      if (FD == 3) { if (T->ord == 3) T->ord = T->prev; else T->prev = T->ord; continue; }
      --ftop;
      if (T->t[0] - T->f[0] == 1 && T->t[1] - T->f[1] == 1 && T->t[2] - T->f[2] == 1) {
        fitem FT = *T;
        /* chunk traversal */
        fitem cst[128];
        int ctop = 0;
        cst[0].ord = cst[0].prev = F->corder_in_file[L];
        for (int d = 0; d < 3; ++d) {
          cst[0].f[d] = FT.f[d] * F->cpf3[L][d];
          cst[0].t[d] = cst[0].f[d] + (int)nextpow2(F->cpf3[L][d]);
        }
        while (ctop >= 0 && !c.rc) {
          fitem* CT = &cst[ctop];
          int CD = (int)(CT->ord & 3);
          CT->ord >>= 2;
          if (CD == 3) { if (CT->ord == 3) CT->ord = CT->prev; else CT->prev = CT->ord; continue; }
          --ctop;
          if (CT->t[0] - CT->f[0] == 1 && CT->t[1] - CT->f[1] == 1 && CT->t[2] - CT->f[2] == 1) {
            int bf[3];
            for (int d = 0; d < 3; ++d) bf[d] = CT->f[d] * F->bpc3[L][d];
            brick_traverse(F->border_in_chunk[L], bf, F->bpc3[L], eb_f, eb_d, vb_f, vb_d, decode_visit, &c);
            continue;
          }
          fitem A = *CT, B = *CT;
          A.t[CD] = CT->f[CD] + (CT->t[CD] - CT->f[CD]) / 2;
          B.f[CD] = A.t[CD];
          int aok = 1, bok = 1;
          for (int d = 0; d < 3; ++d) {
            if (!(B.f[d] < ec_f[d] + ec_d[d] && ec_f[d] < B.t[d])) bok = 0;
            if (!(A.f[d] < ec_f[d] + ec_d[d] && ec_f[d] < A.t[d])) aok = 0;
          }
          if (bok) cst[++ctop] = B;
          if (aok) cst[++ctop] = A;
        }
        (void)vc_f; (void)vc_d; (void)vf_f; (void)vf_d;
        continue;
      }
      fitem A = *T, B = *T;
      A.t[FD] = T->f[FD] + (T->t[FD] - T->f[FD]) / 2;
      B.f[FD] = A.t[FD];
      int aok = 1, bok = 1;
      for (int d = 0; d < 3; ++d) {
        if (!(B.f[d] < efl_f[d] + efl_d[d] && efl_f[d] < B.t[d])) bok = 0;
        if (!(A.f[d] < efl_f[d] + efl_d[d] && efl_f[d] < A.t[d])) aok = 0;
      }
      if (bok) fst[++ftop] = B;
      if (aok) fst[++ftop] = A;
    }
    /* parents of this level are no longer needed */
    for (int i = 0; i < c.np; ++i) free(c.pvols[i]);
    free(c.pkeys); free(c.pvols);
    c.pkeys = c.ckeys; c.pvols = c.cvols; c.np = c.nc;
    c.ckeys = NULL; c.cvols = NULL; c.nc = c.cc = 0;
  }
  for (int i = 0; i < c.np; ++i) free(c.pvols[i]);
  free(c.pkeys); free(c.pvols);
  for (int i = 0; i < D.nfiles; ++i) {
    ofile_in_t* f = &D.files[i];
    for (int k = 0; k < f->nchunks; ++k) { free(f->chunks[k].bricks); free(f->chunks[k].offs); }
    free(f->chunks);
    for (int k = 0; k < f->nexp; ++k) free(f->edata[k]);
    free(f->edata); free(f->eaddrs); free(f->eoffs); free(f->esizes); free(f->edn);
    free(f->bytes);
  }
  free(D.files);
  rc = c.rc;
  free(F);
  return rc;
}
