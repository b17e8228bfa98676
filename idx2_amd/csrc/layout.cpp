// Dataset layout derivation (restates Finalize, Source/Core/idx2Common.cpp:358-754).
#include "layout.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/idx2g.h"

namespace idx2g {

int msb64(uint64_t v) { return v ? 63 - __builtin_clzll(v) : -1; }
int log2ceil(int64_t v) {
  int i = msb64((uint64_t)v);
  return i + (v != (int64_t(1) << i));
}
int64_t next_pow2(int64_t v) { return v == 0 ? 1 : (int64_t(1) << (msb64((uint64_t)(v - 1)) + 1)); }
static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Base-4 encoding of a transform-order string, first character in the low
// digits (EncodeTransformOrder, Source/Core/Wavelet.cpp:406-419).
uint64_t encode_order(const std::string& s) {
  uint64_t r = 0;
  for (int i = (int)s.size() - 1; i >= 0; --i) r = (r << 2) + (s[i] == '+' ? 3 : s[i] - 'X');
  return r;
}

// Expand an order over a power-of-two padded extent: one axis letter per
// halving until the extent is a single brick (Wavelet.cpp:435-460).
static std::string expand_order(uint64_t in, const int n3[3]) {
  int64_t n[3] = {next_pow2(n3[0]), next_pow2(n3[1]), next_pow2(n3[2])};
  std::string out;
  uint64_t saved = in;
  while (n[0] * n[1] * n[2] > 1) {
    int t = int(in & 3);
    in >>= 2;
    if (t == 3) {
      if (in == 3) in = saved; else saved = in;
    } else {
      out.push_back(char('X' + t));
      n[t] >>= 1;
    }
  }
  return out;
}

// Number of axis letters in one pass of the transform (Wavelet.cpp:463-490).
static int one_pass_len(uint64_t in) {
  int len = 0, p3[3] = {1, 1, 1};
  uint64_t saved = in;
  for (;;) {
    int t = int(in & 3);
    in >>= 2;
    if (t == 3) {
      if (in == 3) in = saved; else saved = in;
    } else {
      if (--p3[t] < 0) break;
      ++len;
    }
  }
  return len;
}

static void morton_blocks(const int sbdims[3], bool skip_full, std::vector<BlockInfo>* out) {
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = (sbdims[d] + 3) / 4;
  struct K { uint32_t code; int x, y, z; };
  std::vector<K> keys;
  auto spread = [](uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
  };
  for (int z = 0; z < nb[2]; ++z)
    for (int y = 0; y < nb[1]; ++y)
      for (int x = 0; x < nb[0]; ++x) keys.push_back({spread(x) | (spread(y) << 1) | (spread(z) << 2), x, y, z});
  std::sort(keys.begin(), keys.end(), [](const K& a, const K& b) { return a.code < b.code; });
  out->clear();
  for (auto& k : keys) {
    BlockInfo b;
    int o[3] = {k.x * 4, k.y * 4, k.z * 4};
    int nd = 0;
    for (int d = 0; d < 3; ++d) {
      b.org[d] = o[d];
      b.dims[d] = std::min(4, sbdims[d] - o[d]);
      nd += b.dims[d] > 1;
    }
    b.ndims = nd;
    // full blocks of subband 0 below the coarsest level are coded in the next
    // level (CodedInNextLevel): they carry no exponent and no bits, so they
    // are left out of the table and no kernel spends a step on them (the
    // Morton rank of the blocks that remain keeps their order)
    b.skip = skip_full && b.dims[0] == 4 && b.dims[1] == 4 && b.dims[2] == 4;
    if (!b.skip) out->push_back(b);
  }
}

int Layout::finalize() {
  for (int d = 0; d < 3; ++d)
    if (!is_pow2(brick[d])) return IDX2G_BRICK_SIZE_NOT_POW2;
  for (int d = 0; d < 3; ++d)
    if (dims[d] < brick[d]) return IDX2G_BRICK_SIZE_TOO_BIG;
  if (nlevels == 0) {  // GuessNumLevelsIfNeeded, idx2Common.cpp:666-680
    int b[3] = {brick[0], brick[1], brick[2]};
    auto le = [&](const int* v) { return v[0] <= dims[0] && v[1] <= dims[1] && v[2] <= dims[2]; };
    while (le(b)) {
      for (int d = 0; d < 3; ++d) b[d] = b[d] == 1 ? 1 : b[d] * 2;
      if (le(b)) ++nlevels;
    }
  }
  if (nlevels > kMaxLevels || nlevels < 1) return IDX2G_TOO_MANY_LEVELS;
  // transform order: transformed axes then "++" (idx2Common.cpp:373-387)
  std::string tord;
  for (int d = 0; d < 3; ++d)
    if (brick[d] > 1) tord.push_back(char('X' + d));
  tord += "++";
  tform_order = encode_order(tord);
  tform_full_len = one_pass_len(tform_order);
  // subbands of the extended brick: even (low) / odd (high) samples per
  // transformed axis, index bits X-major (Wavelet.cpp:502-567 for one level)
  int axes[3], na = 0;
  for (int d = 0; d < 3; ++d) {
    ext[d] = brick[d] + (brick[d] > 1);
    if (brick[d] > 1) axes[na++] = d;
  }
  nsb = 1 << na;
  for (int s = 0; s < nsb; ++s) {
    for (int d = 0; d < 3; ++d) {
      sb_lh[s][d] = 0;
      sb[s].from[d] = 0; sb[s].dims[d] = ext[d]; sb[s].strd[d] = 1;
    }
    for (int k = 0; k < na; ++k) {
      int d = axes[k], h = (s >> (na - 1 - k)) & 1;
      sb_lh[s][d] = h;
      sb[s].from[d] = h;
      sb[s].strd[d] = 2;
      sb[s].dims[d] = h ? ext[d] / 2 : (ext[d] + 1) / 2;
    }
  }
  for (int d = 0; d < 3; ++d) group[d] = brick[d] > 1 ? 2 : 1;
  // bricks per level (idx2Common.cpp:463-474)
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = (dims[d] + brick[d] - 1) / brick[d];
  for (int l = 0; l < nlevels; ++l)
    for (int d = 0; d < 3; ++d) {
      nbricks[l][d] = nb[d];
      nb[d] = (nb[d] + group[d] - 1) / group[d];
    }
  // brick orders: leftover bits of the larger axes first, then the transform
  // order repeated (idx2Common.cpp:478-521)
  for (int l = 0; l < nlevels; ++l) {
    int logn[3] = {log2ceil(nbricks[l][0]), log2ceil(nbricks[l][1]), log2ceil(nbricks[l][2])};
    int mn = std::min(logn[0], std::min(logn[1], logn[2]));
    int left[3];
    for (int d = 0; d < 3; ++d) left[d] = logn[d] - (brick[d] > 1 ? mn : 0);
    std::string bo;
    for (int d = 0; d < 3; ++d)
      if (brick[d] == 1)
        while (left[d]-- > 0) bo.push_back(char('X' + d));
    while (!(left[0] <= 0 && left[1] <= 0 && left[2] <= 0))
      for (int d = 0; d < 3; ++d)
        if (left[d]-- > 0) bo.push_back(char('X' + d));
    if (!bo.empty()) bo.push_back('+');
    bo += tord;
    border[l] = encode_order(bo);
    bstr[l] = expand_order(border[l], nbricks[l]);
    if ((int)bstr[l].size() < tform_full_len) return IDX2G_TOO_MANY_LEVELS;
  }
  // bricks per chunk / chunks per file / file orders (idx2Common.cpp:524-623)
  if (bpc_in > 32768) return IDX2G_TOO_MANY_BRICKS_PER_CHUNK;
  if (!is_pow2(bpc_in)) return IDX2G_BRICKS_PER_CHUNK_NOT_POW2;
  if (cpf_in > 4906) return IDX2G_TOO_MANY_CHUNKS_PER_FILE;
  if (!is_pow2(cpf_in)) return IDX2G_CHUNKS_PER_FILE_NOT_POW2;
  bpc[0] = bpc_in;
  cpf[0] = cpf_in;
  for (int l = 0; l < nlevels; ++l) {
    int len = (int)bstr[l].size();
    bpc[l] = 1 << std::min(log2ceil(bpc[0]), len);
    int bic = log2ceil(bpc[l]);
    std::string s(bic, ' ');
    for (int d = 0; d < 3; ++d) bpc3[l][d] = 1;
    for (int j = 0; j < bic; ++j) {
      char c = bstr[l][len - j - 1];
      bpc3[l][c - 'X'] *= 2;
      s[bic - j - 1] = c;
    }
    border_in_chunk[l] = encode_order(s);
    int nchunks3[3];
    for (int d = 0; d < 3; ++d) nchunks3[d] = (nbricks[l][d] + bpc3[l][d] - 1) / bpc3[l][d];
    cpf[l] = 1 << std::min(log2ceil(cpf[0]), len - bic);
    int cif = log2ceil(cpf[l]);
    s.assign(cif, ' ');
    for (int d = 0; d < 3; ++d) cpf3[l][d] = 1;
    for (int j = 0; j < cif; ++j) {
      char c = bstr[l][len - bic - j - 1];
      cpf3[l][c - 'X'] *= 2;
      s[cif - j - 1] = c;
    }
    corder_in_file[l] = encode_order(s);
    for (int d = 0; d < 3; ++d) nfiles3[l][d] = (nchunks3[d] + cpf3[l][d] - 1) / cpf3[l][d];
    int fl = log2ceil(int64_t(1) << (len - bic - cif));
    s.assign(fl, ' ');
    for (int j = 0; j < fl; ++j) s[fl - j - 1] = bstr[l][len - bic - cif - j - 1];
    forder[l] = encode_order(s);
  }
  // directory depths (idx2Common.cpp:626-651)
  if (files_per_dir > 4096) return IDX2G_TOO_MANY_FILES_PER_DIR;
  for (int l = 0; l < nlevels; ++l) {
    bpf[l] = int64_t(bpc[l]) * cpf[l];
    fdd[l].clear();
    int acc = log2ceil(bpf[l]);
    fdd[l].push_back(acc);
    int len = (int)bstr[l].size();
    while (acc < len) {
      int inc = std::min(len - acc, log2ceil(files_per_dir));
      fdd[l].push_back(inc);
      acc += inc;
    }
    if (fdd[l].size() > 4) return IDX2G_TOO_MANY_FILES_PER_DIR;
    std::reverse(fdd[l].begin(), fdd[l].end());
  }
  // limits of this implementation, checked after every check of the
  // reference so that parameters the reference rejects get its code:
  // partial 4^d blocks must have extent 1 (brick >= 8), one bit plane per chunk
  for (int d = 0; d < 3; ++d)
    if (brick[d] > 1 && brick[d] < 8) return IDX2G_NOT_SUPPORTED;
  if (bppc != 1) return IDX2G_NOT_SUPPORTED;
  // blocks of every subband, Morton order (idx2Encode.cpp:333-350)
  for (int l = 0; l < nlevels; ++l)
    for (int s = 0; s < nsb; ++s) {
      morton_blocks(sb[s].dims, s == 0 && l + 1 < nlevels, &blocks[l][s]);
      int n = 0;
      for (auto& b : blocks[l][s]) n += !b.skip;
      ncoded[l][s] = n;
    }
  // CDF5/3 basis norms (GetCdf53NormsFast<16>, Wavelet.h:19-32) and the
  // per-subband weights of ForwardCdf53/InverseCdf53 (Wavelet.cpp:111-177)
  double scal[16], wave[16];
  {
    double n1 = 3, n2 = 23;
    for (int i = 0; i < 16; ++i) {
      scal[i] = std::sqrt(n1 / (1 << (i + 1)));
      n1 = n1 * 4 - 1;
      wave[i] = std::sqrt(n2 / (1 << (i + 5)));
      n2 = n2 * 4 - 33;
    }
  }
  for (int l = 0; l < nlevels; ++l)
    for (int s = 0; s < nsb; ++s) {
      double w[3];
      for (int d = 0; d < 3; ++d) {
        if (ext[d] == 1) w[d] = 1;
        else w[d] = sb_lh[s][d] == 0 ? scal[l] : wave[l];  // Level3Rev: 1 low, 0 high
      }
      wfwd[l][s] = w[0] * w[1] * w[2];
      winv[l][s] = 1.0 / (w[0] * w[1] * w[2]);
    }
  return IDX2G_OK;
}

// GetLinearBrick (Source/Core/idx2Lookup.cpp:12-24): bit interleave, last
// order letter gives the least significant bit.
uint64_t Layout::linear_brick(int l, const int b3in[3]) const {
  uint64_t r = 0;
  int b3[3] = {b3in[0], b3in[1], b3in[2]};
  int n = (int)bstr[l].size();
  for (int i = n - 1; i >= 0; --i) {
    int d = bstr[l][i] - 'X';
    r |= uint64_t(b3[d] & 1) << (n - i - 1);
    b3[d] >>= 1;
  }
  return r;
}

void Layout::spatial_brick(int l, uint64_t lin, int b3[3]) const {
  b3[0] = b3[1] = b3[2] = 0;
  int n = (int)bstr[l].size();
  for (int i = 0; i < n; ++i) {
    int d = bstr[l][i] - 'X';
    b3[d] = (b3[d] << 1) | int((lin >> (n - i - 1)) & 1);
  }
}

int Layout::log2_bpc(int l) const { return log2ceil(bpc[l]); }

// GetAddress (Source/Core/idx2Lookup.h:15-22)
static uint64_t address(uint64_t brick, int shift, int level, int sb, int bp) {
  return (uint64_t(level) << 60) + ((brick >> shift) << 18) + uint64_t(sb << 12) + (uint64_t(bp) & 0xFFF);
}
uint64_t Layout::chunk_address(uint64_t brick, int l, int s, int bpkey) const {
  return address(brick, log2ceil(bpc[l]), l, s, bpkey);
}
uint64_t Layout::file_id(uint64_t brick, int l) const { return address(brick, log2ceil(bpf[l]), l, 0, 0); }

// ConstructFilePath (Source/Core/idx2Lookup.cpp:77-107)
std::string Layout::file_path(const std::string& dir, uint64_t brick, int l) const {
  char buf[64];
  std::string p = dir + "/" + name + "/" + field + "/";
  snprintf(buf, sizeof buf, "/L%02x", l);
  p += buf;
  for (size_t depth = 0; depth + 1 < fdd[l].size(); ++depth) {
    int bitlen = 64 - (int)bstr[l].size() + fdd[l][depth];
    int shift = 64 - bitlen;
    uint64_t first = shift >= 64 ? 0 : (brick & (~0ull << shift)) >> shift;
    snprintf(buf, sizeof buf, "/B%llx", (unsigned long long)first);
    p += buf;
    brick <<= fdd[l][depth];
  }
  return p + ".bin";
}

void Layout::level_bricks(int l, std::vector<uint64_t>* lin, std::vector<int>* xyz) const {
  // walk the linear ids in order and keep those inside the level's brick
  // grid: the result comes out sorted, no per-brick search or sort
  const int n = nbricks[l][0] * nbricks[l][1] * nbricks[l][2];
  lin->resize(n);
  xyz->resize(3 * (size_t)n);
  const uint64_t span = linear_span(l);
  const int nb = (int)bstr[l].size();
  // bit i of the brick string (most significant first) is bit `rank[i]` of
  // coordinate axis[i], rank[i] = the later bits of the same axis; a table
  // per byte of the id holds that byte's share of (x, y, z)
  int axis[64], rank[64];
  for (int i = 0; i < nb; ++i) axis[i] = bstr[l][i] - 'X';
  for (int i = 0; i < nb; ++i) {
    rank[i] = 0;
    for (int j = i + 1; j < nb; ++j) rank[i] += axis[j] == axis[i];
  }
  const int nbytes = (nb + 7) / 8;
  std::vector<int> tab((size_t)nbytes * 256 * 3, 0);
  for (int by = 0; by < nbytes; ++by)
    for (int v = 0; v < 256; ++v)
      for (int bit = 0; bit < 8; ++bit) {
        const int pos = by * 8 + bit, i = nb - 1 - pos;  // id bit pos = string bit i
        if (pos < nb && ((v >> bit) & 1)) tab[((size_t)by * 256 + v) * 3 + axis[i]] |= 1 << rank[i];
      }
  const int* T = tab.data();
  const int n0 = nbricks[l][0], n1 = nbricks[l][1], n2 = nbricks[l][2];
  uint64_t* L = lin->data();
  int* X = xyz->data();
  int k = 0;
  for (uint64_t id = 0; id < span && k < n; ++id) {
    int x = 0, y = 0, z = 0;
    for (int by = 0; by < nbytes; ++by) {
      const int* t = T + ((size_t)by * 256 + ((id >> (8 * by)) & 255)) * 3;
      x |= t[0];
      y |= t[1];
      z |= t[2];
    }
    if (x >= n0 || y >= n1 || z >= n2) continue;
    L[k] = id;
    X[3 * k] = x;
    X[3 * k + 1] = y;
    X[3 * k + 2] = z;
    ++k;
  }
}

// WriteMetaFile (Source/Core/idx2Common.cpp:149-179)
std::string Layout::meta_text() const {
  std::string s;
  char b[512];
  auto add = [&](const char* fmt, auto... a) {
    snprintf(b, sizeof b, fmt, a...);
    s += b;
  };
  std::string tord;
  for (uint64_t t = tform_order; t; t >>= 2) tord.push_back((t & 3) == 3 ? '+' : char('X' + (t & 3)));
  s += "(\n  (common\n    (type \"Simulation\")\n";
  add("    (name \"%s\")\n", name.c_str());
  add("    (field \"%s\")\n", field.c_str());
  add("    (dimensions %d %d %d)\n", dims[0], dims[1], dims[2]);
  add("    (data-type \"%s\")\n", dtype ? "float64" : "float32");
  add("    (min-max %.20f %.20f)\n", vmin, vmax);
  add("    (accuracy %.20f)\n", tolerance);
  s += "  )\n  (format\n";
  add("    (version %d %d)\n", version[0], version[1]);
  add("    (brick-size %d %d %d)\n", brick[0], brick[1], brick[2]);
  add("    (transform-order \"%s\")\n", tord.c_str());
  add("    (num-levels %d)\n", nlevels);
  add("    (transform-passes-per-levels %d)\n", 1);
  add("    (bricks-per-chunk %d)\n", bpc_in);
  add("    (chunks-per-file %d)\n", cpf_in);
  add("    (files-per-directory %d)\n", files_per_dir);
  add("    (bit-planes-per-chunk %d)\n", bppc);
  s += "  )\n)\n";
  return s;
}

void compute_decode_masks(const Layout& L, const int ds[3], DecodeMasks* m) {
  int sp[3];
  for (int d = 0; d < 3; ++d) sp[d] = 1 << ds[d];
  int prev0[3] = {1, 1, 1};
  for (int l = 0; l < L.nlevels; ++l) {
    uint8_t mask = 0xFF;  // bits of absent subbands (2D) stay set, as in the reference
    int cur0[3] = {1, 1, 1};
    for (int s = 0; s < L.nsb; ++s) {
      int f3[3], s3[3];
      for (int d = 0; d < 3; ++d) {
        f3[d] = L.sb[s].from[d] * (l > 0 ? prev0[d] : 1);
        s3[d] = L.sb[s].strd[d] * (l > 0 ? prev0[d] : 1);
        m->spacing[l][s][d] = 1;
      }
      if (s == 0)
        for (int d = 0; d < 3; ++d) cur0[d] = s3[d];
      for (int d = 0; d < 3; ++d) {
        if (f3[d] % sp[d] == 0) {
          // sp and s3 are powers of two, so the spacing is one as well (the
          // decode kernels test block origins against spacing - 1)
          if (s3[d] % sp[d] != 0) m->spacing[l][s][d] = sp[d] / s3[d];
        } else {
          mask = uint8_t(mask & ~(1u << s));
        }
      }
    }
    if (l + 1 < L.nlevels && mask == 1) mask = 0;
    m->mask[l] = mask;
    for (int d = 0; d < 3; ++d) prev0[d] = cur0[d];
  }
}

// Reads the keyed values of the first list of a .idx2 text; the statistics
// appended after it are ignored (Source/Core/sexpr.h:313-354).
int parse_meta(const std::string& text, Layout* L) {
  size_t stop = text.find("Statistics");
  auto find = [&](const char* key) -> const char* {
    std::string pat = std::string("(") + key + " ";
    size_t p = text.find(pat);
    if (p == std::string::npos || (stop != std::string::npos && p > stop)) return nullptr;
    return text.c_str() + p + pat.size();
  };
  const char* p;
  char buf[128];
  if (!(p = find("name")) || sscanf(p, "\"%63[^\"]\"", buf) != 1) return IDX2G_SYNTAX_ERROR;
  L->name = buf;
  if (!(p = find("field")) || sscanf(p, "\"%63[^\"]\"", buf) != 1) return IDX2G_SYNTAX_ERROR;
  L->field = buf;
  if (!(p = find("dimensions")) || sscanf(p, "%d %d %d", &L->dims[0], &L->dims[1], &L->dims[2]) != 3)
    return IDX2G_SYNTAX_ERROR;
  if (!(p = find("data-type")) || sscanf(p, "\"%31[^\"]\"", buf) != 1) return IDX2G_SYNTAX_ERROR;
  if (!strcmp(buf, "float64")) L->dtype = 1;
  else if (!strcmp(buf, "float32")) L->dtype = 0;
  else return IDX2G_TYPE_NOT_SUPPORTED;
  if (!(p = find("accuracy")) || sscanf(p, "%lf", &L->tolerance) != 1) return IDX2G_SYNTAX_ERROR;
  if ((p = find("min-max"))) sscanf(p, "%lf %lf", &L->vmin, &L->vmax);
  if (!(p = find("brick-size")) || sscanf(p, "%d %d %d", &L->brick[0], &L->brick[1], &L->brick[2]) != 3)
    return IDX2G_SYNTAX_ERROR;
  if (!(p = find("num-levels")) || sscanf(p, "%d", &L->nlevels) != 1) return IDX2G_SYNTAX_ERROR;
  if (!(p = find("bricks-per-chunk")) || sscanf(p, "%d", &L->bpc_in) != 1) return IDX2G_SYNTAX_ERROR;
  if (!(p = find("chunks-per-file")) || sscanf(p, "%d", &L->cpf_in) != 1) return IDX2G_SYNTAX_ERROR;
  if (!(p = find("files-per-directory")) || sscanf(p, "%d", &L->files_per_dir) != 1)
    return IDX2G_SYNTAX_ERROR;
  if (!(p = find("bit-planes-per-chunk")) || sscanf(p, "%d", &L->bppc) != 1) return IDX2G_SYNTAX_ERROR;
  if ((p = find("version"))) sscanf(p, "%d %d", &L->version[0], &L->version[1]);
  return IDX2G_OK;
}

}  // namespace idx2g
