// Dataset layout of an idx2 dataset: everything the reference derives in
// Finalize (Source/Core/idx2Common.cpp:725-754) from the user parameters --
// transform order, subbands and norms, brick/chunk/file bit orders per level,
// file paths, chunk addresses, decode masks -- plus the .idx2 metadata text.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace idx2g {

constexpr int kMaxLevels = 16;
constexpr int kBpBias = 1024;  // BitPlaneKeyBias_, Source/Core/idx2Lookup.h:11

struct Grid3 {
  int from[3], dims[3], strd[3];
};

struct BlockInfo {   // one 4^d block of a subband, in Morton order
  int org[3];        // origin in subband samples
  int dims[3];       // 4 or fewer per axis
  int ndims;         // number of axes with dims > 1
  int skip;          // coded in the next level (full subband-0 block below the coarsest level)
};

struct Layout {
  // user parameters
  std::string name, field;
  int dims[3] = {0, 0, 0};
  int dtype = 0;  // 0 f32, 1 f64
  int brick[3] = {32, 32, 32};
  int nlevels = 0;
  double tolerance = 0;
  int bpc_in = 4096, cpf_in = 64, files_per_dir = 64, bppc = 1;
  int version[2] = {1, 0};
  // derived
  int ext[3];                 // extended brick dims (BrickDimsExt3)
  uint64_t tform_order = 0;   // "XYZ++" / "XY++" encoded base 4
  int tform_full_len = 0;
  int nsb = 0;
  Grid3 sb[8];
  int sb_lh[8][3];
  int group[3];
  int nbricks[kMaxLevels][3];
  std::string bstr[kMaxLevels];
  uint64_t border[kMaxLevels];
  int bpc[kMaxLevels], cpf[kMaxLevels];
  int64_t bpf[kMaxLevels];
  int bpc3[kMaxLevels][3], cpf3[kMaxLevels][3], nfiles3[kMaxLevels][3];
  uint64_t border_in_chunk[kMaxLevels], corder_in_file[kMaxLevels], forder[kMaxLevels];
  std::vector<int> fdd[kMaxLevels];
  std::vector<BlockInfo> blocks[kMaxLevels][8];  // per level (skip flag differs) and subband
  int ncoded[kMaxLevels][8];                     // coded blocks per brick-subband
  // forward subband weights (Wx*Wy)*Wz and inverse 1/((Wx*Wy)*Wz) per level and subband
  double wfwd[kMaxLevels][8], winv[kMaxLevels][8];
  double vmin = 0, vmax = 0;

  int finalize();  // returns an idx2g error code
  uint64_t linear_brick(int level, const int b3[3]) const;
  void spatial_brick(int level, uint64_t lin, int b3[3]) const;
  uint64_t chunk_address(uint64_t brick, int level, int sb, int bpkey) const;
  uint64_t file_id(uint64_t brick, int level) const;
  std::string file_path(const std::string& dir, uint64_t brick, int level) const;
  int log2_bpc(int level) const;
  std::string meta_text() const;  // WriteMetaFile text
  // existing bricks of a level in linear order: linear ids and coordinates
  void level_bricks(int level, std::vector<uint64_t>* lin, std::vector<int>* xyz) const;
  uint64_t linear_span(int level) const { return 1ull << bstr[level].size(); }
};

// decode masks and block spacings for a downsampling factor
// (Source/Core/idx2Common.cpp:402-460)
struct DecodeMasks {
  uint8_t mask[kMaxLevels];
  int spacing[kMaxLevels][8][3];
};
void compute_decode_masks(const Layout& L, const int ds[3], DecodeMasks* m);

// parse the first list of a .idx2 file (idx2Common.cpp:182-346)
int parse_meta(const std::string& text, Layout* L);

int msb64(uint64_t v);
int log2ceil(int64_t v);
int64_t next_pow2(int64_t v);
uint64_t encode_order(const std::string& s);

}  // namespace idx2g
