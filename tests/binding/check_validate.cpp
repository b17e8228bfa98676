// Checks include/idx2g_idx2.hpp against the reference library it binds to:
// compiled against the reference's own headers (the static_asserts there
// compare every error code with the reference's enum) and linked with the
// reference's Core objects and libidx2g.  For random encode parameter sets it
// compares the error the reference's Finalize returns (after idx2App's
// SetParams, Applications/idx2App.cpp:216-233) with idx2g_idx2::Validate.
// Each reference call runs in a child process: the reference asserts on some
// inputs (a FilesDirsDepth index past its 4 entries, idx2Common.cpp:636-644);
// such cases are counted and skipped.  Bricks larger than 128 per axis are
// accepted by the reference and refused by libidx2g (NotSupportedInVersion);
// they are counted apart (the random sets here use bricks of at most 128).
#include "idx2.h"
#include "idx2g_idx2.hpp"
#include <stdio.h>
#include <stdlib.h>
#include <sys/wait.h>
#include <unistd.h>
using namespace idx2;
static error<idx2_err_code> RefFinalize(params* P) {
  idx2_file Idx2;
  SetName(&Idx2, P->Meta.Name); SetField(&Idx2, P->Meta.Field); SetVersion(&Idx2, P->Version);
  SetDimensions(&Idx2, P->Meta.Dims3); SetDataType(&Idx2, P->Meta.DType); SetBrickSize(&Idx2, P->BrickDims3);
  SetBricksPerChunk(&Idx2, P->BricksPerChunk); SetChunksPerFile(&Idx2, P->ChunksPerFile);
  SetBitPlanesPerChunk(&Idx2, P->BitPlanesPerChunk); SetNumLevels(&Idx2, (i8)P->NLevels);
  SetTolerance(&Idx2, P->Tolerance); SetFilesPerDirectory(&Idx2, P->FilesPerDir); SetDir(&Idx2, P->OutDir);
  auto e = Finalize(&Idx2, P);
  Dealloc(&Idx2);
  return e;
}
int main(int argc, char** argv) {
  int n = argc > 1 ? atoi(argv[1]) : 2000, bad = 0, aborted = 0, limit = 0;
  srand(7);
  const int dimv[] = {1, 17, 32, 48, 64, 96, 100, 128, 256, 384, 1000};
  const int brv[] = {8, 16, 24, 32, 32, 32, 64, 128};
  const int bpcv[] = {1, 3, 64, 512, 4096, 8192, 65536};
  const int cpfv[] = {1, 3, 64, 4096, 8192};
  const int fpdv[] = {1, 3, 64, 512, 5000};
  for (int i = 0; i < n; ++i) {
    params P;
    snprintf(P.Meta.Name, 64, "N"); snprintf(P.Meta.Field, 64, "f");
    for (int d = 0; d < 3; ++d) P.Meta.Dims3[d] = dimv[rand() % 11];
    if (rand() % 3 == 0) P.Meta.Dims3[2] = 1;
    P.Meta.DType = rand() % 2 ? dtype::float64 : dtype::float32;
    { int b = rand() % 2 ? 32 : brv[rand() % 8]; for (int d = 0; d < 3; ++d) P.BrickDims3[d] = b; }
    if (P.Meta.Dims3[2] == 1 && rand() % 2) P.BrickDims3[2] = 1;
    if (P.Meta.Dims3[2] != 1 && P.BrickDims3[2] == 1) P.BrickDims3[2] = 32;
    P.NLevels = 1 + rand() % 5;
    P.BricksPerChunk = rand() % 2 ? 4096 : bpcv[rand() % 7];
    P.ChunksPerFile = rand() % 2 ? 64 : cpfv[rand() % 5];
    P.FilesPerDir = rand() % 2 ? 64 : fpdv[rand() % 5];
    P.Tolerance = 1e-3;
    P.OutDir = "/tmp/nonexistent_idx2g";
    int fd[2];
    if (pipe(fd)) return 2;
    pid_t pid = fork();
    if (pid == 0) {
      close(fd[0]);
      int rc = int(RefFinalize(&P).Code);
      if (write(fd[1], &rc, sizeof rc) != sizeof rc) _exit(3);
      _exit(0);
    }
    close(fd[1]);
    int rcode = -1;
    const bool got = read(fd[0], &rcode, sizeof rcode) == sizeof rcode;
    close(fd[0]);
    int st = 0;
    waitpid(pid, &st, 0);
    if (!got || !WIFEXITED(st) || WEXITSTATUS(st) != 0) { ++aborted; continue; }
    auto g = idx2g_idx2::Validate(P);
    if (rcode == 0 && int(g.Code) == IDX2G_NOT_SUPPORTED && P.BrickDims3[0] > 128) { ++limit; continue; }
    if (rcode != int(g.Code)) {
      if (bad++ < 20)
        printf("mismatch dims %d %d %d brick %d %d %d lv %d bpc %d cpf %d fpd %d: ref %d ours %d\n", P.Meta.Dims3[0],
               P.Meta.Dims3[1], P.Meta.Dims3[2], P.BrickDims3[0], P.BrickDims3[1], P.BrickDims3[2], P.NLevels,
               P.BricksPerChunk, P.ChunksPerFile, P.FilesPerDir, rcode, int(g.Code));
    }
  }
  printf("%d cases, %d mismatches, %d reference aborts, %d bricks above 128\n", n, bad, aborted, limit);
  return bad != 0;
}
