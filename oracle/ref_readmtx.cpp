// Thin C-ABI shim over the reference's GraphBLAST-style reader
// (algorithm/readMtx.hpp: readMtx :352, coo2csr :247), compiled from the
// reference sources in place (see Makefile target `ref`).  Test
// infrastructure only: it pins the engine's reader semantics 1.
#include <cstdlib>
#include <vector>

#include "readMtx.hpp"

extern "C" int ref_readmtx(const char* path, int directed, int* nrows, int* ncols, int* nnz, int** rowptr,
                           int** colind) {
  std::vector<int> r, c;
  std::vector<float> v;
  readMtx<float>(path, &r, &c, &v, nrows, ncols, nnz, directed, false, nullptr);
  const int n = (int)r.size();
  *rowptr = (int*)std::malloc(sizeof(int) * (*nrows + 1));
  *colind = (int*)std::malloc(sizeof(int) * (n > 0 ? n : 1));
  float* val = (float*)std::malloc(sizeof(float) * (n > 0 ? n : 1));
  coo2csr(*rowptr, *colind, val, r, c, v, *nrows, *ncols);
  std::free(val);
  *nnz = n;
  return 0;
}

extern "C" void ref_free(void* p) { std::free(p); }
