// Thin C-ABI shim over the reference's own Matrix Market reader
// (algorithm/mmio_highlevel.h: mmio_info :7, mmio_data :146), compiled from
// the reference sources in place (see Makefile target `ref`).  Test
// infrastructure only: it pins the engine's reader semantics 0.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mmio_highlevel.h"

extern "C" int ref_mmio_info(const char* path, int* m, int* n, int* nnz, int* sym) {
  return mmio_info<float>(m, n, nnz, sym, path);
}

extern "C" int ref_mmio_data(const char* path, int* rowptr, int* colind, float* val) {
  return mmio_data<float>(rowptr, colind, val, path);
}
