// Shared declarations between the packer, the bsrbmv/bsrbmm operations and
// the drivers (device-side translation units).
#pragma once

#include <hip/hip_runtime.h>

#include "gbo_device.hpp"

namespace gbo {

// Work decomposition of a bit-BSR: the tiles [0, nblocks) are cut into
// chunks of `chunk` consecutive tiles; plan[c] is the first block row that
// chunk c touches (a row that begins before the chunk is "split" and its
// partial results are combined across chunks).
//
// The scratch the operations need besides their inputs and outputs is kept
// in the plan's allocation, after the nchunks plan entries, so that a call
// on a planned matrix allocates nothing (each part 256-byte aligned):
//   slots       kCountSlots x u64   BFS frontier-size slots (zero between calls)
//   count       1 x int64           the folded frontier size
//   partial     nchunks*2*bs x f64  split-row partials of the FP32 sum
//   bfs_counts  bs*(nbr+1) x int32  per-row counts of the dense BFS step
// Calls on one matrix therefore share that scratch: they must be ordered
// (one stream, or synchronised); separately created matrices do not share.
struct PlanLayout {
  size_t slots = 0, count = 0, partial = 0, bfs_counts = 0, bytes = 0;
};
PlanLayout plan_layout(int bs, int64_t nchunks, int32_t nbr, bool scratch);

struct PlanView {
  const int32_t* plan = nullptr;
  int32_t chunk = 0;
  int32_t nchunks = 0;
  int32_t variant = 0;
  bool owned = false;
  unsigned long long* slots = nullptr;
  int64_t* count = nullptr;
  double* partial = nullptr;
  int32_t* bfs_counts = nullptr;  // the dense BFS step's row counts
};

int default_chunk(int bs);
int acquire_plan(const gbo_bsr* A, PlanView* pv, hipStream_t st);
int release_plan(PlanView* pv, hipStream_t st);
// Allocates A->plan (chunk tiles per chunk, 0 = default); with `scratch` the
// allocation also holds the scratch parts above.
int plan_create(gbo_bsr* A, int32_t chunk_tiles, int32_t variant, bool scratch, hipStream_t st);

// A plan acquired for the length of one call, released on every return path.
struct PlanGuard {
  PlanView pv;
  hipStream_t st;
  explicit PlanGuard(hipStream_t s) : st(s) {}
  ~PlanGuard() { (void)release_plan(&pv, st); }
  PlanGuard(const PlanGuard&) = delete;
  PlanGuard& operator=(const PlanGuard&) = delete;
};

// Stream-ordered device allocations of one call, released with hipFreeAsync
// on every return path (success or error).
class AsyncAllocs {
 public:
  explicit AsyncAllocs(hipStream_t st) : st_(st) {}
  ~AsyncAllocs() {
    for (int i = 0; i < n_; ++i) (void)hipFreeAsync(ptrs_[i], st_);
  }
  AsyncAllocs(const AsyncAllocs&) = delete;
  AsyncAllocs& operator=(const AsyncAllocs&) = delete;
  template <class T>
  int alloc(T** p, size_t bytes, const char* who) {
    *p = nullptr;
    if (n_ == kMax) return set_error(GBO_EINVAL, "%s: too many scratch allocations", who);
    void* q = nullptr;
    const hipError_t e = hipMallocAsync(&q, bytes ? bytes : 8, st_);
    if (e != hipSuccess) return set_error(GBO_ENOMEM, "%s: device allocation of %zu bytes failed: %s", who, bytes,
                                          hipGetErrorString(e));
    ptrs_[n_++] = q;
    *p = (T*)q;
    return GBO_OK;
  }

 private:
  static constexpr int kMax = 16;
  hipStream_t st_;
  void* ptrs_[kMax] = {};
  int n_ = 0;
};
int validate_bsr(const gbo_bsr* A, int want_dir, const char* who);

template <class TI, class TO>
int scan_exclusive(const TI* in, TO* out, int64_t n, hipStream_t st);

constexpr uint32_t kInfU32 = 0xFFFFFFFFu;
constexpr int kCountSlots = 1024;

// bsrbmv operation codes.
enum SpmvOp { OP_COUNT = 0, OP_BFS = 1, OP_MINPLUS = 2, OP_SUM = 3, OP_MIN = 4 };

// Whether a BFS step at this block size and plan variant may take the dense
// form (count kernel into per-row counts, then a pack of the nonzero rows):
// bs 8/16 under variant 3, bs 32/64 under every variant.
inline bool dense_bfs_step(int bs, int variant) { return (bs <= 16 && variant == 3) || bs >= 32; }

struct SpmvArgs {
  const int32_t* rowptr;
  const int32_t* colind;
  const void* tiles;
  int32_t nrows, ncols, nbr;
  int64_t T;
  const int32_t* plan;
  int32_t chunk, nchunks;
  int32_t variant;      // bs 8/16 COUNT/BFS kernel variant
  const void* xbits;    // COUNT: x bits; BFS: frontier bits
  const void* visited;  // BFS: visited bits (per block row)
  void* out;            // COUNT int32 y / BFS next bits / MINPLUS uint32 / SUM float / MIN int32
  const uint32_t* w;    // MINPLUS weights in tile bit order (nullable)
  const int64_t* woff;  // MINPLUS per-tile weight offsets
  const void* xval;     // MINPLUS dist / SUM x / MIN x
  unsigned long long* count_slots;  // BFS: newly set bits, summed per slot (kCountSlots, zero on entry)
  int32_t* bfs_counts;  // BFS where dense_bfs_step holds (nullable): int32 scratch of nbr*bs; when set the step runs
                        // the count kernel into it and packs the rows with a nonzero count (dense step)
  double* partial;      // SUM: split-row partials [nchunks][2][bs]
  const uint64_t* active;  // MINPLUS: optional per-block-column activity bits (nullable)
  int sparse_frontier;  // BFS: most frontier words are zero, so a tile is read only where its
                        // block column's frontier word is nonzero (row-group kernel)
  int accumulate;       // COUNT: add into y
  float scale, bias;    // SUM: y = scale * (A x) + bias  (scale 1, bias 0 for the plain op)
};

int spmv_launch(int bs, int op, const SpmvArgs& a, hipStream_t st);

// BFS frontier-size counters.  Every wave adds its count of newly set bits
// to slot (chunk index mod kCountSlots) instead of one shared counter, so
// hundreds of thousands of waves do not serialise on a single address;
// count_fold() sums the slots into one int64 and clears them again.  The
// slots live in the plan's scratch (zeroed when the plan is made).
int count_fold(unsigned long long* slots, int64_t* d_count, hipStream_t st);

}  // namespace gbo
