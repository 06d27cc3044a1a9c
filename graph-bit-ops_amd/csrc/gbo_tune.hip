// Autotuner: picks the block size, chunk size (tiles per wavefront) and bs
// 8/16 kernel variant that minimise the time of a bsrbmv sweep on a given
// graph, by timing each candidate with HIP events.  No reference
// counterpart (the reference fixes the block size at compile time with
// -DBLOCKSIZE, bsrbmv/Makefile:57-67).  Results are cached per graph
// signature (n, nnz, op, candidate set, sampled rowptr hash) for the life of
// the process.
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace {

using Key = std::tuple<int32_t, int64_t, int, int, uint64_t>;
std::map<Key, gbo_tuning> g_cache;
std::mutex g_mu;

uint64_t sample_hash(const int32_t* d_rowptr, int32_t nrows, hipStream_t st) {
  // 64 evenly spaced rowptr entries
  std::vector<int32_t> h(64);
  uint64_t hs = 1469598103934665603ull;
  for (int i = 0; i < 64; ++i) {
    const int64_t idx = (int64_t)nrows * i / 63;
    (void)hipMemcpyAsync(&h[i], d_rowptr + idx, sizeof(int32_t), hipMemcpyDeviceToHost, st);
  }
  (void)hipStreamSynchronize(st);
  for (int32_t v : h) hs = (hs ^ (uint32_t)v) * 1099511628211ull;
  return hs;
}

}  // namespace

using namespace gbo;

extern "C" {

int gbo_autotune(const int32_t* d_rowptr, const int32_t* d_colind, int32_t nrows, int32_t ncols, int op,
                 int bs_mask, int reps, gbo_tuning* out, gbo_stream_t stream) {
  if (!d_rowptr || !out || nrows < 0 || ncols < 0) return set_error(GBO_EINVAL, "gbo_autotune: bad argument");
  if (op != 0 && op != 1) return set_error(GBO_EINVAL, "gbo_autotune: op must be 0 (count) or 1 (bfs step)");
  if (op == 1 && nrows != ncols) return set_error(GBO_EINVAL, "gbo_autotune: bfs step needs a square matrix");
  if ((bs_mask & 0xF) == 0) return set_error(GBO_EINVAL, "gbo_autotune: empty block-size mask");
  if (reps < 1) reps = 3;
  hipStream_t st = stream;
  int32_t nnz = 0;
  GBO_HIP_CHECK(hipMemcpyAsync(&nnz, d_rowptr + nrows, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  const Key key{nrows, (int64_t)nnz, op, bs_mask & 0xF, sample_hash(d_rowptr, nrows, st)};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      *out = it->second;
      return GBO_OK;
    }
  }
  gbo_tuning best{0, 0, 0, 1e30f};
  hipEvent_t e0, e1;
  GBO_HIP_CHECK(hipEventCreate(&e0));
  GBO_HIP_CHECK(hipEventCreate(&e1));
  int rc = GBO_OK;
  for (int k = 0; k < 4 && rc == GBO_OK; ++k) {
    if (!(bs_mask & (1 << k))) continue;
    const int bs = 8 << k;
    gbo_bsr A{};
    rc = gbo_bsr_create_dev(bs, 0, d_rowptr, d_colind, nrows, ncols, &A, st);
    if (rc) break;
    const int64_t nw = (std::max(nrows, ncols) + bs - 1) / bs + 1;
    void *x = nullptr, *y = nullptr, *vis = nullptr;
    int64_t* cnt = nullptr;
    GBO_HIP_CHECK(hipMalloc(&x, nw * bs / 8));
    GBO_HIP_CHECK(hipMalloc(&vis, nw * bs / 8));
    GBO_HIP_CHECK(hipMalloc(&y, std::max<size_t>((size_t)nrows * 4, nw * bs / 8)));
    GBO_HIP_CHECK(hipMalloc((void**)&cnt, sizeof(int64_t)));
    GBO_HIP_CHECK(hipMemsetAsync(x, 0x55, nw * bs / 8, st));   // half the frontier / x bits set
    GBO_HIP_CHECK(hipMemsetAsync(vis, 0, nw * bs / 8, st));
    const int dc = default_chunk(bs);
    const int chunks[3] = {dc, dc * 2, dc / 2};
    // kernel variants that differ for this block size and operation (bs 8
    // and 16 BFS: 3 is the dense count-based step; 2 would run variant 0)
    static const int kVarCount[] = {0, 1, 2, 3}, kVarBfs[] = {0, 1, 3}, kVarOne[] = {0};
    const int* vars = kVarOne;
    int nvar = 1;
    if (bs <= 16) {
      vars = op == 0 ? kVarCount : kVarBfs;
      nvar = op == 0 ? 4 : 3;
    }
    for (int vi = 0; vi < nvar && rc == GBO_OK; ++vi)
      for (int ci = 0; ci < 3 && rc == GBO_OK; ++ci) {
        const int v = vars[vi];
        rc = gbo_bsr_plan_ex(&A, chunks[ci], v, st);
        if (rc) break;
        auto run = [&]() {
          return op == 0 ? gbo_bsrbmv_count(&A, x, (int32_t*)y, 0, st)
                         : gbo_bsrbmv_bfs_step(&A, x, vis, y, cnt, st);
        };
        rc = run();  // warm-up
        if (rc) break;
        GBO_HIP_CHECK(hipEventRecord(e0, st));
        for (int r = 0; r < reps && rc == GBO_OK; ++r) rc = run();
        GBO_HIP_CHECK(hipEventRecord(e1, st));
        GBO_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        GBO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        ms /= reps;
        if (ms < best.ms) best = gbo_tuning{bs, chunks[ci], v, ms};
      }
    (void)hipFree(x);
    (void)hipFree(vis);
    (void)hipFree(y);
    (void)hipFree(cnt);
    gbo_bsr_free(&A);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
  }
  *out = best;
  return GBO_OK;
}

}  // extern "C"
