// Host-side C++ of the engine: error state, Matrix Market readers, synthetic
// graph generators and CSR utilities.  No GPU code here (built with g++ and
// OpenMP); the device side lives in the .hip files.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gbo_internal.hpp"

namespace gbo {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
void clear_error() { g_last_error.clear(); }

template <class T>
static T* host_alloc(size_t n) {
  T* p = static_cast<T*>(std::malloc(std::max<size_t>(n, 1) * sizeof(T)));
  if (!p) throw std::bad_alloc();
  return p;
}

// ------------------------------------------------------------------ CSR build
// Builds a sorted, duplicate-free CSR from per-row buckets filled in parallel.
// `adj` holds row r's entries in [off[r], off[r] + cnt[r]).
static void finalize_rows(int32_t nrows, const std::vector<int64_t>& off, std::vector<int32_t>& adj,
                          std::vector<int64_t>& cnt, int64_t* nnz_out, int32_t** rowptr_out,
                          int32_t** colind_out) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t r = 0; r < nrows; ++r) {
    int32_t* b = adj.data() + off[r];
    int32_t* e = b + cnt[r];
    std::sort(b, e);
    cnt[r] = std::unique(b, e) - b;
  }
  int32_t* rowptr = host_alloc<int32_t>((size_t)nrows + 1);
  int64_t total = 0;
  for (int32_t r = 0; r < nrows; ++r) {
    rowptr[r] = (int32_t)total;
    total += cnt[r];
  }
  rowptr[nrows] = (int32_t)total;
  int32_t* colind = host_alloc<int32_t>((size_t)total);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t r = 0; r < nrows; ++r)
    std::memcpy(colind + rowptr[r], adj.data() + off[r], sizeof(int32_t) * cnt[r]);
  *nnz_out = total;
  *rowptr_out = rowptr;
  *colind_out = colind;
}

// Symmetric CSR from an undirected edge list given as a callable
// edge(k, &u, &v) -> bool (false = no edge at index k).  Self-loops dropped,
// duplicates removed, rows sorted.
template <class EdgeFn>
static int build_symmetric(int32_t n, int64_t m, EdgeFn edge, int64_t* nnz, int32_t** rowptr,
                           int32_t** colind) {
  if (n < 0 || m < 0) return set_error(GBO_EINVAL, "negative size");
  if (n > 0 && (2 * m) >= ((int64_t)1 << 31) * 4)
    return set_error(GBO_ERANGE, "edge count too large");
  std::vector<std::atomic<int64_t>> deg(n);
  for (auto& d : deg) d.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    int32_t u, v;
    if (edge(k, &u, &v) && u != v) {
      deg[u].fetch_add(1, std::memory_order_relaxed);
      deg[v].fetch_add(1, std::memory_order_relaxed);
    }
  }
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int32_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i].load(std::memory_order_relaxed);
  std::vector<int32_t> adj((size_t)off[n]);
  std::vector<std::atomic<int64_t>>& cur = deg;
  for (int32_t i = 0; i < n; ++i) cur[i].store(off[i], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    int32_t u, v;
    if (edge(k, &u, &v) && u != v) {
      adj[cur[u].fetch_add(1, std::memory_order_relaxed)] = v;
      adj[cur[v].fetch_add(1, std::memory_order_relaxed)] = u;
    }
  }
  std::vector<int64_t> cnt((size_t)n);
  for (int32_t i = 0; i < n; ++i) cnt[i] = off[i + 1] - off[i];
  if (off[n] >= ((int64_t)1 << 31)) return set_error(GBO_ERANGE, "nnz exceeds int32 CSR");
  finalize_rows(n, off, adj, cnt, nnz, rowptr, colind);
  return GBO_OK;
}

}  // namespace gbo

using namespace gbo;

extern "C" {

const char* gbo_last_error(void) { return g_last_error.c_str(); }
const char* gbo_version(void) { return "gbo 0.1 (gfx950)"; }
void gbo_free_host(void* p) { std::free(p); }

// ------------------------------------------------------------- generators
// Erdos-Renyi G(n, p): unordered pair {u < v} is an edge iff the 53-bit draw
// for pair index u*n+v is below p * 2^53.  O(n^2): for small n.
int gbo_gen_er(int32_t n, double p, uint64_t seed, int64_t* nnz, int32_t** rp, int32_t** ci) {
  GBO_TRY_BEGIN
  if (n < 0 || !(p >= 0.0 && p <= 1.0) || !nnz || !rp || !ci)
    return set_error(GBO_EINVAL, "gbo_gen_er: bad argument");
  if ((int64_t)n * n > ((int64_t)1 << 34)) return set_error(GBO_ERANGE, "gbo_gen_er: n too large");
  const uint64_t thr = (uint64_t)(p * 9007199254740992.0);
  const int64_t npairs = (int64_t)n * n;
  auto edge = [&](int64_t k, int32_t* u, int32_t* v) {
    *u = (int32_t)(k / n);
    *v = (int32_t)(k % n);
    if (*u >= *v) return false;
    return (rand64(seed, STREAM_ER, (uint64_t)k) >> 11) < thr;
  };
  return build_symmetric(n, npairs, edge, nnz, rp, ci);
  GBO_TRY_END
}

// Erdos-Renyi G(n, m) variant: m draws of (u, v) uniform, self-pairs
// dropped, duplicates merged.  Integer arithmetic only.
int gbo_gen_er_m(int32_t n, int64_t m, uint64_t seed, int64_t* nnz, int32_t** rp, int32_t** ci) {
  GBO_TRY_BEGIN
  if (n <= 0 || m < 0 || !nnz || !rp || !ci) return set_error(GBO_EINVAL, "gbo_gen_er_m: bad argument");
  auto edge = [&](int64_t k, int32_t* u, int32_t* v) {
    const uint64_t r = rand64(seed, STREAM_ER_M, (uint64_t)k);
    *u = (int32_t)rand_below(r, (uint32_t)n);
    *v = (int32_t)rand_below(r << 32, (uint32_t)n);
    return true;
  };
  return build_symmetric(n, m, edge, nnz, rp, ci);
  GBO_TRY_END
}

// R-MAT (Graph500 quadrant recursion, no noise, no permutation): edge k
// descends `scale` levels; level l draws r = rand64(seed, RMAT, k*scale+l)>>11
// and picks quadrant (0,0) if r < A, (0,1) if r < A+B, (1,0) if r < A+B+C,
// else (1,1), with A = a*2^53 etc.  Symmetrized, self-loops dropped,
// duplicates removed.
int gbo_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c, uint64_t seed,
                 int64_t* nnz, int32_t** rp, int32_t** ci) {
  GBO_TRY_BEGIN
  if (scale < 1 || scale > 30 || edgefactor < 0 || !nnz || !rp || !ci)
    return set_error(GBO_EINVAL, "gbo_gen_rmat: bad argument");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return set_error(GBO_EINVAL, "gbo_gen_rmat: bad probabilities");
  const int32_t n = (int32_t)1 << scale;
  const int64_t m = (int64_t)edgefactor << scale;
  const double two53 = 9007199254740992.0;
  const uint64_t ta = (uint64_t)(a * two53), tab = (uint64_t)((a + b) * two53),
                 tabc = (uint64_t)((a + b + c) * two53);
  auto edge = [&](int64_t k, int32_t* u, int32_t* v) {
    uint32_t uu = 0, vv = 0;
    for (int l = 0; l < scale; ++l) {
      const uint64_t r = rand64(seed, STREAM_RMAT, (uint64_t)k * scale + l) >> 11;
      const uint32_t bu = (r >= tab) ? 1u : 0u;
      const uint32_t bv = (r >= ta && r < tab) || (r >= tabc) ? 1u : 0u;
      uu = (uu << 1) | bu;
      vv = (vv << 1) | bv;
    }
    *u = (int32_t)uu;
    *v = (int32_t)vv;
    return true;
  };
  return build_symmetric(n, m, edge, nnz, rp, ci);
  GBO_TRY_END
}

// m x m grid, 8-neighbour (king-move) stencil; vertex id = i*m + j.
int gbo_gen_king(int32_t m, int64_t* nnz, int32_t** rp_out, int32_t** ci_out) {
  GBO_TRY_BEGIN
  if (m < 0 || m > 16384 || !nnz || !rp_out || !ci_out)
    return set_error(GBO_EINVAL, "gbo_gen_king: bad argument");
  const int32_t n = m * m;
  int32_t* rp = host_alloc<int32_t>((size_t)n + 1);
  rp[0] = 0;
  for (int32_t v = 0; v < n; ++v) {
    const int32_t i = v / m, j = v % m;
    int d = 0;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        if ((di || dj) && i + di >= 0 && i + di < m && j + dj >= 0 && j + dj < m) ++d;
    rp[v + 1] = rp[v] + d;
  }
  int32_t* ci = host_alloc<int32_t>((size_t)rp[n]);
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < n; ++v) {
    const int32_t i = v / m, j = v % m;
    int32_t o = rp[v];
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        if ((di || dj) && i + di >= 0 && i + di < m && j + dj >= 0 && j + dj < m)
          ci[o++] = (i + di) * m + (j + dj);
  }
  *nnz = rp[n];
  *rp_out = rp;
  *ci_out = ci;
  return GBO_OK;
  GBO_TRY_END
}

int gbo_gen_edge_weights(int32_t nrows, const int32_t* rp, const int32_t* ci, uint32_t lo,
                         uint32_t hi, uint64_t seed, uint32_t* w) {
  if (nrows < 0 || !rp || !w || lo > hi) return set_error(GBO_EINVAL, "gbo_gen_edge_weights: bad argument");
  const uint64_t span = (uint64_t)hi - lo + 1;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t r = 0; r < nrows; ++r)
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
      const uint32_t u = (uint32_t)std::min(r, ci[k]), v = (uint32_t)std::max(r, ci[k]);
      const uint64_t key = ((uint64_t)u << 32) | v;
      w[k] = lo + (uint32_t)(rand64(seed, STREAM_WEIGHT, key) % span);
    }
  return GBO_OK;
}

int gbo_csr_transpose(int32_t nrows, int32_t ncols, const int32_t* rp, const int32_t* ci,
                      int32_t** tptr, int32_t** tind) {
  GBO_TRY_BEGIN
  if (nrows < 0 || ncols < 0 || !rp || !tptr || !tind) return set_error(GBO_EINVAL, "gbo_csr_transpose: bad argument");
  const int64_t nnz = rp[nrows];
  int32_t* tp = host_alloc<int32_t>((size_t)ncols + 1);
  int32_t* ti = host_alloc<int32_t>((size_t)nnz);
  std::fill(tp, tp + ncols + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) {
    if (ci[k] < 0 || ci[k] >= ncols) {
      std::free(tp);
      std::free(ti);
      return set_error(GBO_EINVAL, "column index out of range");
    }
    tp[ci[k] + 1]++;
  }
  for (int32_t j = 0; j < ncols; ++j) tp[j + 1] += tp[j];
  std::vector<int32_t> cur(tp, tp + ncols);
  for (int32_t r = 0; r < nrows; ++r)
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) ti[cur[ci[k]]++] = r;
  *tptr = tp;
  *tind = ti;
  return GBO_OK;
  GBO_TRY_END
}

int gbo_csr_tril(int32_t nrows, const int32_t* rp, const int32_t* ci, int strict, int64_t* nnz,
                 int32_t** lptr, int32_t** lind) {
  GBO_TRY_BEGIN
  if (nrows < 0 || !rp || !nnz || !lptr || !lind) return set_error(GBO_EINVAL, "gbo_csr_tril: bad argument");
  int32_t* lp = host_alloc<int32_t>((size_t)nrows + 1);
  lp[0] = 0;
  for (int32_t r = 0; r < nrows; ++r) {
    int32_t d = 0;
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) d += strict ? (ci[k] < r) : (ci[k] <= r);
    lp[r + 1] = lp[r] + d;
  }
  int32_t* li = host_alloc<int32_t>((size_t)lp[nrows]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t r = 0; r < nrows; ++r) {
    int32_t o = lp[r];
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
      if (strict ? (ci[k] < r) : (ci[k] <= r)) li[o++] = ci[k];
  }
  *nnz = lp[nrows];
  *lptr = lp;
  *lind = li;
  return GBO_OK;
  GBO_TRY_END
}

}  // extern "C"
