/*
 * oracle.c -- CPU restatement of the reference's bit-BSR graph path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C99 + OpenMP.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ helpers -- */
static inline uint64_t getw(const void* t, int bs, int64_t idx) {
  switch (bs) {
    case 8: return ((const uint8_t*)t)[idx];
    case 16: return ((const uint16_t*)t)[idx];
    case 32: return ((const uint32_t*)t)[idx];
    default: return ((const uint64_t*)t)[idx];
  }
}
static inline void setw(void* t, int bs, int64_t idx, uint64_t v) {
  switch (bs) {
    case 8: ((uint8_t*)t)[idx] = (uint8_t)v; break;
    case 16: ((uint16_t*)t)[idx] = (uint16_t)v; break;
    case 32: ((uint32_t*)t)[idx] = (uint32_t)v; break;
    default: ((uint64_t*)t)[idx] = v; break;
  }
}
static inline int popc(uint64_t x) { return __builtin_popcountll(x); }
static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}
/* index of key in sorted a[lo,hi) or -1 (binarySearch, bsrbmm/bsrbmm.cu:628) */
static int64_t bsearch_i32(const int32_t* a, int64_t lo, int64_t hi, int32_t key) {
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] == key) return mid;
    if (a[mid] > key) hi = mid;
    else lo = mid + 1;
  }
  return -1;
}

/* ---------------------------------------------------------- generators -- */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
uint64_t or_rand64(uint64_t seed, uint64_t stream, uint64_t idx) {
  const uint64_t key = mix64(seed * 0x9e3779b97f4a7c15ULL + stream + 1);
  return mix64(key + (idx + 1) * 0x9e3779b97f4a7c15ULL);
}

int64_t or_gen_er_edges(int32_t n, double p, uint64_t seed, int32_t* u, int32_t* v, int64_t cap) {
  const uint64_t thr = (uint64_t)(p * 9007199254740992.0);
  int64_t m = 0;
  for (int64_t a = 0; a < n; ++a)
    for (int64_t b = a + 1; b < n; ++b)
      if ((or_rand64(seed, 1, (uint64_t)(a * n + b)) >> 11) < thr) {
        if (m < cap) {
          u[m] = (int32_t)a;
          v[m] = (int32_t)b;
        }
        ++m;
      }
  return m;
}

int64_t or_gen_rmat_edges(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, int32_t* u,
                          int32_t* v) {
  const double two53 = 9007199254740992.0;
  const uint64_t ta = (uint64_t)(a * two53), tab = (uint64_t)((a + b) * two53), tabc = (uint64_t)((a + b + c) * two53);
  const int64_t m = (int64_t)ef << scale;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    uint32_t uu = 0, vv = 0;
    for (int l = 0; l < scale; ++l) {
      const uint64_t r = or_rand64(seed, 2, (uint64_t)k * scale + l) >> 11;
      int bu, bv;
      if (r < ta) { bu = 0; bv = 0; }
      else if (r < tab) { bu = 0; bv = 1; }
      else if (r < tabc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      uu = (uu << 1) | (uint32_t)bu;
      vv = (vv << 1) | (uint32_t)bv;
    }
    u[k] = (int32_t)uu;
    v[k] = (int32_t)vv;
  }
  return m;
}

/* ------------------------------------------------------------- packing -- */
/* Sorted distinct block columns of block row I into buf; returns count. */
static int64_t block_cols(int bs, int32_t nrows, const int32_t* rp, const int32_t* ci, int32_t I, int32_t* buf) {
  int64_t k = 0;
  const int64_t r0 = (int64_t)I * bs, r1 = r0 + bs < nrows ? r0 + bs : nrows;
  for (int64_t r = r0; r < r1; ++r)
    for (int32_t q = rp[r]; q < rp[r + 1]; ++q) buf[k++] = ci[q] / bs;
  qsort(buf, (size_t)k, sizeof(int32_t), cmp_i32);
  int64_t u = 0;
  for (int64_t i = 0; i < k; ++i)
    if (u == 0 || buf[u - 1] != buf[i]) buf[u++] = buf[i];
  return u;
}

static int64_t max_block_nnz(int bs, int32_t nrows, const int32_t* rp) {
  const int32_t nbr = (nrows + bs - 1) / bs;
  int64_t mx = 1;
  for (int32_t I = 0; I < nbr; ++I) {
    const int64_t r0 = (int64_t)I * bs, r1 = r0 + bs < nrows ? r0 + bs : nrows;
    const int64_t c = rp[r1] - rp[r0];
    if (c > mx) mx = c;
  }
  return mx;
}

int64_t or_pack_count(int bs, int32_t nrows, const int32_t* rp, const int32_t* ci, int32_t* brp) {
  const int32_t nbr = (nrows + bs - 1) / bs;
  int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)max_block_nnz(bs, nrows, rp));
  brp[0] = 0;
  for (int32_t I = 0; I < nbr; ++I) brp[I + 1] = brp[I] + (int32_t)block_cols(bs, nrows, rp, ci, I, buf);
  free(buf);
  return brp[nbr];
}

void or_pack_fill(int bs, int dir, int32_t nrows, const int32_t* rp, const int32_t* ci, const int32_t* brp,
                  int32_t* bci, void* tiles) {
  const int32_t nbr = (nrows + bs - 1) / bs;
  memset(tiles, 0, (size_t)brp[nbr] * bs * (bs / 8));
  int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)max_block_nnz(bs, nrows, rp));
  for (int32_t I = 0; I < nbr; ++I) {
    const int64_t nt = block_cols(bs, nrows, rp, ci, I, buf);
    for (int64_t k = 0; k < nt; ++k) bci[brp[I] + k] = buf[k];
    const int64_t r0 = (int64_t)I * bs, r1 = r0 + bs < nrows ? r0 + bs : nrows;
    for (int64_t r = r0; r < r1; ++r)
      for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
        const int32_t J = ci[q] / bs, c = ci[q] % bs, rr = (int32_t)(r - r0);
        const int64_t t = brp[I] + bsearch_i32(bci + brp[I], 0, nt, J);
        const int64_t widx = t * bs + (dir == 0 ? rr : c);
        const int bit = bs - 1 - (dir == 0 ? c : rr);
        setw(tiles, bs, widx, getw(tiles, bs, widx) | (1ULL << bit));
      }
  }
  free(buf);
}

void or_pack_vec(int bs, int32_t n, const float* x, void* xbits) {
  const int32_t nw = (n + bs - 1) / bs;
  for (int32_t j = 0; j < nw; ++j) {
    uint64_t w = 0;
    for (int k = 0; k < bs; ++k) {
      const int64_t i = (int64_t)j * bs + k;
      w = (w << 1) | (uint64_t)((i < n) && x[i] > 0.0f);
    }
    setw(xbits, bs, j, w);
  }
}

void or_pack_values(int bs, int32_t nrows, const int32_t* rp, const int32_t* ci, const uint32_t* val,
                    const int32_t* brp, const int32_t* bci, const void* tiles, uint32_t* tval, int64_t* woff) {
  const int32_t nbr = (nrows + bs - 1) / bs;
  const int64_t T = brp[nbr];
  woff[0] = 0;
  for (int64_t t = 0; t < T; ++t) {
    int64_t s = 0;
    for (int r = 0; r < bs; ++r) s += popc(getw(tiles, bs, t * bs + r));
    woff[t + 1] = woff[t] + s;
  }
  for (int64_t i = 0; i < woff[T]; ++i) tval[i] = 0xFFFFFFFFu;
  /* The k-th set bit of a tile in (row, column) order holds the minimum
   * value among the CSR entries at that position. */
  for (int32_t I = 0; I < nbr; ++I) {
    const int64_t r0 = (int64_t)I * bs, r1 = r0 + bs < nrows ? r0 + bs : nrows;
    for (int64_t r = r0; r < r1; ++r)
      for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
        const int32_t J = ci[q] / bs, c = ci[q] % bs, rr = (int32_t)(r - r0);
        const int64_t t = brp[I] + bsearch_i32(bci + brp[I], 0, brp[I + 1] - brp[I], J);
        int64_t rank = 0;
        for (int k = 0; k < rr; ++k) rank += popc(getw(tiles, bs, t * bs + k));
        const uint64_t w = getw(tiles, bs, t * bs + rr);
        rank += c > 0 ? popc(w >> (bs - c)) : 0; /* bits of columns < c */
        uint32_t* slot = &tval[woff[t] + rank];
        if (val[q] < *slot) *slot = val[q];
      }
  }
}

/* --------------------------------------------------------------- bsrbmv -- */
void or_bsrbmv_count(int bs, int32_t nrows, int32_t r0, int32_t r1, const int32_t* brp, const int32_t* bci,
                     const void* tiles, const void* xbits, int32_t* y) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t I = r0; I < r1; ++I) {
    int32_t acc[64];
    for (int r = 0; r < bs; ++r) acc[r] = 0;
    for (int32_t t = brp[I]; t < brp[I + 1]; ++t) {
      const uint64_t x = getw(xbits, bs, bci[t]);
      for (int r = 0; r < bs; ++r) acc[r] += popc(getw(tiles, bs, (int64_t)t * bs + r) & x);
    }
    for (int r = 0; r < bs; ++r) {
      const int64_t i = (int64_t)I * bs + r;
      if (i < nrows) y[i] = acc[r];
    }
  }
}

int64_t or_bsrbmv_bfs_step(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                           const void* frontier, const void* visited, void* next) {
  const int32_t nbr = (nrows + bs - 1) / bs;
  int64_t total = 0;
  for (int32_t I = 0; I < nbr; ++I) {
    uint64_t word = 0;
    for (int r = 0; r < bs; ++r) {
      int64_t cnt = 0;
      for (int32_t t = brp[I]; t < brp[I + 1]; ++t)
        cnt += popc(getw(tiles, bs, (int64_t)t * bs + r) & getw(frontier, bs, bci[t]));
      if (cnt > 0) word |= 1ULL << (bs - 1 - r);
    }
    word &= ~getw(visited, bs, I);
    if (bs < 64) word &= (1ULL << bs) - 1;
    setw(next, bs, I, word);
    total += popc(word);
  }
  return total;
}

void or_bsrbmv_minplus(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                       const uint32_t* w, const int64_t* woff, const uint32_t* dist, uint32_t* out) {
  const int32_t nbr = (nrows + bs - 1) / bs;
  for (int32_t I = 0; I < nbr; ++I)
    for (int r = 0; r < bs; ++r) {
      const int64_t i = (int64_t)I * bs + r;
      if (i >= nrows) break;
      uint32_t best = dist[i];
      for (int32_t t = brp[I]; t < brp[I + 1]; ++t) {
        int64_t rank = 0;
        if (w)
          for (int k = 0; k < r; ++k) rank += popc(getw(tiles, bs, (int64_t)t * bs + k));
        const uint64_t wd = getw(tiles, bs, (int64_t)t * bs + r);
        for (int c = 0; c < bs; ++c)
          if ((wd >> (bs - 1 - c)) & 1) {
            const uint32_t dv = dist[(int64_t)bci[t] * bs + c];
            const uint32_t wt = w ? w[woff[t] + rank] : 1u;
            ++rank;
            const uint32_t cand = (dv > 0xFFFFFFFFu - wt) ? 0xFFFFFFFFu : dv + wt;
            if (cand < best) best = cand;
          }
      }
      out[i] = best;
    }
}

void or_bsrbmv_sum_f64(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                       const double* x, double* y) {
  const int32_t nbr = (nrows + bs - 1) / bs;
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t I = 0; I < nbr; ++I)
    for (int r = 0; r < bs; ++r) {
      const int64_t i = (int64_t)I * bs + r;
      if (i >= nrows) break;
      double s = 0.0;
      for (int32_t t = brp[I]; t < brp[I + 1]; ++t) {
        const uint64_t wd = getw(tiles, bs, (int64_t)t * bs + r);
        for (int c = 0; c < bs; ++c)
          if ((wd >> (bs - 1 - c)) & 1) s += x[(int64_t)bci[t] * bs + c];
      }
      y[i] = s;
    }
}

void or_bsrbmv_min_i32(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                       const int32_t* x, int32_t* y) {
  const int32_t nbr = (nrows + bs - 1) / bs;
  for (int32_t I = 0; I < nbr; ++I)
    for (int r = 0; r < bs; ++r) {
      const int64_t i = (int64_t)I * bs + r;
      if (i >= nrows) break;
      int32_t m = INT32_MAX;
      for (int32_t t = brp[I]; t < brp[I + 1]; ++t) {
        const uint64_t wd = getw(tiles, bs, (int64_t)t * bs + r);
        for (int c = 0; c < bs; ++c)
          if ((wd >> (bs - 1 - c)) & 1) {
            const int32_t v = x[(int64_t)bci[t] * bs + c];
            if (v < m) m = v;
          }
      }
      y[i] = m;
    }
}

/* --------------------------------------------------------------- bsrbmm -- */
uint64_t or_bsrbmm_masked(int bs, int32_t nbr, const int32_t* arp, const int32_t* aci, const void* atiles,
                          const int32_t* brp, const int32_t* bci, const void* btiles) {
  uint64_t sum = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sum)
  for (int32_t I = 0; I < nbr; ++I)
    for (int32_t i = arp[I]; i < arp[I + 1]; ++i) {
      const int32_t K = aci[i];
      for (int32_t j = arp[I]; j < arp[I + 1]; ++j) {
        const int32_t J = aci[j];
        const int64_t tb = bsearch_i32(bci, brp[J], brp[J + 1], K);
        if (tb < 0) continue;
        for (int a = 0; a < bs; ++a) {
          const uint64_t m = getw(atiles, bs, (int64_t)i * bs + a);
          const uint64_t x = getw(atiles, bs, (int64_t)j * bs + a);
          for (int k = 0; k < bs; ++k)
            if ((m >> (bs - 1 - k)) & 1) sum += (uint64_t)popc(x & getw(btiles, bs, tb * bs + k));
        }
      }
    }
  return sum;
}

uint64_t or_bsrbmm_sum(int bs, int32_t nbr, const int32_t* arp, const int32_t* aci, const void* atiles,
                       const int32_t* brp, const int32_t* bci, const void* btiles) {
  uint64_t sum = 0;
  (void)bci;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sum)
  for (int32_t I = 0; I < nbr; ++I)
    for (int32_t ta = arp[I]; ta < arp[I + 1]; ++ta) {
      const int32_t J = aci[ta];
      for (int32_t tb = brp[J]; tb < brp[J + 1]; ++tb)
        for (int a = 0; a < bs; ++a) {
          const uint64_t x = getw(atiles, bs, (int64_t)ta * bs + a);
          for (int k = 0; k < bs; ++k) sum += (uint64_t)popc(x & getw(btiles, bs, (int64_t)tb * bs + k));
        }
    }
  return sum;
}

/* C = A * B over the integers, element-level CSR (what bmm32_sparse_symbolic
 * + bmm32_sparse_numeric, bsrbmm/bsrbmm.cu:1273,1311, set out to produce:
 * C's row pointer, then per row the columns with a nonzero product, each with
 * its count sum_k A_ik B_kj as the value).  Gustavson with a dense marker per
 * thread; columns ascending.  First call with ci == NULL fills rp and returns
 * nnz; second call fills ci/val. */
int64_t or_spgemm(int32_t m, const int32_t* arp, const int32_t* aci, const int32_t* brp, const int32_t* bci,
                  int32_t ncols, int32_t* rp, int32_t* ci, int32_t* val) {
  if (!ci) {
#pragma omp parallel
    {
      int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncols > 0 ? ncols : 1));
      for (int32_t j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
      for (int32_t i = 0; i < m; ++i) {
        int32_t c = 0;
        for (int32_t p = arp[i]; p < arp[i + 1]; ++p)
          for (int32_t q = brp[aci[p]]; q < brp[aci[p] + 1]; ++q)
            if (mark[bci[q]] != i) {
              mark[bci[q]] = i;
              ++c;
            }
        rp[i + 1] = c;
      }
      free(mark);
    }
    rp[0] = 0;
    for (int32_t i = 0; i < m; ++i) rp[i + 1] += rp[i];
    return rp[m];
  }
#pragma omp parallel
  {
    int32_t* acc = (int32_t*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(int32_t));
#pragma omp for schedule(dynamic, 256)
    for (int32_t i = 0; i < m; ++i) {
      int32_t c = rp[i];
      for (int32_t p = arp[i]; p < arp[i + 1]; ++p)
        for (int32_t q = brp[aci[p]]; q < brp[aci[p] + 1]; ++q) {
          if (acc[bci[q]] == 0) ci[c++] = bci[q];
          acc[bci[q]]++;
        }
      qsort(ci + rp[i], (size_t)(c - rp[i]), sizeof(int32_t), cmp_i32);
      for (int32_t k = rp[i]; k < c; ++k) {
        val[k] = acc[ci[k]];
        acc[ci[k]] = 0;
      }
    }
    free(acc);
  }
  return rp[m];
}

/* -------------------------------------------------------------- drivers -- */
int32_t or_bfs_bsr(int bs, int32_t n, const int32_t* brp, const int32_t* bci, const void* tiles, int32_t src,
                   int32_t* level, int64_t* sizes, int32_t max_levels) {
  const int32_t nbr = (n + bs - 1) / bs;
  const size_t wb = (size_t)bs / 8;
  void* f = calloc((size_t)nbr + 1, wb);
  void* vis = calloc((size_t)nbr + 1, wb);
  void* nx = calloc((size_t)nbr + 1, wb);
  for (int32_t i = 0; i < n; ++i) level[i] = -1;
  level[src] = 0;
  setw(f, bs, src / bs, 1ULL << (bs - 1 - src % bs));
  if (sizes && max_levels > 0) sizes[0] = 1;
  int32_t lvl = 0;
  for (;;) {
    for (int32_t I = 0; I < nbr; ++I) setw(vis, bs, I, getw(vis, bs, I) | getw(f, bs, I)); /* OR (bfs.cu:511) */
    const int64_t cnt = or_bsrbmv_bfs_step(bs, n, brp, bci, tiles, f, vis, nx);
    if (cnt == 0) break;
    ++lvl;
    if (sizes && lvl < max_levels) sizes[lvl] = cnt;
    for (int32_t I = 0; I < nbr; ++I) {
      const uint64_t w = getw(nx, bs, I);
      for (int k = 0; k < bs; ++k)
        if ((w >> (bs - 1 - k)) & 1) level[(int64_t)I * bs + k] = lvl;
    }
    void* tmp = f;
    f = nx;
    nx = tmp;
  }
  free(f);
  free(vis);
  free(nx);
  return lvl + 1;
}

/* ------------------------------------------------------------- textbook -- */
int32_t or_bfs_queue(int32_t n, const int32_t* rp, const int32_t* ci, int32_t src, int32_t* level) {
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < n; ++i) level[i] = -1;
  int64_t h = 0, t = 0;
  level[src] = 0;
  q[t++] = src;
  int32_t maxl = 0;
  while (h < t) {
    const int32_t u = q[h++];
    for (int32_t k = rp[u]; k < rp[u + 1]; ++k) {
      const int32_t v = ci[k];
      if (level[v] < 0) {
        level[v] = level[u] + 1;
        if (level[v] > maxl) maxl = level[v];
        q[t++] = v;
      }
    }
  }
  free(q);
  return maxl + 1;
}

typedef struct {
  uint64_t key; /* dist << 32 | vertex */
} hitem;

static void heap_push(hitem* h, int64_t* n, uint64_t key) {
  int64_t i = (*n)++;
  h[i].key = key;
  while (i > 0) {
    const int64_t p = (i - 1) / 2;
    if (h[p].key <= h[i].key) break;
    hitem tmp = h[p];
    h[p] = h[i];
    h[i] = tmp;
    i = p;
  }
}
static uint64_t heap_pop(hitem* h, int64_t* n) {
  const uint64_t top = h[0].key;
  h[0] = h[--(*n)];
  int64_t i = 0;
  for (;;) {
    const int64_t l = 2 * i + 1, r = l + 1;
    int64_t m = i;
    if (l < *n && h[l].key < h[m].key) m = l;
    if (r < *n && h[r].key < h[m].key) m = r;
    if (m == i) break;
    hitem tmp = h[m];
    h[m] = h[i];
    h[i] = tmp;
    i = m;
  }
  return top;
}

void or_dijkstra(int32_t n, const int32_t* rp, const int32_t* ci, const uint32_t* w, int32_t src, uint32_t* dist) {
  for (int32_t i = 0; i < n; ++i) dist[i] = 0xFFFFFFFFu;
  hitem* h = (hitem*)malloc(sizeof(hitem) * (size_t)(rp[n] + 2));
  int64_t hn = 0;
  dist[src] = 0;
  heap_push(h, &hn, (uint64_t)src);
  while (hn > 0) {
    const uint64_t key = heap_pop(h, &hn);
    const uint32_t d = (uint32_t)(key >> 32);
    const int32_t u = (int32_t)(key & 0xFFFFFFFFu);
    if (d != dist[u]) continue;
    for (int32_t k = rp[u]; k < rp[u + 1]; ++k) {
      const int32_t v = ci[k];
      const uint64_t nd = (uint64_t)d + (w ? w[k] : 1u);
      if (nd < dist[v]) {
        dist[v] = (uint32_t)nd;
        heap_push(h, &hn, (nd << 32) | (uint32_t)v);
      }
    }
  }
  free(h);
}

int32_t or_pagerank(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* outdeg, double alpha, double tol,
                    int32_t max_iter, double* p) {
  double* x = (double*)malloc(sizeof(double) * (size_t)n);
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  for (int32_t i = 0; i < n; ++i) p[i] = 1.0 / n;
  int32_t it = 0;
  while (it < max_iter) {
    for (int32_t j = 0; j < n; ++j) x[j] = outdeg[j] > 0 ? alpha * p[j] / outdeg[j] : 0.0;
    double delta = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : delta)
    for (int32_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += x[ci[k]];
      y[i] = s + (1.0 - alpha) / n;
      delta += fabs(y[i] - p[i]);
    }
    memcpy(p, y, sizeof(double) * (size_t)n);
    ++it;
    if (delta < tol) break;
  }
  free(x);
  free(y);
  return it;
}

static int32_t uf_find(int32_t* par, int32_t u) {
  while (par[u] != u) {
    par[u] = par[par[u]];
    u = par[u];
  }
  return u;
}

void or_cc(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* label) {
  for (int32_t i = 0; i < n; ++i) label[i] = i;
  for (int32_t u = 0; u < n; ++u)
    for (int32_t k = rp[u]; k < rp[u + 1]; ++k) {
      const int32_t a = uf_find(label, u), b = uf_find(label, ci[k]);
      if (a < b) label[b] = a;
      else if (b < a) label[a] = b;
    }
  for (int32_t u = 0; u < n; ++u) label[u] = uf_find(label, u);
}

uint64_t or_tc(int32_t n, const int32_t* rp, const int32_t* ci) {
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (int32_t i = 0; i < n; ++i) {
    int32_t ie = rp[i];
    while (ie < rp[i + 1] && ci[ie] < i) ++ie; /* N_L(i) = [rp[i], ie) */
    for (int32_t q = rp[i]; q < ie; ++q) {
      const int32_t k = ci[q];
      int32_t a = rp[i], b = rp[k];
      while (a < ie && b < rp[k + 1] && ci[b] < k) {
        if (ci[a] < ci[b]) ++a;
        else if (ci[a] > ci[b]) ++b;
        else {
          ++total;
          ++a;
          ++b;
        }
      }
    }
  }
  return total;
}

/* ----------------------------------------------- personalized PageRank -- */
/* The map of algorithm/pr.cu:449-516 (p' = alpha * sum p[j]/outdeg[j] +
 * teleport) with a per-column personalization vector, and its adjoint; see
 * oracle.h.  y is the product term for row i and column k. */
static inline double ppr_y(int mode, const int32_t* rp, const int32_t* ci, const int32_t* outdeg, int32_t i,
                           int32_t K, int32_t k, const double* Xd, const float* Xf) {
  double y = 0.0;
  for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
    const int64_t j = ci[q];
    const double x = Xd ? Xd[j * K + k] : (double)Xf[j * K + k];
    if (mode == 0) {
      if (outdeg[j] > 0) y += x / (double)outdeg[j];
    } else {
      y += x;
    }
  }
  if (mode == 1) y = outdeg[i] > 0 ? y / (double)outdeg[i] : 0.0;
  return y;
}

int32_t or_ppr_solve(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* outdeg, double alpha, int32_t K,
                     int mode, const double* R, double* X, double tol, int32_t max_iter) {
  const double c = mode == 0 ? 1.0 - alpha : 1.0;
  double* Y = (double*)malloc(sizeof(double) * (size_t)n * K + 8);
  double* rn = (double*)calloc((size_t)K, sizeof(double));
  double* d = (double*)calloc((size_t)K, sizeof(double));
  for (int64_t i = 0; i < n; ++i)
    for (int32_t k = 0; k < K; ++k) rn[k] += fabs(R[i * K + k]);
  int32_t it = 0, done = 0;
  while (!done && it < max_iter) {
    for (int32_t k = 0; k < K; ++k) d[k] = 0.0;
#pragma omp parallel
    {
      double* dl = (double*)calloc((size_t)K, sizeof(double));
#pragma omp for schedule(dynamic, 1024)
      for (int32_t i = 0; i < n; ++i)
        for (int32_t k = 0; k < K; ++k) {
          const double t = alpha * ppr_y(mode, rp, ci, outdeg, i, K, k, X, NULL) + c * R[(int64_t)i * K + k];
          dl[k] += fabs(t - X[(int64_t)i * K + k]);
          Y[(int64_t)i * K + k] = t;
        }
#pragma omp critical
      for (int32_t k = 0; k < K; ++k) d[k] += dl[k];
      free(dl);
    }
    memcpy(X, Y, sizeof(double) * (size_t)n * K);
    ++it;
    double m = 0.0;
    for (int32_t k = 0; k < K; ++k) {
      const double rel = rn[k] > 0.0 ? d[k] / rn[k] : d[k];
      if (rel > m) m = rel;
    }
    done = m < tol;
  }
  free(Y);
  free(rn);
  free(d);
  return done ? it : -1;
}

void or_ppr_check(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* outdeg, double alpha, int32_t K,
                  int mode, const float* X, const float* R, const float* L, double* res_l1, double* res_inf,
                  double* dot, double* y_l1) {
  const double c = mode == 0 ? 1.0 - alpha : 1.0;
  double* acc = (double*)calloc((size_t)4 * K, sizeof(double)); /* l1, inf, dot, y_l1 */
#pragma omp parallel
  {
    double* a = (double*)calloc((size_t)4 * K, sizeof(double));
#pragma omp for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i)
      for (int32_t k = 0; k < K; ++k) {
        const double y = ppr_y(mode, rp, ci, outdeg, i, K, k, NULL, X);
        const double r = (double)R[(int64_t)i * K + k];
        const double e = fabs((double)X[(int64_t)i * K + k] - (alpha * y + c * r));
        a[k] += e;
        if (e > a[K + k]) a[K + k] = e;
        if (L) {
          a[2 * K + k] += (double)L[(int64_t)i * K + k] * (y - r);
          a[3 * K + k] += fabs(y - r);
        }
      }
#pragma omp critical
    for (int32_t k = 0; k < K; ++k) {
      acc[k] += a[k];
      if (a[K + k] > acc[K + k]) acc[K + k] = a[K + k];
      acc[2 * K + k] += a[2 * K + k];
      acc[3 * K + k] += a[3 * K + k];
    }
    free(a);
  }
  for (int32_t k = 0; k < K; ++k) {
    if (res_l1) res_l1[k] = acc[k];
    if (res_inf) res_inf[k] = acc[K + k];
    if (dot) dot[k] = acc[2 * K + k];
    if (y_l1) y_l1[k] = acc[3 * K + k];
  }
  free(acc);
}
