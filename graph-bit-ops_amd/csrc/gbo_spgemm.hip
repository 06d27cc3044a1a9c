// General bit-BSR SpGEMM C = A * B with a bit-BSR output (the product that
// bmm32_sparse_symbolic / bmm32_sparse_numeric, bsrbmm/bsrbmm.cu:1273,1311,
// set out to compute serially, one block row per thread block and a binary
// search per candidate tile).
//
// Operands are row-packed (word r of a tile = tile row r, MSB = column 0).
// A tile pair (t, u) with t = A(I,K) and u = B(K,J) is *productive* when
// some column k of A(I,K) is set in a row and row k of B(K,J) is nonzero:
// colmask(t) & rowmask(u) != 0.  The Boolean product of a productive pair is
// nonzero, an OR of nonzero words stays nonzero, and an unproductive pair
// adds nothing, so C has a tile (I,J) exactly when some productive pair
// lands there -- the tiles csr2bsr would make of the product, decided by
// one AND of two bs-bit masks per pair, without multiplying tiles.
//
// Expand-sort-compress on the device:
//   1. masks: colmask of every A tile, rowmask of every B tile (one group of
//      bs lanes per tile; OR by shuffles, rowmask by a ballot);
//   2. candidates: A tile t pairs with every tile of B's block row col(t);
//      their exclusive scan numbers the candidate pairs, and waves take
//      1024 consecutive candidates each (balanced whatever the skew; a lane
//      finds its A tile by a binary search bracketed by its wave's first and
//      last tile);
//   3. the productive pairs are compacted (count per wave, scan, write),
//      keyed (I, J);
//   4. a radix sort by key groups each C tile's pairs; the first pair of a
//      key run starts a C tile, so C's block columns, rowptr and segment
//      starts come from one scan of the run heads;
//   5. numeric: one group of bs lanes per C tile, lane r ORs, for every pair
//      of the tile's segment and every set bit k of A(I,K) row r, row k of
//      B(K,J); counts: for every set bit j of C row r, the sum over the
//      segment of popc(A(I,K) row r & column j of B(K,J)), columns of B
//      coming from a transposed copy of B's tiles.
//
// The symbolic phase runs steps 1-4 on the keys alone; the numeric phase
// and the count phase redo the expansion with the pair indices attached.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {
namespace {

constexpr int kPerWave = 1024;  // candidate pairs per wave

// colmask[t] = OR of A tile t's words; rowmask[u] = bit BS-1-r set iff row r
// of B tile u is nonzero.  Group of BS lanes per tile.
template <int BS>
__global__ __launch_bounds__(256) void k_masks(const word_t<BS>* __restrict__ tiles, int64_t T, bool rows,
                                               word_t<BS>* __restrict__ mask) {
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int64_t t = (int64_t)wave_index() * G + lane / BS;
  const int r = lane % BS;
  const uint64_t w = t < T ? (uint64_t)tiles[t * BS + r] : 0ull;
  uint64_t m;
  if (rows) {
    const uint64_t b = __ballot(w != 0);
    m = brev_low<BS>((b >> ((lane / BS) * BS)) & low_mask<BS>());
  } else {
    m = w;
    for (int off = 1; off < BS; off <<= 1) m |= __shfl_xor(m, off);
  }
  if (r == 0 && t < T) mask[t] = (word_t<BS>)m;
}

// Block row of every tile: the first I with rowptr[I + 1] > t (a binary
// search per tile, so a hub row is not one thread's loop).
__global__ __launch_bounds__(256) void k_tile_rows_of(const int32_t* __restrict__ rowptr, int32_t nbr, int64_t T,
                                                      int32_t* __restrict__ trow) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  int32_t lo = 0, hi = nbr - 1;
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if ((int64_t)rowptr[mid + 1] <= t) lo = mid + 1;
    else hi = mid;
  }
  trow[t] = lo;
}

// Candidate pairs of A tile t: the tiles of B's block row col(t).
__global__ __launch_bounds__(256) void k_cand_count(const int32_t* __restrict__ a_colind, int64_t T,
                                                    const int32_t* __restrict__ b_rowptr, int64_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int32_t K = a_colind[t];
  cnt[t] = b_rowptr[K + 1] - b_rowptr[K];
}

// Last index t in [lo, hi] with coff[t] <= p (coff nondecreasing, coff[lo] <= p).
__device__ __forceinline__ int64_t last_le(const int64_t* __restrict__ coff, int64_t lo, int64_t hi, int64_t p) {
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo + 1) / 2;
    if (coff[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct PairArgs {
  const int32_t* a_colind;
  const int32_t* a_trow;   // block row of each A tile
  const void* a_mask;      // colmask per A tile
  const int32_t* b_rowptr;
  const int32_t* b_colind;
  const void* b_mask;      // rowmask per B tile
  const int64_t* coff;     // exclusive scan of candidate counts, T + 1 entries
  int64_t T, P;            // A tiles, candidate pairs
  int jbits;               // key = (I << jbits) | J
  int64_t* wcount;         // COUNT: productive pairs per wave
  const int64_t* woff;     // EMIT: exclusive scan of wcount
  uint64_t* keys;          // EMIT
  uint64_t* vals;          // EMIT (nullable): (t << 32) | u
};

// MODE 0 counts the productive pairs of each wave's candidates, MODE 1
// writes their keys (and pair indices) compacted in candidate order.
template <int BS, int MODE>
__global__ __launch_bounds__(256) void k_pairs(const PairArgs a) {
  using W = word_t<BS>;
  const int lane = lane_id();
  const int64_t w = wave_index();
  const int64_t p0 = w * kPerWave;
  if (p0 >= a.P) return;
  const int64_t p1 = min(p0 + (int64_t)kPerWave, a.P);
  // the wave's A tiles lie in [t_first, t_last]
  const int64_t t_first = last_le(a.coff, 0, a.T - 1, p0);
  const int64_t t_last = last_le(a.coff, t_first, a.T - 1, p1 - 1);
  const W* am = (const W*)a.a_mask;
  const W* bm = (const W*)a.b_mask;
  int64_t pos = MODE == 1 ? a.woff[w] : 0;
  int64_t total = 0;
  for (int64_t base = p0; base < p1; base += kWave) {
    const int64_t p = base + lane;
    bool prod = false;
    int64_t t = 0, u = 0;
    if (p < p1) {
      t = last_le(a.coff, t_first, t_last, p);
      const int32_t K = a.a_colind[t];
      u = a.b_rowptr[K] + (p - a.coff[t]);
      prod = ((uint64_t)am[t] & (uint64_t)bm[u]) != 0;
    }
    const uint64_t b = __ballot(prod);
    if (MODE == 0) {
      total += __popcll(b);
    } else {
      if (prod) {
        const int64_t at = pos + __popcll(b & ((1ull << lane) - 1ull));
        a.keys[at] = ((uint64_t)(uint32_t)a.a_trow[t] << a.jbits) | (uint64_t)(uint32_t)a.b_colind[u];
        if (a.vals) a.vals[at] = ((uint64_t)t << 32) | (uint64_t)u;
      }
      pos += __popcll(b);
    }
  }
  if (MODE == 0 && lane == 0) a.wcount[w] = total;
}

// head[i] = 1 where a run of equal keys starts.
__global__ __launch_bounds__(256) void k_heads(const uint64_t* __restrict__ keys, int64_t P, int32_t* __restrict__ head) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// For each run head i (cid = its exclusive rank): the C tile's block row and
// column, and its segment start.
__global__ __launch_bounds__(256) void k_runs(const uint64_t* __restrict__ keys, const int32_t* __restrict__ head,
                                              const int64_t* __restrict__ rank, int64_t P, int jbits,
                                              int32_t* __restrict__ crow, int32_t* __restrict__ ccol,
                                              int64_t* __restrict__ seg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P || !head[i]) return;
  const int64_t c = rank[i];
  crow[c] = (int32_t)(keys[i] >> jbits);
  if (ccol) ccol[c] = (int32_t)(keys[i] & ((1ull << jbits) - 1ull));
  if (seg) seg[c] = i;
}

// rowptr[I] = first C tile whose block row is >= I, for I in [0, nbr].
__global__ __launch_bounds__(256) void k_row_bounds(const int32_t* __restrict__ crow, int64_t nc, int32_t nbr,
                                                    int32_t* __restrict__ rowptr) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I > nbr) return;
  int64_t lo = 0, hi = nc;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (crow[mid] < (int32_t)I) lo = mid + 1;
    else hi = mid;
  }
  rowptr[I] = (int32_t)lo;
}

// Boolean product of each C tile from its segment of pairs.  Group of BS
// lanes per C tile, lane r = tile row r.
template <int BS>
__global__ __launch_bounds__(256) void k_numeric(const uint64_t* __restrict__ vals, const int64_t* __restrict__ seg,
                                                 int64_t nc, const word_t<BS>* __restrict__ at,
                                                 const word_t<BS>* __restrict__ bt, word_t<BS>* __restrict__ ct) {
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int64_t c = (int64_t)wave_index() * G + lane / BS;
  const int r = lane % BS;
  if (c >= nc) return;
  uint64_t acc = 0;
  for (int64_t i = seg[c]; i < seg[c + 1]; ++i) {
    const uint64_t v = vals[i];
    const int64_t t = (int64_t)(v >> 32), u = (int64_t)(v & 0xFFFFFFFFull);
    uint64_t a = (uint64_t)at[t * BS + r];
    while (a) {
      const int p = 63 - __builtin_clzll(a);  // highest set bit: column BS-1-p
      a &= ~(1ull << p);
      acc |= (uint64_t)bt[u * BS + (BS - 1 - p)];
    }
  }
  ct[c * BS + r] = (word_t<BS>)acc;
}

template <int BS>
__global__ __launch_bounds__(256) void k_tile_popc64(const word_t<BS>* __restrict__ tiles, int64_t T,
                                                     int64_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  int64_t s = 0;
  for (int r = 0; r < BS; ++r) s += popcw(tiles[t * BS + r]);
  cnt[t] = s;
}

// Integer product at the set bits of C: lane r walks the set bits j of C
// row r (ascending column) and sums popc(A(I,K) row r & column j of B(K,J))
// over the segment; values go out in tile bit order (rows in order, columns
// ascending inside a row) from valoff[c].
template <int BS>
__global__ __launch_bounds__(256) void k_counts(const uint64_t* __restrict__ vals, const int64_t* __restrict__ seg,
                                                int64_t nc, const word_t<BS>* __restrict__ at,
                                                const word_t<BS>* __restrict__ btc,  // B's tiles, column form
                                                const word_t<BS>* __restrict__ ct, const int64_t* __restrict__ valoff,
                                                int32_t* __restrict__ out) {
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int64_t c = (int64_t)wave_index() * G + lane / BS;
  const int r = lane % BS;
  const uint64_t crow = c < nc ? (uint64_t)ct[c * BS + r] : 0ull;
  // bits of the rows before r inside this tile (group prefix sum)
  int before = popcw(crow);
  for (int off = 1; off < BS; off <<= 1) {
    const int o = __shfl_up(before, off, BS);
    if (r >= off) before += o;
  }
  before -= popcw(crow);
  if (c >= nc) return;
  int64_t pos = valoff[c] + before;
  uint64_t m = crow;
  while (m) {
    const int p = 63 - __builtin_clzll(m);
    m &= ~(1ull << p);
    const int j = BS - 1 - p;
    int32_t s = 0;
    for (int64_t i = seg[c]; i < seg[c + 1]; ++i) {
      const uint64_t v = vals[i];
      const int64_t t = (int64_t)(v >> 32), u = (int64_t)(v & 0xFFFFFFFFull);
      s += popcw((word_t<BS>)((uint64_t)at[t * BS + r] & (uint64_t)btc[u * BS + j]));
    }
    out[pos++] = s;
  }
}

// Column-form copy of row-form tiles: one wave takes 64/BS tiles, lane l
// holds word l % BS of tile l / BS, one ballot per column k gathers that
// column over the rows.
template <int BS>
__global__ __launch_bounds__(256) void k_transpose_tiles(const word_t<BS>* __restrict__ in, int64_t T,
                                                         word_t<BS>* __restrict__ out) {
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int64_t t = (int64_t)wave_index() * G + lane / BS;
  const int r = lane % BS;
  const uint64_t w = t < T ? (uint64_t)in[t * BS + r] : 0ull;
  const int g = lane / BS;
  for (int k = 0; k < BS; ++k) {
    const uint64_t b = __ballot((w >> (BS - 1 - k)) & 1ull);  // bit l: row l % BS of tile l / BS has column k
    if (r == k && t < T) out[t * BS + k] = (word_t<BS>)brev_low<BS>((b >> (g * BS)) & low_mask<BS>());
  }
}

unsigned bits_for(int64_t n) {  // smallest b with 2^b > n (at least 1)
  unsigned b = 1;
  while (b < 63 && (1ll << b) <= n) ++b;
  return b;
}

unsigned grid_for(int64_t items, int per_block) {
  return (unsigned)std::max<int64_t>(1, (items + per_block - 1) / per_block);
}

int check_operands(const gbo_bsr* A, const gbo_bsr* B, const char* who) {
  int rc = validate_bsr(A, 0, who);
  if (rc) return rc;
  if ((rc = validate_bsr(B, 0, who))) return rc;
  if (A->bs != B->bs) return set_error(GBO_EINVAL, "%s: block sizes differ (%d, %d)", who, A->bs, B->bs);
  if (A->ncols != B->nrows) return set_error(GBO_EINVAL, "%s: inner dimensions differ (%d, %d)", who, A->ncols, B->nrows);
  if (A->nblocks >= ((int64_t)1 << 32) || B->nblocks >= ((int64_t)1 << 32))
    return set_error(GBO_ERANGE, "%s: too many tiles", who);
  return GBO_OK;
}

// The sorted productive pairs of A * B: keys ((I << jbits) | J) and, when
// `with_vals`, the pair indices ((t << 32) | u), in key order.  Buffers are
// owned by `mem`.
struct Pairs {
  uint64_t* keys = nullptr;
  uint64_t* vals = nullptr;
  int64_t P = 0;
  int jbits = 1;
};

template <int BS>
int sorted_pairs(const gbo_bsr* A, const gbo_bsr* B, bool with_vals, AsyncAllocs& mem, Pairs* out, hipStream_t st) {
  using W = word_t<BS>;
  const int64_t T = A->nblocks;
  out->jbits = (int)bits_for(std::max<int32_t>(B->nblockcols, 1) - 1);
  if (T == 0 || B->nblocks == 0) return GBO_OK;
  AsyncAllocs tmp(st);
  W *amask = nullptr, *bmask = nullptr;
  int32_t* trow = nullptr;
  int64_t *cnt = nullptr, *coff = nullptr;
  int rc;
  if ((rc = tmp.alloc(&amask, sizeof(W) * (size_t)T, "gbo_bsrbmm")) ||
      (rc = tmp.alloc(&bmask, sizeof(W) * (size_t)B->nblocks, "gbo_bsrbmm")) ||
      (rc = tmp.alloc(&trow, sizeof(int32_t) * (size_t)T, "gbo_bsrbmm")) ||
      (rc = tmp.alloc(&cnt, sizeof(int64_t) * (size_t)T, "gbo_bsrbmm")) ||
      (rc = tmp.alloc(&coff, sizeof(int64_t) * ((size_t)T + 1), "gbo_bsrbmm")))
    return rc;
  constexpr int G = kWave / BS;
  k_masks<BS><<<grid_for((T + G - 1) / G, 4), 256, 0, st>>>((const W*)A->tiles, T, false, amask);
  k_masks<BS><<<grid_for((B->nblocks + G - 1) / G, 4), 256, 0, st>>>((const W*)B->tiles, B->nblocks, true, bmask);
  k_tile_rows_of<<<grid_for(T, 256), 256, 0, st>>>(A->rowptr, A->nblockrows, T, trow);
  k_cand_count<<<grid_for(T, 256), 256, 0, st>>>(A->colind, T, B->rowptr, cnt);
  GBO_LAUNCH_CHECK();
  if ((rc = scan_exclusive<int64_t, int64_t>(cnt, coff, T, st))) return rc;
  int64_t Pc = 0;
  GBO_HIP_CHECK(hipMemcpyAsync(&Pc, coff + T, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if (Pc == 0) return GBO_OK;
  const int64_t nw = (Pc + kPerWave - 1) / kPerWave;
  int64_t *wcount = nullptr, *woff = nullptr;
  if ((rc = tmp.alloc(&wcount, sizeof(int64_t) * (size_t)nw, "gbo_bsrbmm")) ||
      (rc = tmp.alloc(&woff, sizeof(int64_t) * ((size_t)nw + 1), "gbo_bsrbmm")))
    return rc;
  PairArgs pa{};
  pa.a_colind = A->colind;
  pa.a_trow = trow;
  pa.a_mask = amask;
  pa.b_rowptr = B->rowptr;
  pa.b_colind = B->colind;
  pa.b_mask = bmask;
  pa.coff = coff;
  pa.T = T;
  pa.P = Pc;
  pa.jbits = out->jbits;
  pa.wcount = wcount;
  k_pairs<BS, 0><<<grid_for(nw, 4), 256, 0, st>>>(pa);
  GBO_LAUNCH_CHECK();
  if ((rc = scan_exclusive<int64_t, int64_t>(wcount, woff, nw, st))) return rc;
  int64_t P = 0;
  GBO_HIP_CHECK(hipMemcpyAsync(&P, woff + nw, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if (P == 0) return GBO_OK;
  uint64_t *k0 = nullptr, *v0 = nullptr;
  if ((rc = tmp.alloc(&k0, sizeof(uint64_t) * (size_t)P, "gbo_bsrbmm")) ||
      (rc = mem.alloc(&out->keys, sizeof(uint64_t) * (size_t)P, "gbo_bsrbmm")))
    return rc;
  if (with_vals && ((rc = tmp.alloc(&v0, sizeof(uint64_t) * (size_t)P, "gbo_bsrbmm")) ||
                    (rc = mem.alloc(&out->vals, sizeof(uint64_t) * (size_t)P, "gbo_bsrbmm"))))
    return rc;
  pa.woff = woff;
  pa.keys = k0;
  pa.vals = v0;
  k_pairs<BS, 1><<<grid_for(nw, 4), 256, 0, st>>>(pa);
  GBO_LAUNCH_CHECK();
  const unsigned end_bit = (unsigned)out->jbits + bits_for(std::max<int32_t>(A->nblockrows, 1) - 1);
  size_t tb = 0;
  void* sort_tmp = nullptr;
  if (with_vals) {
    GBO_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, k0, out->keys, v0, out->vals, (size_t)P, 0, end_bit, st));
  } else {
    GBO_HIP_CHECK(rocprim::radix_sort_keys(nullptr, tb, k0, out->keys, (size_t)P, 0, end_bit, st));
  }
  if ((rc = tmp.alloc(&sort_tmp, std::max<size_t>(tb, 8), "gbo_bsrbmm"))) return rc;
  if (with_vals) {
    GBO_HIP_CHECK(rocprim::radix_sort_pairs(sort_tmp, tb, k0, out->keys, v0, out->vals, (size_t)P, 0, end_bit, st));
  } else {
    GBO_HIP_CHECK(rocprim::radix_sort_keys(sort_tmp, tb, k0, out->keys, (size_t)P, 0, end_bit, st));
  }
  out->P = P;
  return GBO_OK;
}

// Runs of the sorted keys: C's block rows (and columns, segment starts) per
// C tile; *nc = number of C tiles.
int key_runs(const Pairs& pr, AsyncAllocs& mem, int32_t** crow, int32_t* ccol, int64_t** seg, int64_t* nc,
             hipStream_t st) {
  AsyncAllocs tmp(st);
  int32_t* head = nullptr;
  int64_t* rank = nullptr;
  int rc;
  if ((rc = tmp.alloc(&head, sizeof(int32_t) * (size_t)pr.P, "gbo_bsrbmm")) ||
      (rc = tmp.alloc(&rank, sizeof(int64_t) * ((size_t)pr.P + 1), "gbo_bsrbmm")))
    return rc;
  k_heads<<<grid_for(pr.P, 256), 256, 0, st>>>(pr.keys, pr.P, head);
  GBO_LAUNCH_CHECK();
  if ((rc = scan_exclusive<int32_t, int64_t>(head, rank, pr.P, st))) return rc;
  GBO_HIP_CHECK(hipMemcpyAsync(nc, rank + pr.P, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if ((rc = mem.alloc(crow, sizeof(int32_t) * (size_t)std::max<int64_t>(*nc, 1), "gbo_bsrbmm"))) return rc;
  if (seg) {
    if ((rc = mem.alloc(seg, sizeof(int64_t) * ((size_t)*nc + 1), "gbo_bsrbmm"))) return rc;
    GBO_HIP_CHECK(hipMemcpyAsync(*seg + *nc, &pr.P, sizeof(int64_t), hipMemcpyHostToDevice, st));
  }
  k_runs<<<grid_for(pr.P, 256), 256, 0, st>>>(pr.keys, head, rank, pr.P, pr.jbits, *crow, ccol, seg ? *seg : nullptr);
  GBO_LAUNCH_CHECK();
  // the host value copied above must outlive the copy
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  return GBO_OK;
}

template <int BS>
int symbolic_t(const gbo_bsr* A, const gbo_bsr* B, int32_t* d_rowptr, int64_t* nblocks, hipStream_t st) {
  AsyncAllocs mem(st);
  Pairs pr;
  int rc = sorted_pairs<BS>(A, B, false, mem, &pr, st);
  if (rc) return rc;
  int64_t nc = 0;
  int32_t* crow = nullptr;
  if (pr.P > 0 && (rc = key_runs(pr, mem, &crow, nullptr, nullptr, &nc, st))) return rc;
  if (nc >= ((int64_t)1 << 31)) return set_error(GBO_ERANGE, "gbo_bsrbmm_symbolic: %lld tiles exceed int32", (long long)nc);
  k_row_bounds<<<grid_for((int64_t)A->nblockrows + 1, 256), 256, 0, st>>>(crow, nc, A->nblockrows, d_rowptr);
  GBO_LAUNCH_CHECK();
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  *nblocks = nc;
  return GBO_OK;
}

// Numeric phase; with `counts` also the integer product at C's set bits
// (valoff must then hold C's per-tile bit offsets).
template <int BS>
int numeric_t(const gbo_bsr* A, const gbo_bsr* B, int64_t nc_expect, int32_t* d_colind, word_t<BS>* d_tiles,
              const int64_t* d_valoff, int32_t* d_counts, hipStream_t st) {
  using W = word_t<BS>;
  AsyncAllocs mem(st);
  Pairs pr;
  int rc = sorted_pairs<BS>(A, B, true, mem, &pr, st);
  if (rc) return rc;
  int64_t nc = 0;
  int32_t* crow = nullptr;
  int64_t* seg = nullptr;
  if (pr.P > 0 && (rc = key_runs(pr, mem, &crow, d_counts ? nullptr : d_colind, &seg, &nc, st))) return rc;
  if (nc != nc_expect)
    return set_error(GBO_EINVAL, "gbo_bsrbmm: C has %lld tiles, the given structure %lld", (long long)nc,
                     (long long)nc_expect);
  if (nc == 0) return GBO_OK;
  constexpr int G = kWave / BS;
  if (!d_counts) {
    k_numeric<BS><<<grid_for((nc + G - 1) / G, 4), 256, 0, st>>>(pr.vals, seg, nc, (const W*)A->tiles,
                                                                 (const W*)B->tiles, d_tiles);
    GBO_LAUNCH_CHECK();
  } else {
    W* btc = nullptr;
    if ((rc = mem.alloc(&btc, sizeof(W) * (size_t)B->nblocks * BS, "gbo_bsrbmm_values"))) return rc;
    k_transpose_tiles<BS><<<grid_for((B->nblocks + G - 1) / G, 4), 256, 0, st>>>((const W*)B->tiles, B->nblocks, btc);
    GBO_LAUNCH_CHECK();
    k_counts<BS><<<grid_for((nc + G - 1) / G, 4), 256, 0, st>>>(pr.vals, seg, nc, (const W*)A->tiles, btc,
                                                                (const W*)d_tiles, d_valoff, d_counts);
    GBO_LAUNCH_CHECK();
  }
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  return GBO_OK;
}

template <int BS>
int valoff_t(const word_t<BS>* tiles, int64_t nc, int64_t* d_valoff, int64_t* nnz, hipStream_t st) {
  AsyncAllocs mem(st);
  int64_t* cnt = nullptr;
  int rc = mem.alloc(&cnt, sizeof(int64_t) * (size_t)std::max<int64_t>(nc, 1), "gbo_bsrbmm_numeric");
  if (rc) return rc;
  if (nc > 0) {
    k_tile_popc64<BS><<<grid_for(nc, 256), 256, 0, st>>>(tiles, nc, cnt);
    GBO_LAUNCH_CHECK();
  }
  if ((rc = scan_exclusive<int64_t, int64_t>(cnt, d_valoff, nc, st))) return rc;
  if (nnz) {
    GBO_HIP_CHECK(hipMemcpyAsync(nnz, d_valoff + nc, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    GBO_HIP_CHECK(hipStreamSynchronize(st));
  }
  return GBO_OK;
}

}  // namespace
}  // namespace gbo

using namespace gbo;

#define GBO_BS_DISPATCH(bs, call8, call16, call32, call64) \
  ((bs) == 8 ? (call8) : (bs) == 16 ? (call16) : (bs) == 32 ? (call32) : (call64))

extern "C" {

int gbo_bsrbmm_symbolic(const gbo_bsr* A, const gbo_bsr* B, int32_t* d_C_rowptr, int64_t* C_nblocks,
                        gbo_stream_t stream) {
  GBO_TRY_BEGIN
  int rc = check_operands(A, B, "gbo_bsrbmm_symbolic");
  if (rc) return rc;
  if (!d_C_rowptr || !C_nblocks) return set_error(GBO_EINVAL, "gbo_bsrbmm_symbolic: null output");
  hipStream_t st = stream;
  return GBO_BS_DISPATCH(A->bs, symbolic_t<8>(A, B, d_C_rowptr, C_nblocks, st),
                         symbolic_t<16>(A, B, d_C_rowptr, C_nblocks, st), symbolic_t<32>(A, B, d_C_rowptr, C_nblocks, st),
                         symbolic_t<64>(A, B, d_C_rowptr, C_nblocks, st));
  GBO_TRY_END
}

int gbo_bsrbmm_numeric(const gbo_bsr* A, const gbo_bsr* B, const int32_t* d_C_rowptr, int64_t C_nblocks,
                       int32_t* d_C_colind, void* d_C_tiles, int64_t* d_C_valoff, int64_t* C_nnz,
                       gbo_stream_t stream) {
  GBO_TRY_BEGIN
  int rc = check_operands(A, B, "gbo_bsrbmm_numeric");
  if (rc) return rc;
  if (!d_C_rowptr || C_nblocks < 0 || (C_nblocks > 0 && (!d_C_colind || !d_C_tiles)))
    return set_error(GBO_EINVAL, "gbo_bsrbmm_numeric: bad output");
  hipStream_t st = stream;
  const int bs = A->bs;
  rc = GBO_BS_DISPATCH(
      bs, numeric_t<8>(A, B, C_nblocks, d_C_colind, (uint8_t*)d_C_tiles, nullptr, nullptr, st),
      numeric_t<16>(A, B, C_nblocks, d_C_colind, (uint16_t*)d_C_tiles, nullptr, nullptr, st),
      numeric_t<32>(A, B, C_nblocks, d_C_colind, (uint32_t*)d_C_tiles, nullptr, nullptr, st),
      numeric_t<64>(A, B, C_nblocks, d_C_colind, (uint64_t*)d_C_tiles, nullptr, nullptr, st));
  if (rc || !d_C_valoff) return rc;
  return GBO_BS_DISPATCH(bs, valoff_t<8>((const uint8_t*)d_C_tiles, C_nblocks, d_C_valoff, C_nnz, st),
                         valoff_t<16>((const uint16_t*)d_C_tiles, C_nblocks, d_C_valoff, C_nnz, st),
                         valoff_t<32>((const uint32_t*)d_C_tiles, C_nblocks, d_C_valoff, C_nnz, st),
                         valoff_t<64>((const uint64_t*)d_C_tiles, C_nblocks, d_C_valoff, C_nnz, st));
  GBO_TRY_END
}

int gbo_bsrbmm_values(const gbo_bsr* A, const gbo_bsr* B, const gbo_bsr* C, const int64_t* d_C_valoff,
                      int32_t* d_C_counts, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  int rc = check_operands(A, B, "gbo_bsrbmm_values");
  if (rc) return rc;
  if ((rc = validate_bsr(C, 0, "gbo_bsrbmm_values(C)"))) return rc;
  if (C->bs != A->bs || C->nrows != A->nrows || C->ncols != B->ncols)
    return set_error(GBO_EINVAL, "gbo_bsrbmm_values: C is not the %d x %d product at bs %d", A->nrows, B->ncols,
                     A->bs);
  if (C->nblocks > 0 && (!d_C_valoff || !d_C_counts)) return set_error(GBO_EINVAL, "gbo_bsrbmm_values: null output");
  hipStream_t st = stream;
  return GBO_BS_DISPATCH(
      A->bs, numeric_t<8>(A, B, C->nblocks, nullptr, (uint8_t*)C->tiles, d_C_valoff, d_C_counts, st),
      numeric_t<16>(A, B, C->nblocks, nullptr, (uint16_t*)C->tiles, d_C_valoff, d_C_counts, st),
      numeric_t<32>(A, B, C->nblocks, nullptr, (uint32_t*)C->tiles, d_C_valoff, d_C_counts, st),
      numeric_t<64>(A, B, C->nblocks, nullptr, (uint64_t*)C->tiles, d_C_valoff, d_C_counts, st));
  GBO_TRY_END
}

int gbo_bsr_bit_offsets(const gbo_bsr* A, int64_t* d_off, int64_t* nnz, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  int rc = validate_bsr(A, -1, "gbo_bsr_bit_offsets");
  if (rc) return rc;
  if (!d_off) return set_error(GBO_EINVAL, "gbo_bsr_bit_offsets: null output");
  hipStream_t st = stream;
  return GBO_BS_DISPATCH(A->bs, valoff_t<8>((const uint8_t*)A->tiles, A->nblocks, d_off, nnz, st),
                         valoff_t<16>((const uint16_t*)A->tiles, A->nblocks, d_off, nnz, st),
                         valoff_t<32>((const uint32_t*)A->tiles, A->nblocks, d_off, nnz, st),
                         valoff_t<64>((const uint64_t*)A->tiles, A->nblocks, d_off, nnz, st));
  GBO_TRY_END
}

int gbo_bsrbmm(const gbo_bsr* A, const gbo_bsr* B, gbo_bsr* C, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  if (!C) return set_error(GBO_EINVAL, "gbo_bsrbmm: null output");
  std::memset(C, 0, sizeof(*C));
  int rc = check_operands(A, B, "gbo_bsrbmm");
  if (rc) return rc;
  gbo_bsr R{};
  R.bs = A->bs;
  R.dir = 0;
  R.nrows = A->nrows;
  R.ncols = B->ncols;
  R.nblockrows = A->nblockrows;
  R.nblockcols = B->nblockcols;
  GBO_HIP_CHECK(hipMalloc((void**)&R.rowptr, sizeof(int32_t) * ((size_t)R.nblockrows + 1)));
  rc = gbo_bsrbmm_symbolic(A, B, R.rowptr, &R.nblocks, stream);
  if (rc == GBO_OK) {
    const size_t tile_bytes = (size_t)R.nblocks * R.bs * (R.bs / 8);
    hipError_t e1 = hipMalloc((void**)&R.colind, sizeof(int32_t) * std::max<int64_t>(R.nblocks, 1));
    hipError_t e2 = hipMalloc(&R.tiles, std::max<size_t>(tile_bytes, 8));
    if (e1 != hipSuccess || e2 != hipSuccess)
      rc = set_error(GBO_ENOMEM, "gbo_bsrbmm: cannot allocate %zu tile bytes", tile_bytes);
  }
  if (rc == GBO_OK)
    rc = gbo_bsrbmm_numeric(A, B, R.rowptr, R.nblocks, R.colind, R.tiles, nullptr, nullptr, stream);
  if (rc != GBO_OK) {
    gbo_bsr_free(&R);
    return rc;
  }
  R.plan_variant = (R.bs <= 16) ? 3 : 0;
  *C = R;
  return GBO_OK;
  GBO_TRY_END
}

}  // extern "C"
