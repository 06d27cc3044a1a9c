// CSR -> bit-BSR packing on the GPU (replaces csr2bsr_batch_{N}, algorithm/
// csr2bsr_batch_bsrbmv.cu:180-947, and csr2bsr_batch_{N}_{row,col},
// bsrbmm/csr2bsr_batch.cu), vector packing (ToBit{N}Row) and the work
// decomposition ("plan") used by every bsrbmv/bsrbmm operation.
//
// The reference packs one block row per host-loop iteration through dense
// float tiles (cusparseScsr2bsr) and a ballot kernel.  Here one wavefront
// packs 64/bs block rows at once: lane r of a group follows CSR row r of its
// block row, the group takes the minimum pending block column (a k-way merge
// of the bs sorted rows), and every lane ORs its entries of that block
// column into its own row word -- which is directly word r of the tile in
// the reference layout (MSB = column 0).  A block row of more than 2048
// nonzeros is split into pieces (block-column ranges) merged by groups of
// their own.  Two passes: count tiles per piece, scan, then fill
// colind/tiles.  No dense staging.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {

// ------------------------------------------------------------------ scan
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanThreads * kScanPer;

template <class TI>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const TI* __restrict__ in, int64_t n,
                                                              int64_t* __restrict__ bsum) {
  __shared__ int64_t sh[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t i = base + threadIdx.x + (int64_t)k * kScanThreads;
    if (i < n) s += (int64_t)in[i];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

// Exclusive scan of one tile; thread t owns elements [base+t*P, base+t*P+P).
// In-place safe: each element is read and written by the same thread.
template <class TI, class TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const TI* in, int64_t n,
                                                             const int64_t* __restrict__ boff,
                                                             TO* out) {
  __shared__ int64_t sh[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  int64_t v[kScanPer];
  int64_t s = 0;
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t i = base + k;
    v[k] = (i < n) ? (int64_t)in[i] : 0;
    s += v[k];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int64_t o = ((int)threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += o;
    __syncthreads();
  }
  int64_t run = sh[threadIdx.x] - s + (boff ? boff[blockIdx.x] : 0);
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t i = base + k;
    if (i < n) out[i] = (TO)run;
    run += v[k];
    if (i == n - 1) out[n] = (TO)run;
  }
}

// out[0..n] = exclusive prefix sums of in[0..n-1]; out[n] = total.
template <class TI, class TO>
int scan_exclusive(const TI* in, TO* out, int64_t n, hipStream_t st) {
  if (n <= 0) {
    GBO_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(TO), st));
    return GBO_OK;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    k_scan_apply<TI, TO><<<1, kScanThreads, 0, st>>>(in, n, nullptr, out);
    GBO_LAUNCH_CHECK();
    return GBO_OK;
  }
  int64_t* bsum = nullptr;
  GBO_HIP_CHECK(hipMallocAsync((void**)&bsum, sizeof(int64_t) * (2 * nb + 1), st));
  int64_t* boff = bsum + nb;
  k_scan_reduce<TI><<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, bsum);
  GBO_LAUNCH_CHECK();
  int rc = scan_exclusive<int64_t, int64_t>(bsum, boff, nb, st);
  if (rc == GBO_OK) {
    k_scan_apply<TI, TO><<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, boff, out);
    GBO_LAUNCH_CHECK();
  }
  GBO_HIP_CHECK(hipFreeAsync(bsum, st));
  return rc;
}

template int scan_exclusive<int32_t, int32_t>(const int32_t*, int32_t*, int64_t, hipStream_t);
template int scan_exclusive<int32_t, int64_t>(const int32_t*, int64_t*, int64_t, hipStream_t);
template int scan_exclusive<int64_t, int64_t>(const int64_t*, int64_t*, int64_t, hipStream_t);

// ---------------------------------------------------------------- packing
enum { PACK_COUNT = 0, PACK_ROW = 1, PACK_COL = 2, PACK_VALUES = 3 };

// Error flags raised by the packer (device int, OR-ed).
enum { PACKERR_COLRANGE = 1, PACKERR_UNSORTED = 2 };

// A block row with more than kPieceNnz nonzeros is packed in pieces: ranges
// of block columns [lo, hi), one group each, cut at evenly spaced entries of
// the block row's longest CSR row.  A single group's merge over the hub
// block row of R-MAT s22 (129,447 tiles at bs 8, one dependent step per
// tile) had set the time of each whole packing pass (~75 ms).
constexpr int32_t kPieceNnz = 2048;

// The pieces: piece p is block row row[p] over block columns [lo[p], hi[p])
// (lo = INT_MIN: from the row's start, hi = INT_MAX: to its end).  Without a
// table (row == nullptr) piece p is block row p, whole.
struct PackPieces {
  const int32_t* row = nullptr;
  const int32_t* lo = nullptr;
  const int32_t* hi = nullptr;
  const int64_t* base = nullptr;  // fill: first tile of each piece
  int32_t* cnt = nullptr;         // count: tiles of each piece
  int32_t n = 0;
};

// First index q in [q0, q1) with ci[q] / BS >= J (q1 if none).  On rows that
// are not sorted the result is still monotone in J, so the pieces of a row
// cover it exactly and the count pass sees every adjacent pair once.
template <int BS>
__device__ __forceinline__ int32_t first_blockcol_at_least(const int32_t* __restrict__ ci, int32_t q0, int32_t q1,
                                                           int32_t J) {
  while (q0 < q1) {
    const int32_t mid = q0 + (q1 - q0) / 2;
    if (ci[mid] / BS < J) q0 = mid + 1;
    else q1 = mid;
  }
  return q0;
}

template <int BS, int MODE>
__global__ __launch_bounds__(256) void k_pack(const int32_t* __restrict__ rp,
                                              const int32_t* __restrict__ ci, int32_t nrows,
                                              int32_t ncols, int32_t nbr,
                                              int32_t* __restrict__ bsr_rp,  // counts (COUNT) or rowptr
                                              int32_t* __restrict__ out_col,
                                              word_t<BS>* __restrict__ out_tiles,
                                              const uint32_t* __restrict__ val,
                                              const int64_t* __restrict__ woff,
                                              uint32_t* __restrict__ tval, int* __restrict__ err,
                                              const PackPieces pc) {
  using W = word_t<BS>;
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int g = lane / BS, r = lane % BS;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t npc = pc.row ? (int64_t)pc.n : (int64_t)nbr;
  const int64_t pi = wave * G + g;  // piece of this group
  if (wave * G >= npc) return;      // wave-uniform
  const bool blk_ok = pi < npc;
  const int64_t I = pc.row ? (blk_ok ? (int64_t)pc.row[pi] : 0) : pi;
  const int64_t row = I * BS + r;
  int32_t p = 0, pe = 0;
  if (blk_ok && row < nrows) {
    const int32_t r0 = rp[row], r1 = rp[row + 1];
    p = r0;
    pe = r1;
    if (pc.row) {
      const int32_t lo = pc.lo[pi], hi = pc.hi[pi];
      if (lo != INT_MIN) p = first_blockcol_at_least<BS>(ci, r0, r1, lo);
      if (hi != INT_MAX) pe = first_blockcol_at_least<BS>(ci, r0, r1, hi);
      if (pe < p) pe = p;  // only on rows that are not sorted (flagged below)
      // the pair across the piece's start (pairs inside it are checked by the merge)
      if (MODE == PACK_COUNT && p > r0 && p < r1 && ci[p - 1] / BS > ci[p] / BS) atomicOr(err, PACKERR_UNSORTED);
    }
  }
  auto blockcol = [&](int32_t q) -> int32_t {
    const int32_t c = ci[q];
    if (MODE == PACK_COUNT && (c < 0 || c >= ncols)) {
      atomicOr(err, PACKERR_COLRANGE);
      return INT_MAX - 1;  // poison: still terminates
    }
    return c / BS;
  };
  int32_t cur = (p < pe) ? blockcol(p) : INT_MAX;
  int64_t base = 0;
  if (MODE != PACK_COUNT && blk_ok) base = pc.base ? pc.base[pi] : bsr_rp[I];
  int32_t k = 0;
  while (true) {
    const int32_t m = group_min<BS>(cur);
    if (__all(m == INT_MAX)) break;
    uint64_t w = 0;
    int32_t p0 = p;
    if (m != INT_MAX) {
      while (cur == m) {
        w |= 1ull << (BS - 1 - (ci[p] % BS));
        ++p;
        const int32_t nxt = (p < pe) ? blockcol(p) : INT_MAX;
        if (MODE == PACK_COUNT && nxt < cur) atomicOr(err, PACKERR_UNSORTED);
        cur = nxt;
      }
    }
    if constexpr (MODE == PACK_ROW) {
      if (m != INT_MAX) {
        out_tiles[(base + k) * BS + r] = (W)w;
        if (r == 0) out_col[base + k] = m;
      }
    } else if constexpr (MODE == PACK_COL) {
      // Transpose the group's bs x bs bit tile with one ballot per column:
      // ballot bit (g*BS + r) = row r has column kk.
      uint64_t mine = 0;
      for (int kk = 0; kk < BS; ++kk) {
        const uint64_t b = __ballot((w >> (BS - 1 - kk)) & 1ull);
        const uint64_t grp = (b >> (g * BS)) & low_mask<BS>();
        if (kk == r) mine = brev_low<BS>(grp);
      }
      if (m != INT_MAX) {
        out_tiles[(base + k) * BS + r] = (W)mine;
        if (r == 0) out_col[base + k] = m;
      }
    } else if constexpr (MODE == PACK_VALUES) {
      // Rank of this row's first bit inside the tile = popcounts of the
      // rows above it (exclusive prefix over the group).
      int32_t cnt = (m != INT_MAX) ? popcw(w) : 0;
      int32_t incl = cnt;
      for (int off = 1; off < BS; off <<= 1) {
        const int32_t o = __shfl_up(incl, off, BS);
        if (r >= off) incl += o;
      }
      if (m != INT_MAX) {
        int64_t pos = woff[base + k] + (incl - cnt);
        for (int32_t q = p0; q < p; ++q) {
          if (q > p0 && ci[q] == ci[q - 1]) {
            tval[pos - 1] = min(tval[pos - 1], val[q]);
          } else {
            tval[pos++] = val[q];
          }
        }
      }
    }
    if (m != INT_MAX) ++k;
  }
  if constexpr (MODE == PACK_COUNT) {
    if (blk_ok && r == 0) {
      if (pc.cnt) pc.cnt[pi] = k;
      else bsr_rp[I] = k;
    }
  }
}

// Pieces per block row.
__global__ __launch_bounds__(256) void k_piece_n(const int32_t* __restrict__ rp, int32_t nrows, int32_t nbr, int bs,
                                                 int32_t* __restrict__ pn) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nbr) return;
  const int64_t r0 = I * bs, r1 = min(r0 + bs, (int64_t)nrows);
  const int64_t nnz = (int64_t)rp[r1] - rp[r0];
  pn[I] = (int32_t)max((int64_t)1, (nnz + kPieceNnz - 1) / kPieceNnz);
}

// Piece table: block row I's pieces from poff[I], cut at evenly spaced
// entries of its longest CSR row.
__global__ __launch_bounds__(256) void k_piece_fill(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                    int32_t nrows, int32_t nbr, int bs,
                                                    const int32_t* __restrict__ poff, int32_t* __restrict__ prow,
                                                    int32_t* __restrict__ plo, int32_t* __restrict__ phi) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nbr) return;
  const int32_t o = poff[I], P = poff[I + 1] - poff[I];
  int32_t best = 0, len = -1;
  const int64_t r0 = I * bs, r1 = min(r0 + bs, (int64_t)nrows);
  if (P > 1) {
    for (int64_t rr = r0; rr < r1; ++rr) {
      const int32_t l = rp[rr + 1] - rp[rr];
      if (l > len) {
        len = l;
        best = rp[rr];
      }
    }
  }
  for (int32_t j = 0; j < P; ++j) {
    prow[o + j] = (int32_t)I;
    plo[o + j] = j == 0 ? INT_MIN : ci[best + (int32_t)((int64_t)j * len / P)] / bs;
    phi[o + j] = j == P - 1 ? INT_MAX : ci[best + (int32_t)((int64_t)(j + 1) * len / P)] / bs;
  }
}

// rowptr[I] = first tile of block row I's first piece (rowptr[nbr] = total).
__global__ __launch_bounds__(256) void k_piece_rowptr(const int32_t* __restrict__ poff, const int64_t* __restrict__ toff,
                                                      int32_t nbr, int32_t* __restrict__ rowptr) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I <= nbr) rowptr[I] = (int32_t)toff[poff[I]];
}

template <int BS>
__global__ __launch_bounds__(256) void k_tile_popc(const word_t<BS>* __restrict__ tiles, int64_t T,
                                                   int64_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  int64_t s = 0;
  for (int r = 0; r < BS; ++r) s += popcw(tiles[t * BS + r]);
  cnt[t] = s;
}

template <int BS>
__global__ __launch_bounds__(256) void k_pack_vec(const float* __restrict__ x, word_t<BS>* __restrict__ xb,
                                                  int32_t n, int32_t nw) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nw) return;
  uint64_t w = 0;
  for (int k = 0; k < BS; ++k) {
    const int64_t i = (int64_t)j * BS + k;
    if (i < n && x[i] > 0.0f) w |= 1ull << (BS - 1 - k);
  }
  xb[j] = (word_t<BS>)w;
}

// First block row of every chunk of `chunk` tiles: the first row R with
// rowptr[R] >= b, or the row containing tile b when that row starts before b.
__global__ __launch_bounds__(256) void k_plan(const int32_t* __restrict__ rowptr, int32_t nbr,
                                              int32_t chunk, int32_t nchunks,
                                              int32_t* __restrict__ plan) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t b = (int64_t)c * chunk;
  int32_t lo = 0, hi = nbr;  // search in rowptr[0..nbr]
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if ((int64_t)rowptr[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  int32_t R = lo;
  if (R > nbr || (int64_t)rowptr[R] > b) R -= 1;
  if (R >= nbr) R = nbr > 0 ? nbr - 1 : 0;
  if (R < 0) R = 0;
  plan[c] = R;
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Tiles per wavefront chunk when the plan call gives none: 256 at every block
// size.  At bs 64 (one tile per wave instruction, scalar per-tile reads) 256
// beat 128 for the count (R-MAT s22 4.17-4.30 against 4.28-4.36 ms) and the
// sum (4.78 against 5.12 ms).
int default_chunk(int /*bs*/) { return 256; }

static bool bs_ok(int bs) { return bs == 8 || bs == 16 || bs == 32 || bs == 64; }

template <int BS, int MODE>
static int launch_pack(const int32_t* rp, const int32_t* ci, int32_t nrows, int32_t ncols, int32_t nbr,
                       int32_t* bsr_rp, int32_t* out_col, void* out_tiles, const uint32_t* val,
                       const int64_t* woff, uint32_t* tval, int* err, const PackPieces& pc, hipStream_t st) {
  constexpr int G = kWave / BS;
  const int64_t units = pc.row ? (int64_t)pc.n : (int64_t)nbr;
  const int64_t waves = (units + G - 1) / G;
  const int64_t blocks = (waves + 3) / 4;
  if (blocks == 0) return GBO_OK;
  k_pack<BS, MODE><<<(unsigned)blocks, 256, 0, st>>>(rp, ci, nrows, ncols, nbr, bsr_rp, out_col,
                                                     (word_t<BS>*)out_tiles, val, woff, tval, err, pc);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

template <int MODE>
static int dispatch_pack(int bs, const int32_t* rp, const int32_t* ci, int32_t nrows, int32_t ncols,
                         int32_t nbr, int32_t* bsr_rp, int32_t* out_col, void* out_tiles,
                         const uint32_t* val, const int64_t* woff, uint32_t* tval, int* err,
                         const PackPieces& pc, hipStream_t st) {
  switch (bs) {
    case 8: return launch_pack<8, MODE>(rp, ci, nrows, ncols, nbr, bsr_rp, out_col, out_tiles, val, woff, tval, err, pc, st);
    case 16: return launch_pack<16, MODE>(rp, ci, nrows, ncols, nbr, bsr_rp, out_col, out_tiles, val, woff, tval, err, pc, st);
    case 32: return launch_pack<32, MODE>(rp, ci, nrows, ncols, nbr, bsr_rp, out_col, out_tiles, val, woff, tval, err, pc, st);
    case 64: return launch_pack<64, MODE>(rp, ci, nrows, ncols, nbr, bsr_rp, out_col, out_tiles, val, woff, tval, err, pc, st);
  }
  return set_error(GBO_EUNSUPPORTED, "block size %d not supported (8, 16, 32, 64)", bs);
}

static int check_err_flag(int* d_err, hipStream_t st, const char* who) {
  int h = 0;
  GBO_HIP_CHECK(hipMemcpyAsync(&h, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if (h & PACKERR_COLRANGE) return set_error(GBO_EINVAL, "%s: column index out of range", who);
  if (h & PACKERR_UNSORTED) return set_error(GBO_EINVAL, "%s: CSR rows must be sorted ascending", who);
  return GBO_OK;
}

// The pieces of every block row, their tile counts (a count pass that also
// checks the rows: columns in range, block columns ascending) and the first
// tile of each (toff, npieces + 1 entries); *total = tiles.  Buffers are
// owned by `mem`; poff (nbr + 1) is each block row's first piece.
struct PieceSet {
  PackPieces pc;
  int32_t* poff = nullptr;
  int64_t* toff = nullptr;
  int64_t total = 0;
};

static int make_pieces(int bs, const int32_t* d_rp, const int32_t* d_ci, int32_t nrows, int32_t ncols, int32_t nbr,
                       AsyncAllocs& mem, PieceSet* ps, const char* who, hipStream_t st) {
  int rc;
  int32_t *poff = nullptr, *prow = nullptr, *plo = nullptr, *phi = nullptr, *pcnt = nullptr;
  int64_t* toff = nullptr;
  int* d_err = nullptr;
  if ((rc = mem.alloc(&poff, sizeof(int32_t) * ((size_t)nbr + 1), who)) || (rc = mem.alloc(&d_err, sizeof(int), who)))
    return rc;
  GBO_HIP_CHECK(hipMemsetAsync(d_err, 0, sizeof(int), st));
  if (nbr > 0) {
    k_piece_n<<<(unsigned)(((int64_t)nbr + 255) / 256), 256, 0, st>>>(d_rp, nrows, nbr, bs, poff);
    GBO_LAUNCH_CHECK();
  }
  if ((rc = scan_exclusive<int32_t, int32_t>(poff, poff, nbr, st))) return rc;
  int32_t np = 0;
  GBO_HIP_CHECK(hipMemcpyAsync(&np, poff + nbr, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if ((rc = mem.alloc(&prow, sizeof(int32_t) * std::max<int64_t>(np, 1), who)) ||
      (rc = mem.alloc(&plo, sizeof(int32_t) * std::max<int64_t>(np, 1), who)) ||
      (rc = mem.alloc(&phi, sizeof(int32_t) * std::max<int64_t>(np, 1), who)) ||
      (rc = mem.alloc(&pcnt, sizeof(int32_t) * std::max<int64_t>(np, 1), who)) ||
      (rc = mem.alloc(&toff, sizeof(int64_t) * ((size_t)np + 1), who)))
    return rc;
  if (nbr > 0) {
    k_piece_fill<<<(unsigned)(((int64_t)nbr + 255) / 256), 256, 0, st>>>(d_rp, d_ci, nrows, nbr, bs, poff, prow, plo,
                                                                         phi);
    GBO_LAUNCH_CHECK();
  }
  PackPieces pc;
  pc.row = prow;
  pc.lo = plo;
  pc.hi = phi;
  pc.cnt = pcnt;
  pc.n = np;
  if ((rc = dispatch_pack<PACK_COUNT>(bs, d_rp, d_ci, nrows, ncols, nbr, nullptr, nullptr, nullptr, nullptr,
                                      nullptr, nullptr, d_err, pc, st)))
    return rc;
  if ((rc = check_err_flag(d_err, st, who))) return rc;
  if ((rc = scan_exclusive<int32_t, int64_t>(pcnt, toff, np, st))) return rc;
  GBO_HIP_CHECK(hipMemcpyAsync(&ps->total, toff + np, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if (ps->total >= ((int64_t)1 << 31))
    return set_error(GBO_ERANGE, "tile count %lld exceeds int32", (long long)ps->total);
  pc.cnt = nullptr;
  pc.base = toff;
  ps->pc = pc;
  ps->poff = poff;
  ps->toff = toff;
  return GBO_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

PlanLayout plan_layout(int bs, int64_t nchunks, int32_t nbr, bool scratch) {
  PlanLayout L;
  size_t off = align256(sizeof(int32_t) * (size_t)nchunks);
  if (scratch) {
    L.slots = off;
    off += align256(sizeof(unsigned long long) * kCountSlots);
    L.count = off;
    off += 256;
    L.partial = off;
    off += align256(sizeof(double) * 2 * (size_t)nchunks * bs);
    L.bfs_counts = off;
    off += align256(sizeof(int32_t) * (size_t)bs * ((size_t)nbr + 1));
  }
  L.bytes = off;
  return L;
}

// Points the view's scratch pointers into a plan allocation made with scratch.
static void view_scratch(PlanView* pv, int bs, int32_t nbr) {
  const PlanLayout L = plan_layout(bs, pv->nchunks, nbr, true);
  char* base = (char*)const_cast<int32_t*>(pv->plan);
  pv->slots = (unsigned long long*)(base + L.slots);
  pv->count = (int64_t*)(base + L.count);
  pv->partial = (double*)(base + L.partial);
  pv->bfs_counts = (int32_t*)(base + L.bfs_counts);
}

// Plan entries plus (optionally) zeroed count slots in one allocation.
static int plan_fill(const gbo_bsr* A, int32_t chunk, int64_t nch, bool scratch, int32_t* plan, hipStream_t st) {
  k_plan<<<(unsigned)((nch + 255) / 256), 256, 0, st>>>(A->rowptr, A->nblockrows, chunk, (int32_t)nch, plan);
  GBO_LAUNCH_CHECK();
  if (scratch) {
    const PlanLayout L = plan_layout(A->bs, nch, A->nblockrows, true);
    unsigned long long* slots = (unsigned long long*)((char*)plan + L.slots);
    k_fill_u64<<<kCountSlots / 256, 256, 0, st>>>(slots, kCountSlots, 0ull);
    GBO_LAUNCH_CHECK();
  }
  return GBO_OK;
}

// Plan for a matrix: uses A->plan (made by gbo_bsr_plan, with its scratch)
// when present, else builds a temporary one with scratch (stream-ordered
// allocation, released by release_plan).
int acquire_plan(const gbo_bsr* A, PlanView* pv, hipStream_t st) {
  if (A->plan && A->plan_chunk > 0 && A->plan_nchunks > 0) {
    pv->plan = A->plan;
    pv->chunk = A->plan_chunk;
    pv->nchunks = A->plan_nchunks;
    pv->variant = A->plan_variant;
    pv->owned = false;
    view_scratch(pv, A->bs, A->nblockrows);
    return GBO_OK;
  }
  const int32_t chunk = default_chunk(A->bs);
  const int64_t nch = std::max<int64_t>(1, (A->nblocks + chunk - 1) / chunk);
  if (nch >= INT_MAX) return set_error(GBO_ERANGE, "too many chunks");
  const PlanLayout L = plan_layout(A->bs, nch, A->nblockrows, true);
  int32_t* plan = nullptr;
  GBO_HIP_CHECK(hipMallocAsync((void**)&plan, L.bytes, st));
  pv->plan = plan;
  pv->chunk = chunk;
  pv->nchunks = (int32_t)nch;
  pv->variant = A->plan_variant;
  pv->owned = true;
  view_scratch(pv, A->bs, A->nblockrows);
  return plan_fill(A, chunk, nch, true, plan, st);
}

int plan_create(gbo_bsr* A, int32_t chunk_tiles, int32_t variant, bool scratch, hipStream_t st) {
  const int32_t chunk = chunk_tiles > 0 ? chunk_tiles : default_chunk(A->bs);
  const int64_t nch = std::max<int64_t>(1, (A->nblocks + chunk - 1) / chunk);
  if (nch >= INT_MAX) return set_error(GBO_ERANGE, "gbo_bsr_plan: too many chunks");
  gbo_bsr_plan_free(A);
  const PlanLayout L = plan_layout(A->bs, nch, A->nblockrows, scratch);
  GBO_HIP_CHECK(hipMalloc((void**)&A->plan, L.bytes));
  A->plan_chunk = chunk;
  A->plan_nchunks = (int32_t)nch;
  A->plan_variant = variant;
  const int rc = plan_fill(A, chunk, nch, scratch, A->plan, st);
  if (rc) gbo_bsr_plan_free(A);
  return rc;
}

int release_plan(PlanView* pv, hipStream_t st) {
  if (pv->owned && pv->plan) GBO_HIP_CHECK(hipFreeAsync(const_cast<int32_t*>(pv->plan), st));
  pv->plan = nullptr;
  pv->owned = false;
  return GBO_OK;
}

int validate_bsr(const gbo_bsr* A, int want_dir, const char* who) {
  if (!A) return set_error(GBO_EINVAL, "%s: null matrix", who);
  if (!bs_ok(A->bs)) return set_error(GBO_EUNSUPPORTED, "%s: block size %d not supported", who, A->bs);
  if (want_dir >= 0 && A->dir != want_dir)
    return set_error(GBO_EINVAL, "%s: matrix must have dir %d", who, want_dir);
  if (A->nrows < 0 || A->ncols < 0 || A->nblocks < 0)
    return set_error(GBO_EINVAL, "%s: negative size", who);
  if (A->nblockrows != (A->nrows + A->bs - 1) / A->bs || A->nblockcols != (A->ncols + A->bs - 1) / A->bs)
    return set_error(GBO_EINVAL, "%s: nblockrows/nblockcols inconsistent with nrows/ncols", who);
  if (A->nblocks >= ((int64_t)1 << 31)) return set_error(GBO_ERANGE, "%s: nblocks exceeds int32", who);
  if (!A->rowptr || (A->nblocks > 0 && (!A->colind || !A->tiles)))
    return set_error(GBO_EINVAL, "%s: null device array", who);
  return GBO_OK;
}

}  // namespace gbo

using namespace gbo;


extern "C" {

int gbo_csr2bsr_nnzb_dev(int bs, const int32_t* d_rp, const int32_t* d_ci, int32_t nrows, int32_t ncols,
                         int32_t* d_bsrRowPtr, int64_t* nblocks, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  if (!bs_ok(bs)) return set_error(GBO_EUNSUPPORTED, "block size %d not supported (8, 16, 32, 64)", bs);
  if (nrows < 0 || ncols < 0 || !d_rp || !d_bsrRowPtr || !nblocks)
    return set_error(GBO_EINVAL, "gbo_csr2bsr_nnzb_dev: bad argument");
  hipStream_t st = stream;
  const int32_t nbr = (nrows + bs - 1) / bs;
  AsyncAllocs mem(st);
  PieceSet ps;
  int rc = make_pieces(bs, d_rp, d_ci, nrows, ncols, nbr, mem, &ps, "gbo_csr2bsr_nnzb_dev", st);
  if (rc != GBO_OK) return rc;
  k_piece_rowptr<<<(unsigned)(((int64_t)nbr + 1 + 255) / 256), 256, 0, st>>>(ps.poff, ps.toff, nbr, d_bsrRowPtr);
  GBO_LAUNCH_CHECK();
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  *nblocks = ps.total;
  return GBO_OK;
  GBO_TRY_END
}

int gbo_csr2bsr_fill_dev(int bs, int dir, const int32_t* d_rp, const int32_t* d_ci, int32_t nrows,
                         int32_t ncols, const int32_t* d_bsrRowPtr, int32_t* d_bsrColInd, void* d_tiles,
                         gbo_stream_t stream) {
  GBO_TRY_BEGIN
  if (!bs_ok(bs)) return set_error(GBO_EUNSUPPORTED, "block size %d not supported (8, 16, 32, 64)", bs);
  if (dir != 0 && dir != 1) return set_error(GBO_EINVAL, "dir must be 0 (row) or 1 (col)");
  if (nrows < 0 || ncols < 0 || !d_rp || !d_bsrRowPtr) return set_error(GBO_EINVAL, "gbo_csr2bsr_fill_dev: bad argument");
  hipStream_t st = stream;
  const int32_t nbr = (nrows + bs - 1) / bs;
  // the tile offsets come from the pieces' counts (the same as d_bsrRowPtr
  // at block-row starts); the rows are checked again on the way
  AsyncAllocs mem(st);
  PieceSet ps;
  int rc = make_pieces(bs, d_rp, d_ci, nrows, ncols, nbr, mem, &ps, "gbo_csr2bsr_fill_dev", st);
  if (rc != GBO_OK) return rc;
  if (ps.total > 0 && (!d_bsrColInd || !d_tiles)) return set_error(GBO_EINVAL, "gbo_csr2bsr_fill_dev: null output");
  int32_t given = 0;  // the caller sized colind and tiles by rowptr[nbr]
  GBO_HIP_CHECK(hipMemcpyAsync(&given, d_bsrRowPtr + nbr, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if (given != ps.total)
    return set_error(GBO_EINVAL, "gbo_csr2bsr_fill_dev: rowptr[nblockrows] = %d, the CSR has %lld tiles", given,
                     (long long)ps.total);
  if (dir == 0)
    rc = dispatch_pack<PACK_ROW>(bs, d_rp, d_ci, nrows, ncols, nbr, const_cast<int32_t*>(d_bsrRowPtr), d_bsrColInd,
                                 d_tiles, nullptr, nullptr, nullptr, nullptr, ps.pc, st);
  else
    rc = dispatch_pack<PACK_COL>(bs, d_rp, d_ci, nrows, ncols, nbr, const_cast<int32_t*>(d_bsrRowPtr), d_bsrColInd,
                                 d_tiles, nullptr, nullptr, nullptr, nullptr, ps.pc, st);
  return rc;
  GBO_TRY_END
}

int gbo_bsr_create_dev(int bs, int dir, const int32_t* d_rp, const int32_t* d_ci, int32_t nrows,
                       int32_t ncols, gbo_bsr* out, gbo_stream_t stream) {
  if (!out) return set_error(GBO_EINVAL, "gbo_bsr_create_dev: null output");
  std::memset(out, 0, sizeof(*out));
  if (!bs_ok(bs)) return set_error(GBO_EUNSUPPORTED, "block size %d not supported (8, 16, 32, 64)", bs);
  if (dir != 0 && dir != 1) return set_error(GBO_EINVAL, "dir must be 0 (row) or 1 (col)");
  const int32_t nbr = (nrows + bs - 1) / bs;
  gbo_bsr A{};
  A.bs = bs;
  A.dir = dir;
  A.nrows = nrows;
  A.ncols = ncols;
  A.nblockrows = nbr;
  A.nblockcols = (ncols + bs - 1) / bs;
  GBO_HIP_CHECK(hipMalloc((void**)&A.rowptr, sizeof(int32_t) * ((size_t)nbr + 1)));
  int rc = gbo_csr2bsr_nnzb_dev(bs, d_rp, d_ci, nrows, ncols, A.rowptr, &A.nblocks, stream);
  if (rc != GBO_OK) {
    (void)hipFree(A.rowptr);
    return rc;
  }
  const size_t tile_bytes = (size_t)A.nblocks * bs * (bs / 8);
  hipError_t e1 = hipMalloc((void**)&A.colind, sizeof(int32_t) * std::max<int64_t>(A.nblocks, 1));
  hipError_t e2 = hipMalloc(&A.tiles, std::max<size_t>(tile_bytes, 8));
  if (e1 != hipSuccess || e2 != hipSuccess) {
    gbo_bsr_free(&A);
    return set_error(GBO_ENOMEM, "gbo_bsr_create: cannot allocate %zu tile bytes", tile_bytes);
  }
  rc = gbo_csr2bsr_fill_dev(bs, dir, d_rp, d_ci, nrows, ncols, A.rowptr, A.colind, A.tiles, stream);
  if (rc != GBO_OK) {
    gbo_bsr_free(&A);
    return rc;
  }
  // variant 3: the fastest bs 8 count kernel on R-MAT s22 and, at bs 8/16,
  // the set-bit list min / min-plus (bs 16 count runs its variant 0)
  A.plan_variant = (bs <= 16) ? 3 : 0;
  *out = A;
  return GBO_OK;
}

// Uploads a host CSR, sorting a copy of any unsorted row first.
static int upload_csr(const int32_t* h_rp, const int32_t* h_ci, int32_t nrows, int32_t ncols,
                      int32_t** d_rp, int32_t** d_ci) {
  if (nrows < 0 || ncols < 0 || !h_rp) return set_error(GBO_EINVAL, "bad CSR");
  if (h_rp[0] != 0) return set_error(GBO_EINVAL, "rowptr[0] must be 0");
  for (int32_t r = 0; r < nrows; ++r)
    if (h_rp[r + 1] < h_rp[r]) return set_error(GBO_EINVAL, "rowptr not monotone at row %d", r);
  const int64_t nnz = h_rp[nrows];
  if (nnz > 0 && !h_ci) return set_error(GBO_EINVAL, "null colind");
  bool sorted = true;
  for (int32_t r = 0; r < nrows && sorted; ++r)
    for (int32_t k = h_rp[r] + 1; k < h_rp[r + 1]; ++k)
      if (h_ci[k] < h_ci[k - 1]) {
        sorted = false;
        break;
      }
  std::vector<int32_t> copy;
  const int32_t* src = h_ci;
  if (!sorted) {
    copy.assign(h_ci, h_ci + nnz);
    for (int32_t r = 0; r < nrows; ++r) std::sort(copy.begin() + h_rp[r], copy.begin() + h_rp[r + 1]);
    src = copy.data();
  }
  GBO_HIP_CHECK(hipMalloc((void**)d_rp, sizeof(int32_t) * ((size_t)nrows + 1)));
  GBO_HIP_CHECK(hipMalloc((void**)d_ci, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
  GBO_HIP_CHECK(hipMemcpy(*d_rp, h_rp, sizeof(int32_t) * ((size_t)nrows + 1), hipMemcpyHostToDevice));
  if (nnz) GBO_HIP_CHECK(hipMemcpy(*d_ci, src, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  return GBO_OK;
}

int gbo_csr2bsr_nnzb(int bs, const int32_t* h_rp, const int32_t* h_ci, int32_t nrows, int32_t ncols,
                     int64_t* nblocks) {
  GBO_TRY_BEGIN
  if (!bs_ok(bs)) return set_error(GBO_EUNSUPPORTED, "block size %d not supported (8, 16, 32, 64)", bs);
  if (!nblocks) return set_error(GBO_EINVAL, "null nblocks");
  int32_t *d_rp = nullptr, *d_ci = nullptr, *d_brp = nullptr;
  int rc = upload_csr(h_rp, h_ci, nrows, ncols, &d_rp, &d_ci);
  const int32_t nbr = (nrows + bs - 1) / bs;
  if (rc == GBO_OK && hipMalloc((void**)&d_brp, sizeof(int32_t) * ((size_t)nbr + 1)) != hipSuccess)
    rc = set_error(GBO_ENOMEM, "allocation failed");
  if (rc == GBO_OK) rc = gbo_csr2bsr_nnzb_dev(bs, d_rp, d_ci, nrows, ncols, d_brp, nblocks, nullptr);
  (void)hipFree(d_rp);
  (void)hipFree(d_ci);
  (void)hipFree(d_brp);
  return rc;
  GBO_TRY_END
}

int gbo_csr2bsr_batch(int bs, int dir, const int32_t* h_rp, const int32_t* h_ci, int32_t nrows,
                      int32_t ncols, int64_t nnz, int32_t* d_bsrRowPtr, int32_t* d_bsrColInd,
                      void* d_tiles, int32_t nblockrows, int64_t nblocks, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  if (!bs_ok(bs)) return set_error(GBO_EUNSUPPORTED, "block size %d not supported (8, 16, 32, 64)", bs);
  if (!h_rp || h_rp[nrows] != nnz) return set_error(GBO_EINVAL, "gbo_csr2bsr_batch: nnz != rowptr[nrows]");
  if (nblockrows != (nrows + bs - 1) / bs) return set_error(GBO_EINVAL, "gbo_csr2bsr_batch: nblockrows != ceil(nrows/bs)");
  int32_t *d_rp = nullptr, *d_ci = nullptr;
  int rc = upload_csr(h_rp, h_ci, nrows, ncols, &d_rp, &d_ci);
  int64_t got = 0;
  if (rc == GBO_OK) rc = gbo_csr2bsr_nnzb_dev(bs, d_rp, d_ci, nrows, ncols, d_bsrRowPtr, &got, stream);
  if (rc == GBO_OK && got != nblocks)
    rc = set_error(GBO_EINVAL, "gbo_csr2bsr_batch: nblocks %lld does not match the %lld tiles of the matrix",
                   (long long)nblocks, (long long)got);
  if (rc == GBO_OK) rc = gbo_csr2bsr_fill_dev(bs, dir, d_rp, d_ci, nrows, ncols, d_bsrRowPtr, d_bsrColInd, d_tiles, stream);
  if (rc == GBO_OK) GBO_HIP_CHECK(hipStreamSynchronize(stream));
  (void)hipFree(d_rp);
  (void)hipFree(d_ci);
  return rc;
  GBO_TRY_END
}

int gbo_bsr_create(int bs, int dir, const int32_t* h_rp, const int32_t* h_ci, int32_t nrows, int32_t ncols,
                   gbo_bsr* out, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  int32_t *d_rp = nullptr, *d_ci = nullptr;
  int rc = upload_csr(h_rp, h_ci, nrows, ncols, &d_rp, &d_ci);
  if (rc == GBO_OK) rc = gbo_bsr_create_dev(bs, dir, d_rp, d_ci, nrows, ncols, out, stream);
  if (rc == GBO_OK) GBO_HIP_CHECK(hipStreamSynchronize(stream));
  (void)hipFree(d_rp);
  (void)hipFree(d_ci);
  return rc;
  GBO_TRY_END
}

void gbo_bsr_plan_free(gbo_bsr* A) {
  if (!A) return;
  if (A->plan) (void)hipFree(A->plan);
  A->plan = nullptr;
  A->plan_chunk = 0;
  A->plan_nchunks = 0;
}

void gbo_bsr_free(gbo_bsr* A) {
  if (!A) return;
  gbo_bsr_plan_free(A);
  if (A->rowptr) (void)hipFree(A->rowptr);
  if (A->colind) (void)hipFree(A->colind);
  if (A->tiles) (void)hipFree(A->tiles);
  A->rowptr = nullptr;
  A->colind = nullptr;
  A->tiles = nullptr;
  A->nblocks = 0;
}

int gbo_bsr_plan(gbo_bsr* A, int32_t chunk_tiles, gbo_stream_t stream) {
  return gbo_bsr_plan_ex(A, chunk_tiles, A ? A->plan_variant : 0, stream);
}

int gbo_bsr_plan_ex(gbo_bsr* A, int32_t chunk_tiles, int32_t variant, gbo_stream_t stream) {
  int rc = validate_bsr(A, -1, "gbo_bsr_plan");
  if (rc) return rc;
  if (chunk_tiles < 0) return set_error(GBO_EINVAL, "gbo_bsr_plan: negative chunk");
  if (variant < 0 || variant > 3) return set_error(GBO_EINVAL, "gbo_bsr_plan: variant must be 0, 1, 2 or 3");
  return plan_create(A, chunk_tiles, variant, true, stream);
}

int gbo_csr2bsr_values_u32(const gbo_bsr* A, const int32_t* d_rp, const int32_t* d_ci, const uint32_t* d_val,
                           uint32_t* d_tval, int64_t* d_woff, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_csr2bsr_values_u32");
  if (rc) return rc;
  if (!d_rp || !d_val || !d_tval || !d_woff) return set_error(GBO_EINVAL, "gbo_csr2bsr_values_u32: null argument");
  hipStream_t st = stream;
  const int64_t T = A->nblocks;
  int64_t* cnt = nullptr;
  GBO_HIP_CHECK(hipMallocAsync((void**)&cnt, sizeof(int64_t) * std::max<int64_t>(T, 1), st));
  if (T > 0) {
    const unsigned blocks = (unsigned)((T + 255) / 256);
    switch (A->bs) {
      case 8: k_tile_popc<8><<<blocks, 256, 0, st>>>((const uint8_t*)A->tiles, T, cnt); break;
      case 16: k_tile_popc<16><<<blocks, 256, 0, st>>>((const uint16_t*)A->tiles, T, cnt); break;
      case 32: k_tile_popc<32><<<blocks, 256, 0, st>>>((const uint32_t*)A->tiles, T, cnt); break;
      case 64: k_tile_popc<64><<<blocks, 256, 0, st>>>((const uint64_t*)A->tiles, T, cnt); break;
    }
    GBO_LAUNCH_CHECK();
  }
  rc = scan_exclusive<int64_t, int64_t>(cnt, d_woff, T, st);
  (void)hipFreeAsync(cnt, st);
  if (rc) return rc;
  AsyncAllocs mem(st);
  PieceSet ps;
  if ((rc = make_pieces(A->bs, d_rp, d_ci, A->nrows, A->ncols, A->nblockrows, mem, &ps, "gbo_csr2bsr_values_u32", st)))
    return rc;
  if (ps.total != T) return set_error(GBO_EINVAL, "gbo_csr2bsr_values_u32: the CSR has %lld tiles, the matrix %lld",
                                      (long long)ps.total, (long long)T);
  return dispatch_pack<PACK_VALUES>(A->bs, d_rp, d_ci, A->nrows, A->ncols, A->nblockrows, A->rowptr, nullptr,
                                    nullptr, d_val, d_woff, d_tval, nullptr, ps.pc, st);
}

int gbo_pack_vec(int bs, const float* d_x, void* d_xbits, int32_t n, gbo_stream_t stream) {
  if (!bs_ok(bs)) return set_error(GBO_EUNSUPPORTED, "block size %d not supported (8, 16, 32, 64)", bs);
  if (n < 0 || (n > 0 && (!d_x || !d_xbits))) return set_error(GBO_EINVAL, "gbo_pack_vec: bad argument");
  const int32_t nw = (n + bs - 1) / bs;
  if (nw == 0) return GBO_OK;
  const unsigned blocks = (unsigned)((nw + 255) / 256);
  switch (bs) {
    case 8: k_pack_vec<8><<<blocks, 256, 0, stream>>>(d_x, (uint8_t*)d_xbits, n, nw); break;
    case 16: k_pack_vec<16><<<blocks, 256, 0, stream>>>(d_x, (uint16_t*)d_xbits, n, nw); break;
    case 32: k_pack_vec<32><<<blocks, 256, 0, stream>>>(d_x, (uint32_t*)d_xbits, n, nw); break;
    case 64: k_pack_vec<64><<<blocks, 256, 0, stream>>>(d_x, (uint64_t*)d_xbits, n, nw); break;
  }
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

}  // extern "C"
