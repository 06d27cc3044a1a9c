// bsrbmm: bit-BSR x bit-BSR products reduced to a scalar.
//
// Triangle counting (bmm{N}_sparse_masked_v5, bsrbmm/bsrbmm.cu:969,1214):
//   count = sum_{i,k} A_ik (A B)_ik,  A = L packed by rows, B = L^T packed
//   by columns (word k of B tile (J,K) = column k of that tile).
// The reference walks, per mask tile (I,K), every tile (I,J) of block row I
// and binary-searches B's block row J for K -- O(deg^2 log) per block row.
// Here the block-column structure of B is transposed once (a stable radix
// sort of the tiles by block column on the device: csr2csc on blocks) so
// that the J's that contribute to mask tile (I,K) are the intersection of
// two sorted lists: block row I of A and block column K of B.  Each wave
// takes a fixed chunk of mask tiles; its 64 lanes look up 64 entries of the
// shorter list at a time in the longer one (a binary search, or one load
// from the bitmap of a heavy row's block columns), or, when both lists have
// bitmaps, AND the two maps 64 words at a time.  Each match J adds
// sum_{a,k} M[a][k] (A_IJ B_JK)[a][k] = sum over the set bits (a,j) of A_IJ
// of popc(M[a] & B_JK row j) -- one popcount per set bit of A_IJ, with B's
// tiles turned into row form once per call.
//
// Plain product sum (bmm{N}_sparse, bsrbmm/bsrbmm.cu:296-555):
//   sum_{i,k} (A B)_ik = sum_j colsum_A(j) * rowsum_B(j),
// two linear passes over the tiles instead of the reference's tile-pair loop.
#include <algorithm>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {

struct TcArgs {
  const int32_t* a_rowptr;
  const int32_t* a_colind;
  const void* a_tiles;
  int32_t nbr;
  int64_t T;
  const int32_t* plan;
  int32_t chunk, nchunks;
  const int32_t* bt_ptr;   // block-CSC of B: column K lists [bt_ptr[K], bt_ptr[K+1])
  const int32_t* bt_row;   //   block row J of each entry (ascending inside a column)
  const int32_t* bt_tile;  //   tile index in B of each entry
  const void* b_tiles;
  unsigned long long* count;
  // heavy block rows of A (many tiles): a bitmap of the row's block columns
  // and the rank of each 64-bit word, so membership of J in the row and
  // J's tile index are two loads instead of a binary search
  const int32_t* hslot;    // per block row: slot of its bitmap, or -1 (nullable)
  const uint4* hmap;       // [slots][hwords]: x, y = the word's 64 bits (low,
                           // high), z = set bits in the words before it
  int64_t hwords;
  // per block column K of B: the slot whose bitmap holds B's column K list
  // (B = L^T, so column K of B is row K of L), or -1 (nullable)
  const int32_t* hcol;
  // bs 32/64 (nullable): the set bits of every A tile, row-major (entry =
  // row * BS + column), from aoff[t]; a match whose A tile has at most
  // kLaneMatchBits of them is counted by its own lane
  const int32_t* aoff;
  const uint16_t* abits;
  int vec16;  // bs 16: the tiles of A and B are 16-byte aligned (vector loads)
};

// A match (I,J,K) at bs 32/64 whose A tile (I,J) has at most this many set
// bits is counted by one lane (a popcount per bit from the tile's bit list)
// instead of by a group of bs lanes, one per tile row: R-MAT tiles hold 2-3
// bits, so the group form left all but a few of its lanes idle.
constexpr int32_t kLaneMatchBits = 16;

// A mask tile (I,K) whose row I and column K both have bitmaps intersects
// them by an AND of the two maps over the words both lists span when that
// takes at most 1/kAndRatio of the passes of the key lookups (64 words or 64
// keys a pass).  R-MAT s20: 5-13 % faster than the lookups alone with 1 or 2,
// the key lookups' gathers being dearer than the AND's coalesced loads.
constexpr int64_t kAndRatio = 2;

// First index in [lo, hi) of sorted arr whose value is >= key.
__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* __restrict__ arr, int64_t lo, int64_t hi, int32_t key) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (arr[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Row-packed copy of column-packed tiles (same positions): word j of the
// output holds row j of the tile, bit BS-1-k = column k.  One wave takes
// 64/BS tiles; lane l holds column word l % BS of tile l / BS, and one
// ballot per row j collects that row over the columns of every tile.
template <int BS>
__global__ __launch_bounds__(256) void k_tile_rows(const word_t<BS>* __restrict__ in, int64_t T,
                                                   word_t<BS>* __restrict__ out) {
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int64_t t = (int64_t)wave_index() * G + lane / BS;  // tile of this lane
  const int k = lane % BS;                                  // its column word
  const uint64_t w = t < T ? (uint64_t)in[t * BS + k] : 0ull;
  const int g = lane / BS;
#pragma unroll 8
  for (int j = 0; j < BS; ++j) {
    const uint64_t b = __ballot((w >> (BS - 1 - j)) & 1ull);  // bit l: column l % BS of tile l / BS has row j
    const uint64_t mine = (b >> (g * BS)) & low_mask<BS>();
    if (k == j && t < T) out[t * BS + j] = (word_t<BS>)brev_low<BS>(mine);
  }
}

// Triangles contributed by one matching (I,J,K): with A = tile (I,J) of L,
// M = mask tile (I,K) and R = B's tile (J,K) in row form,
//   sum_{a,k} M[a][k] sum_j A[a][j] R[j][k] = sum_{(a,j) in A} popc(M[a] & R[j]),
// one popcount per set bit of A.
template <int BS>
__device__ __forceinline__ uint64_t tc_hit(const word_t<BS>* __restrict__ A, const word_t<BS>* __restrict__ M,
                                           const word_t<BS>* __restrict__ R, bool vec16) {
  uint64_t s = 0;
  if constexpr (BS == 16) {
    if (vec16) {
    // A and M as four 64-bit words each (word q = rows 4q..4q+3, row 4q+i
    // in bits 16i..16i+15), all loaded together, instead of a chain of
    // row loads; R rows are gathered per set bit of A
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 a0 = *(const u32x4*)A, a1 = *((const u32x4*)A + 1);
    const u32x4 m0 = *(const u32x4*)M, m1 = *((const u32x4*)M + 1);
    const uint64_t aw[4] = {a0.x | ((uint64_t)a0.y << 32), a0.z | ((uint64_t)a0.w << 32),
                            a1.x | ((uint64_t)a1.y << 32), a1.z | ((uint64_t)a1.w << 32)};
    const uint64_t mw[4] = {m0.x | ((uint64_t)m0.y << 32), m0.z | ((uint64_t)m0.w << 32),
                            m1.x | ((uint64_t)m1.y << 32), m1.z | ((uint64_t)m1.w << 32)};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint64_t wa = aw[q];
      while (wa) {
        const int p = __builtin_ctzll(wa);
        wa &= wa - 1;
        const int i = p >> 4, col = 15 - (p & 15);  // row 4q+i, column col
        s += __popc((uint32_t)((mw[q] >> (16 * i)) & 0xFFFFu) & (uint32_t)R[col]);
      }
    }
    return s;
    }
  }
  if constexpr (BS == 8) {
    uint64_t a = *(const uint64_t*)A;  // byte r = row r, bit 7 - c = column c
    const uint64_t m = *(const uint64_t*)M, r = *(const uint64_t*)R;
    while (a) {
      const int p = __builtin_ctzll(a);
      a &= a - 1;
      const int row = p >> 3, col = 7 - (p & 7);
      s += __popc((uint32_t)((m >> (8 * row)) & (r >> (8 * col)) & 0xFFu));
    }
  } else {
    for (int row = 0; row < BS; ++row) {
      uint64_t wa = (uint64_t)A[row];
      if (!wa) continue;
      const uint64_t ma = (uint64_t)M[row];
      if (!ma) continue;
      while (wa) {
        const int p = __builtin_ctzll(wa);
        wa &= wa - 1;
        s += (uint64_t)popcw((word_t<BS>)(ma & (uint64_t)R[BS - 1 - p]));
      }
    }
  }
  return s;
}

// Eight waves per SIMD: the kernel is a chain of dependent loads, and
// without the bound its scalar registers (TcArgs) cost one wave per SIMD
// (a few of them move to vector-register lanes instead).
template <int BS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_tc(const TcArgs a) {
  using W = word_t<BS>;
  const int lane = lane_id();
  const int c = wave_index();
  if (c >= a.nchunks) return;
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const W* __restrict__ At = (const W*)a.a_tiles;
  const W* __restrict__ Rt = (const W*)a.b_tiles;  // B's tiles in row form
  uint64_t total = 0;

  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.a_rowptr[Rl];
      rel = a.a_rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && ((int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const int32_t hs = a.hslot ? a.hslot[R0 + j] : -1;  // wave-uniform
      const uint4* __restrict__ hm = hs >= 0 ? a.hmap + (int64_t)hs * a.hwords : nullptr;
      for (int64_t tm = s; tm < te; ++tm) {  // mask tile (I, K)
        const int32_t K = a.a_colind[tm];
        // Intersect block row I of A (J ascending) with block column K of B;
        // every lane that finds a match adds that (I,J,K) term itself.
        const int64_t la0 = rs, la1 = re;
        const int64_t lb0 = a.bt_ptr[K], lb1 = a.bt_ptr[K + 1];
        // a heavy row walks B's column and looks each J up in its bitmap; a
        // short row facing a heavy column walks its own tiles and looks them
        // up in the column's bitmap
        const bool use_map = hs >= 0 && (lb1 - lb0) <= (la1 - la0);
        const bool a_short = !use_map && (la1 - la0) <= (lb1 - lb0);
        const int32_t hk = (a_short && a.hcol) ? a.hcol[K] : -1;
        const uint4* __restrict__ cm = hk >= 0 ? a.hmap + (int64_t)hk * a.hwords : nullptr;
        const int64_t s0 = a_short ? la0 : lb0, s1 = a_short ? la1 : lb1;
        const int64_t l0 = a_short ? lb0 : la0, l1 = a_short ? lb1 : la1;
        const int32_t* skeys = a_short ? a.a_colind : a.bt_row;
        const int32_t* lkeys = a_short ? a.bt_row : a.a_colind;
        // bs 32/64: one lane per tile row; mask row r stays in a register
        const uint64_t mrow = (BS >= 32 && lane < (kWave / BS) * BS) ? (uint64_t)At[tm * BS + lane % BS] : 0ull;
        // adds the matches (tA, tB) of the lanes that have one (tA >= 0); every
        // lane of the wave calls it together
        auto process = [&](int64_t tA, int64_t tB) {
          if constexpr (BS <= 16) {
            // each lane adds its own match
            if (tA >= 0) total += tc_hit<BS>(At + tA * BS, At + tm * BS, Rt + tB * BS, a.vec16 != 0);
          } else {
            // 64/BS matches at a time, one per group of BS lanes (lane = row)
            constexpr int G = kWave / BS;
            const int grp = lane / BS, r = lane % BS;
            // a sparse A tile: this lane counts the match alone, from the
            // tile's bit list (mask row a, B row j per set bit (a, j))
            if (tA >= 0 && a.abits) {
              const int32_t o0 = a.aoff[tA], o1 = a.aoff[tA + 1];
              if (o1 - o0 <= kLaneMatchBits) {
                for (int32_t k = o0; k < o1; ++k) {
                  const uint32_t e = a.abits[k];
                  total += (uint64_t)popcw((W)((uint64_t)At[tm * BS + e / BS] & (uint64_t)Rt[tB * BS + e % BS]));
                }
                tA = -1;
              }
            }
            // the other matches: 64/BS at a time, one per group of BS lanes
            uint64_t hits = __ballot(tA >= 0);
            while (hits) {
              int hsel = -1;
#pragma unroll
              for (int g2 = 0; g2 < G; ++g2) {
                if (hits) {
                  const int h = __builtin_ctzll(hits);
                  hits &= hits - 1;
                  if (g2 == grp) hsel = h;
                }
              }
              const int src = hsel >= 0 ? hsel : 0;
              const int64_t hA = __shfl(tA, src), hB = __shfl(tB, src);
              if (hsel >= 0) {
                uint64_t wa = (uint64_t)At[hA * BS + r];
                if (mrow) {
                  while (wa) {
                    const int p = __builtin_ctzll(wa);
                    wa &= wa - 1;
                    total += (uint64_t)popcw((W)(mrow & (uint64_t)Rt[hB * BS + (BS - 1 - p)]));
                  }
                }
              }
            }
          }
        };
        // both rows have bitmaps: the common J are the set bits of the AND of
        // the two maps over the words both lists span
        const int32_t hK = (hs >= 0 && a.hcol) ? a.hcol[K] : -1;
        if (hK >= 0 && la1 > la0 && lb1 > lb0) {
          const int32_t jlo = max(a.a_colind[la0], a.bt_row[lb0]);
          const int32_t jhi = min(a.a_colind[la1 - 1], a.bt_row[lb1 - 1]);
          const int64_t len = min(la1 - la0, lb1 - lb0);
          const int64_t wpass = jhi >= jlo ? ((jhi >> 6) - (jlo >> 6) + kWave) / kWave : 0;
          if (wpass * kAndRatio <= (len + kWave - 1) / kWave) {
            const uint4* __restrict__ km = a.hmap + (int64_t)hK * a.hwords;
            for (int32_t w0 = jlo >> 6; jhi >= jlo && w0 <= (jhi >> 6); w0 += kWave) {
              const int32_t w = w0 + lane;
              uint64_t x = 0, wi = 0, wk = 0;
              uint32_t ri = 0, rk = 0;
              if (w <= (jhi >> 6)) {
                const uint4 ei = hm[w], ek = km[w];
                wi = (uint64_t)ei.x | ((uint64_t)ei.y << 32);
                wk = (uint64_t)ek.x | ((uint64_t)ek.y << 32);
                ri = ei.z;
                rk = ek.z;
                x = wi & wk;
              }
              // the chunk's matches spread over the lanes: match p is set bit
              // k of the word of lane o (prefix sums of the words' popcounts)
              const int cnt = __popcll(x);
              const int incl = wave_incl_scan(cnt);
              const int tot = __builtin_amdgcn_readlane(incl, kWave - 1);
              for (int mb = 0; mb < tot; mb += kWave) {
                const int p = mb + lane;
                int o = 0;
#pragma unroll
                for (int step = kWave / 2; step >= 1; step >>= 1)
                  if (__shfl(incl, o + step - 1) <= p) o += step;
                o = o < kWave ? o : kWave - 1;
                const int k = p - (__shfl(incl, o) - __shfl(cnt, o));
                uint64_t xo = __shfl(x, o);
                const uint64_t wio = __shfl(wi, o), wko = __shfl(wk, o);
                const uint32_t rio = __shfl(ri, o), rko = __shfl(rk, o);
                int64_t tA = -1, tB = -1;
                if (p < tot) {
                  // bit position of the k-th set bit of xo (k < popc(xo))
                  int b = 0;
#pragma unroll
                  for (int step = 32; step >= 1; step >>= 1) {
                    if (__popcll(xo & ((1ull << (b + step)) - 1ull)) <= k) b += step;  // b + step <= 63
                  }
                  const uint64_t below = (1ull << b) - 1ull;
                  tA = la0 + (int64_t)rio + __popcll(wio & below);
                  tB = a.bt_tile[lb0 + (int64_t)rko + __popcll(wko & below)];
                }
                process(tA, tB);
              }
            }
            continue;
          }
        }
        for (int64_t base = s0; base < s1; base += kWave) {
          const int64_t q = base + lane;
          int64_t tA = -1, tB = -1;
          if (q < s1) {
            const int32_t J = skeys[q];
            if (use_map) {
              const uint4 e = hm[J >> 6];  // word and rank in one load
              const uint64_t wd = (uint64_t)e.x | ((uint64_t)e.y << 32);
              if ((wd >> (J & 63)) & 1ull) {
                tA = la0 + (int64_t)e.z + __popcll(wd & ((1ull << (J & 63)) - 1ull));
                tB = a.bt_tile[q];
              }
            } else if (hk >= 0) {
              const uint4 e = cm[J >> 6];
              const uint64_t wd = (uint64_t)e.x | ((uint64_t)e.y << 32);
              if ((wd >> (J & 63)) & 1ull) {
                tA = q;
                tB = a.bt_tile[lb0 + (int64_t)e.z + __popcll(wd & ((1ull << (J & 63)) - 1ull))];
              }
            } else {
              const int64_t pos = lower_bound_i32(lkeys, l0, l1, J);
              if (pos < l1 && lkeys[pos] == J) {
                tA = a_short ? q : pos;
                tB = a_short ? a.bt_tile[pos] : a.bt_tile[q];
              }
            }
          }
          process(tA, tB);
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
  for (int off = 1; off < kWave; off <<= 1) total += __shfl_xor(total, off);
  if (lane == 0 && total) atomicAdd(a.count, (unsigned long long)total);
}

// Set bits per tile (a group of BS lanes per tile, lane = word).
template <int BS>
__global__ __launch_bounds__(256) void k_tile_bitcount(const word_t<BS>* __restrict__ tiles, int64_t T,
                                                       int32_t* __restrict__ cnt) {
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int64_t t = (int64_t)wave_index() * G + lane / BS;
  int c = t < T ? popcw(tiles[t * BS + lane % BS]) : 0;
  for (int off = 1; off < BS; off <<= 1) c += __shfl_xor(c, off);
  if (lane % BS == 0 && t < T) cnt[t] = c;
}

// Set-bit list of every tile from off[t]: entries row * BS + column, rows in
// order, columns ascending.
template <int BS>
__global__ __launch_bounds__(256) void k_tile_bitlist(const word_t<BS>* __restrict__ tiles, int64_t T,
                                                      const int32_t* __restrict__ off, uint16_t* __restrict__ list) {
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int64_t t = (int64_t)wave_index() * G + lane / BS;
  const int r = lane % BS;
  const uint64_t w = t < T ? (uint64_t)tiles[t * BS + r] : 0ull;
  const int pc = __popcll(w);
  int before = pc;  // set bits of this tile's rows before r
  for (int o = 1; o < BS; o <<= 1) {
    const int v = __shfl_up(before, o, BS);
    if (r >= o) before += v;
  }
  before -= pc;
  if (t >= T) return;
  int32_t at = off[t] + before;
  uint64_t m = w;
  while (m) {
    const int p = 63 - __builtin_clzll(m);  // highest set bit: column BS-1-p
    m &= ~(1ull << p);
    list[at++] = (uint16_t)(r * BS + (BS - 1 - p));
  }
}

// Bitmaps of the heavy rows: block column J of row rows[h] sets bit J of
// slot h (one workgroup per heavy row, its tiles taken strided).  Each
// 16-byte map entry holds a 64-bit word (x, y) and, after k_heavy_rank, the
// set bits of the words before it (z).
__global__ __launch_bounds__(256) void k_heavy_bits(const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colind,
                                                    const int32_t* __restrict__ rows, int64_t hwords,
                                                    uint4* __restrict__ map) {
  const int32_t I = rows[blockIdx.x];
  uint4* m = map + (int64_t)blockIdx.x * hwords;
  for (int64_t t = rowptr[I] + threadIdx.x; t < rowptr[I + 1]; t += blockDim.x) {
    const int32_t J = colind[t];
    atomicOr((unsigned long long*)&m[J >> 6], 1ull << (J & 63));
  }
}

// hcol[K] = h when B's column K list (bt_row over [bt_ptr[K], bt_ptr[K+1]))
// equals row K = rows[h] of A, so that slot h's bitmap and ranks also
// locate J inside B's column K (one workgroup per slot).
__global__ __launch_bounds__(256) void k_heavy_colcheck(const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind,
                                                        const int32_t* __restrict__ bt_ptr,
                                                        const int32_t* __restrict__ bt_row,
                                                        const int32_t* __restrict__ rows, int32_t* __restrict__ hcol) {
  __shared__ int same;
  const int32_t K = rows[blockIdx.x];
  const int32_t a0 = rowptr[K], n = rowptr[K + 1] - a0, b0 = bt_ptr[K];
  if (threadIdx.x == 0) same = (bt_ptr[K + 1] - b0 == n);
  __syncthreads();
  if (same) {
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x)
      if (colind[a0 + i] != bt_row[b0 + i]) same = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0 && same) hcol[K] = (int32_t)blockIdx.x;
}

// rank (z) of every word = set bits of the slot's words before it (one
// workgroup per slot: each thread sums a contiguous run of words, a block
// scan of the 256 run totals, then the runs are written out).
__global__ __launch_bounds__(256) void k_heavy_rank(int64_t hwords, uint4* __restrict__ map) {
  __shared__ int32_t part[256];
  uint4* m = map + (int64_t)blockIdx.x * hwords;
  const int64_t per = (hwords + 255) / 256;
  const int64_t w0 = min((int64_t)threadIdx.x * per, hwords), w1 = min(w0 + per, hwords);
  int32_t s = 0;
  for (int64_t w = w0; w < w1; ++w) s += __popc(m[w].x) + __popc(m[w].y);
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t acc = 0;
    for (int k = 0; k < 256; ++k) {
      const int32_t v = part[k];
      part[k] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int32_t acc = part[threadIdx.x];
  for (int64_t w = w0; w < w1; ++w) {
    const int32_t c = __popc(m[w].x) + __popc(m[w].y);
    m[w].z = (uint32_t)acc;
    acc += c;
  }
}

// colsum (rows=false): S[J*BS + c] += #rows of tile (I,J) with column c set.
// rowsum of a column-direction B (rows=true): word k of tile (J,K) holds
// column k with bit (BS-1-j) = row j, so S[J*BS + j] += #k with bit set --
// the same per-bit count over the words of the tile, indexed by the tile's
// block row instead of its block column.
template <int BS>
__global__ __launch_bounds__(256) void k_bitcount_by_col(const int32_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ colind,
                                                         const word_t<BS>* __restrict__ tiles, int64_t T,
                                                         int32_t nbr, bool by_block_row,
                                                         unsigned long long* __restrict__ S) {
  const int lane = lane_id();
  const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (t >= T) return;
  const uint64_t w = (lane < BS) ? (uint64_t)tiles[t * BS + lane] : 0ull;
  int32_t J;
  if (by_block_row) {
    // block row of tile t: upper_bound(rowptr, t) - 1
    int32_t lo = 0, hi = nbr;
    while (lo < hi) {
      const int32_t mid = lo + (hi - lo) / 2;
      if ((int64_t)rowptr[mid + 1] <= t) lo = mid + 1;
      else hi = mid;
    }
    J = lo;
  } else {
    J = colind[t];
  }
  int mycnt = 0;
  for (int cc = 0; cc < BS; ++cc) {
    const uint64_t b = __ballot((w >> (BS - 1 - cc)) & 1ull);
    if (lane == cc) mycnt = __popcll(b);
  }
  if (lane < BS && mycnt) atomicAdd(&S[(int64_t)J * BS + lane], (unsigned long long)mycnt);
}

__global__ __launch_bounds__(256) void k_dot_u64(const unsigned long long* __restrict__ x,
                                                 const unsigned long long* __restrict__ y, int64_t n,
                                                 unsigned long long* out) {
  uint64_t s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  for (int off = 1; off < kWave; off <<= 1) s += __shfl_xor(s, off);
  if (lane_id() == 0 && s) atomicAdd(out, (unsigned long long)s);
}

__global__ __launch_bounds__(256) void k_iota_i32(int32_t* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}

// ptr[c] = first position of sorted keys >= c, for c in [0, nbc].
__global__ __launch_bounds__(256) void k_col_bounds(const int32_t* __restrict__ keys, int64_t T, int32_t nbc,
                                                    int32_t* __restrict__ ptr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > nbc) return;
  ptr[c] = (int32_t)lower_bound_i32(keys, 0, T, (int32_t)c);
}

// row[i] = block row of tile tile[i]: the last J with rowptr[J] <= tile.
__global__ __launch_bounds__(256) void k_tile_block_row(const int32_t* __restrict__ rowptr, int32_t nbr,
                                                        const int32_t* __restrict__ tile, int64_t T,
                                                        int32_t* __restrict__ row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T) return;
  const int32_t t = tile[i];
  int32_t lo = 0, hi = nbr;  // first J with rowptr[J + 1] > t
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if (rowptr[mid + 1] <= t) lo = mid + 1;
    else hi = mid;
  }
  row[i] = lo;
}

}  // namespace gbo

using namespace gbo;

// Block-CSC of a block structure on the device (csr2csc on blocks): a stable
// radix sort of the tile indices by block column.  The tiles are stored in
// block-row order, so every column's list comes out with ascending rows.
// The three outputs are owned by `mem` (the caller's allocations).
static int block_transpose(const gbo_bsr* B, AsyncAllocs& mem, int32_t** d_ptr, int32_t** d_row, int32_t** d_tile,
                           hipStream_t st) {
  const int32_t nbr = B->nblockrows, nbc = B->nblockcols;
  const int64_t T = B->nblocks;
  const size_t tb = sizeof(int32_t) * (size_t)std::max<int64_t>(T, 1);
  int rc;
  if ((rc = mem.alloc(d_ptr, sizeof(int32_t) * ((size_t)nbc + 1), "block_transpose")) ||
      (rc = mem.alloc(d_row, tb, "block_transpose")) || (rc = mem.alloc(d_tile, tb, "block_transpose")))
    return rc;
  AsyncAllocs tmpmem(st);
  int32_t *keys = nullptr, *idx = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  unsigned bits = 1;
  while (bits < 31 && (1ll << bits) <= nbc) ++bits;
  if ((rc = tmpmem.alloc(&keys, tb, "block_transpose")) || (rc = tmpmem.alloc(&idx, tb, "block_transpose"))) return rc;
  k_iota_i32<<<(unsigned)((T + 255) / 256), 256, 0, st>>>(idx, T);
  GBO_LAUNCH_CHECK();
  GBO_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, B->colind, keys, idx, *d_tile, (size_t)T, 0, bits, st));
  if ((rc = tmpmem.alloc(&tmp, std::max<size_t>(tmp_bytes, 1), "block_transpose"))) return rc;
  GBO_HIP_CHECK(rocprim::radix_sort_pairs(tmp, tmp_bytes, B->colind, keys, idx, *d_tile, (size_t)T, 0, bits, st));
  k_col_bounds<<<(unsigned)(((int64_t)nbc + 1 + 255) / 256), 256, 0, st>>>(keys, T, nbc, *d_ptr);
  GBO_LAUNCH_CHECK();
  k_tile_block_row<<<(unsigned)((T + 255) / 256), 256, 0, st>>>(B->rowptr, nbr, *d_tile, T, *d_row);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

extern "C" {

int gbo_bsrbmm_tc(const gbo_bsr* L, const gbo_bsr* LT, uint64_t* d_count, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  int rc = validate_bsr(L, 0, "gbo_bsrbmm_tc(A)");
  if (rc) return rc;
  rc = validate_bsr(LT, 1, "gbo_bsrbmm_tc(B)");
  if (rc) return rc;
  if (L->bs != LT->bs) return set_error(GBO_EINVAL, "gbo_bsrbmm_tc: block sizes differ");
  if (L->nrows != L->ncols) return set_error(GBO_EINVAL, "gbo_bsrbmm_tc: the mask A must be square");
  if (L->ncols != LT->nrows || L->nrows != LT->ncols)
    return set_error(GBO_EINVAL, "gbo_bsrbmm_tc: B must be %d x %d", L->ncols, L->nrows);
  if (!d_count) return set_error(GBO_EINVAL, "gbo_bsrbmm_tc: null count");
  hipStream_t st = stream;
  GBO_HIP_CHECK(hipMemsetAsync(d_count, 0, sizeof(uint64_t), st));
  if (L->nblocks == 0 || LT->nblocks == 0) return GBO_OK;
  AsyncAllocs mem(st);
  int32_t *bt_ptr = nullptr, *bt_row = nullptr, *bt_tile = nullptr;
  if ((rc = block_transpose(LT, mem, &bt_ptr, &bt_row, &bt_tile, st))) return rc;
  // B's tiles in row form (same positions), for one popcount per set bit of A
  void* rows = nullptr;
  const size_t tbytes = (size_t)LT->nblocks * LT->bs * LT->bs / 8;
  if ((rc = mem.alloc(&rows, std::max<size_t>(tbytes, 8), "gbo_bsrbmm_tc"))) return rc;
  {
    const int64_t waves = (LT->nblocks * LT->bs + kWave - 1) / kWave;
    const unsigned tb = (unsigned)((waves + 3) / 4);
    switch (LT->bs) {
      case 8: k_tile_rows<8><<<tb, 256, 0, st>>>((const uint8_t*)LT->tiles, LT->nblocks, (uint8_t*)rows); break;
      case 16: k_tile_rows<16><<<tb, 256, 0, st>>>((const uint16_t*)LT->tiles, LT->nblocks, (uint16_t*)rows); break;
      case 32: k_tile_rows<32><<<tb, 256, 0, st>>>((const uint32_t*)LT->tiles, LT->nblocks, (uint32_t*)rows); break;
      case 64: k_tile_rows<64><<<tb, 256, 0, st>>>((const uint64_t*)LT->tiles, LT->nblocks, (uint64_t*)rows); break;
    }
    GBO_LAUNCH_CHECK();
  }
  TcArgs a{};
  a.a_rowptr = L->rowptr;
  a.a_colind = L->colind;
  a.a_tiles = L->tiles;
  a.nbr = L->nblockrows;
  a.T = L->nblocks;
  a.bt_ptr = bt_ptr;
  a.bt_row = bt_row;
  a.bt_tile = bt_tile;
  a.b_tiles = rows;
  a.count = (unsigned long long*)d_count;
  a.vec16 = (((uintptr_t)L->tiles) & 15) == 0;
  // bs 32/64: set bits per A tile and their scan; the total is read at the
  // synchronization below, and the bit lists are built when the tiles are
  // sparse (at most 8 set bits per tile on average)
  int32_t* aoff = nullptr;
  int32_t nbits = 0;
  // copies into host memory (nbits, hrp) are waited for on every return
  // until the explicit synchronization below
  struct SyncOnExit {
    hipStream_t st;
    bool armed = false;
    ~SyncOnExit() {
      if (armed) (void)hipStreamSynchronize(st);
    }
  } pending{st};
  std::vector<int32_t> hrp((size_t)L->nblockrows + 1);
  if (L->bs >= 32) {
    const int64_t T = L->nblocks;
    if ((rc = mem.alloc(&aoff, sizeof(int32_t) * ((size_t)T + 1), "gbo_bsrbmm_tc"))) return rc;
    const int64_t waves = (T * L->bs + kWave - 1) / kWave;
    const unsigned gb = (unsigned)((waves + 3) / 4);
    if (L->bs == 32) k_tile_bitcount<32><<<gb, 256, 0, st>>>((const uint32_t*)L->tiles, T, aoff);
    else k_tile_bitcount<64><<<gb, 256, 0, st>>>((const uint64_t*)L->tiles, T, aoff);
    GBO_LAUNCH_CHECK();
    if ((rc = scan_exclusive<int32_t, int32_t>(aoff, aoff, T, st))) return rc;
    pending.armed = true;
    GBO_HIP_CHECK(hipMemcpyAsync(&nbits, aoff + T, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  }
  // heavy rows (more than kHeavyTiles tiles) get a block-column bitmap, the
  // largest rows first, within a budget of device memory
  {
    constexpr int32_t kHeavyTiles = 64;
    constexpr int64_t kHeavyBudget = (int64_t)1 << 30;
    const int32_t nbr = L->nblockrows;
    const int64_t hwords = ((int64_t)L->nblockcols + 63) / 64;
    pending.armed = true;
    GBO_HIP_CHECK(hipMemcpyAsync(hrp.data(), L->rowptr, sizeof(int32_t) * hrp.size(), hipMemcpyDeviceToHost, st));
    GBO_HIP_CHECK(hipStreamSynchronize(st));
    pending.armed = false;
    if (aoff && (double)nbits <= 8.0 * (double)L->nblocks) {
      const int64_t T = L->nblocks;
      uint16_t* abits = nullptr;
      if ((rc = mem.alloc(&abits, sizeof(uint16_t) * (size_t)std::max<int32_t>(nbits, 1), "gbo_bsrbmm_tc")))
        return rc;
      const int64_t waves = (T * L->bs + kWave - 1) / kWave;
      const unsigned gb = (unsigned)((waves + 3) / 4);
      if (L->bs == 32) k_tile_bitlist<32><<<gb, 256, 0, st>>>((const uint32_t*)L->tiles, T, aoff, abits);
      else k_tile_bitlist<64><<<gb, 256, 0, st>>>((const uint64_t*)L->tiles, T, aoff, abits);
      GBO_LAUNCH_CHECK();
      a.aoff = aoff;
      a.abits = abits;
    }
    std::vector<int32_t> heavy;
    for (int32_t I = 0; I < nbr; ++I)
      if (hrp[(size_t)I + 1] - hrp[(size_t)I] > kHeavyTiles) heavy.push_back(I);
    std::sort(heavy.begin(), heavy.end(), [&](int32_t x, int32_t y) {
      return hrp[(size_t)x + 1] - hrp[(size_t)x] > hrp[(size_t)y + 1] - hrp[(size_t)y];
    });
    const int64_t per_row = hwords * (int64_t)sizeof(uint4);
    const size_t nh = (size_t)std::min<int64_t>((int64_t)heavy.size(), kHeavyBudget / std::max<int64_t>(per_row, 1));
    heavy.resize(nh);
    if (nh > 0) {
      std::vector<int32_t> slot((size_t)nbr, -1);
      for (size_t h = 0; h < nh; ++h) slot[(size_t)heavy[h]] = (int32_t)h;
      SyncOnExit sources{st};  // slot and heavy are sources of the copies below
      int32_t *d_slot = nullptr, *d_rows = nullptr;
      uint4* d_map = nullptr;
      if ((rc = mem.alloc(&d_slot, sizeof(int32_t) * (size_t)nbr, "gbo_bsrbmm_tc")) ||
          (rc = mem.alloc(&d_rows, sizeof(int32_t) * nh, "gbo_bsrbmm_tc")) ||
          (rc = mem.alloc(&d_map, sizeof(uint4) * nh * (size_t)hwords, "gbo_bsrbmm_tc")))
        return rc;
      sources.armed = true;
      GBO_HIP_CHECK(hipMemcpyAsync(d_slot, slot.data(), sizeof(int32_t) * (size_t)nbr, hipMemcpyHostToDevice, st));
      GBO_HIP_CHECK(hipMemcpyAsync(d_rows, heavy.data(), sizeof(int32_t) * nh, hipMemcpyHostToDevice, st));
      GBO_HIP_CHECK(hipMemsetAsync(d_map, 0, sizeof(uint4) * nh * (size_t)hwords, st));
      k_heavy_bits<<<(unsigned)nh, 256, 0, st>>>(L->rowptr, L->colind, d_rows, hwords, d_map);
      GBO_LAUNCH_CHECK();
      k_heavy_rank<<<(unsigned)nh, 256, 0, st>>>(hwords, d_map);
      GBO_LAUNCH_CHECK();
      // L is square, so block column K of B = L^T is block row K of L: the
      // row bitmaps serve B's heavy columns once the lists are checked equal
      int32_t* d_hcol = nullptr;
      if ((rc = mem.alloc(&d_hcol, sizeof(int32_t) * (size_t)nbr, "gbo_bsrbmm_tc"))) return rc;
      GBO_HIP_CHECK(hipMemsetAsync(d_hcol, 0xFF, sizeof(int32_t) * (size_t)nbr, st));
      k_heavy_colcheck<<<(unsigned)nh, 256, 0, st>>>(L->rowptr, L->colind, bt_ptr, bt_row, d_rows, d_hcol);
      GBO_LAUNCH_CHECK();
      a.hcol = d_hcol;
      // the host vectors above are sources of asynchronous copies
      GBO_HIP_CHECK(hipStreamSynchronize(st));
      sources.armed = false;
      a.hslot = d_slot;
      a.hmap = d_map;
      a.hwords = hwords;
    }
  }
  // TC work per mask tile is far larger than a bsrbmv tile: small chunks
  // (a plan without the bsrbmv scratch), made last so that no early return
  // above leaves it allocated
  gbo_bsr tmp = *L;
  tmp.plan = nullptr;
  tmp.plan_chunk = 0;
  if ((rc = plan_create(&tmp, 16, 0, false, st))) return rc;
  a.plan = tmp.plan;
  a.chunk = tmp.plan_chunk;
  a.nchunks = tmp.plan_nchunks;
  const unsigned blocks = (unsigned)(((int64_t)a.nchunks + 3) / 4);
  switch (L->bs) {
    case 8: k_tc<8><<<blocks, 256, 0, st>>>(a); break;
    case 16: k_tc<16><<<blocks, 256, 0, st>>>(a); break;
    case 32: k_tc<32><<<blocks, 256, 0, st>>>(a); break;
    case 64: k_tc<64><<<blocks, 256, 0, st>>>(a); break;
  }
  if (hipGetLastError() != hipSuccess) rc = set_error(GBO_EDEVICE, "k_tc launch failed");
  // the plan is a plain allocation: wait for the kernel before freeing it
  (void)hipStreamSynchronize(st);
  gbo_bsr_plan_free(&tmp);
  return rc;
  GBO_TRY_END
}

int gbo_bsrbmm_sum(const gbo_bsr* A, const gbo_bsr* B, uint64_t* d_sum, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmm_sum(A)");
  if (rc) return rc;
  rc = validate_bsr(B, 1, "gbo_bsrbmm_sum(B)");
  if (rc) return rc;
  if (A->bs != B->bs) return set_error(GBO_EINVAL, "gbo_bsrbmm_sum: block sizes differ");
  if (A->ncols != B->nrows) return set_error(GBO_EINVAL, "gbo_bsrbmm_sum: inner dimensions differ");
  if (!d_sum) return set_error(GBO_EINVAL, "gbo_bsrbmm_sum: null output");
  hipStream_t st = stream;
  const int bs = A->bs;
  const int64_t n = (int64_t)A->nblockcols * bs;
  AsyncAllocs mem(st);
  unsigned long long* S = nullptr;
  if ((rc = mem.alloc(&S, sizeof(unsigned long long) * 2 * std::max<int64_t>(n, 1), "gbo_bsrbmm_sum"))) return rc;
  GBO_HIP_CHECK(hipMemsetAsync(S, 0, sizeof(unsigned long long) * 2 * std::max<int64_t>(n, 1), st));
  GBO_HIP_CHECK(hipMemsetAsync(d_sum, 0, sizeof(uint64_t), st));
  auto launch = [&](const gbo_bsr* M, bool by_row, unsigned long long* out) {
    if (M->nblocks == 0) return;
    const unsigned blocks = (unsigned)((M->nblocks + 3) / 4);
    switch (bs) {
      case 8: k_bitcount_by_col<8><<<blocks, 256, 0, st>>>(M->rowptr, M->colind, (const uint8_t*)M->tiles, M->nblocks, M->nblockrows, by_row, out); break;
      case 16: k_bitcount_by_col<16><<<blocks, 256, 0, st>>>(M->rowptr, M->colind, (const uint16_t*)M->tiles, M->nblocks, M->nblockrows, by_row, out); break;
      case 32: k_bitcount_by_col<32><<<blocks, 256, 0, st>>>(M->rowptr, M->colind, (const uint32_t*)M->tiles, M->nblocks, M->nblockrows, by_row, out); break;
      case 64: k_bitcount_by_col<64><<<blocks, 256, 0, st>>>(M->rowptr, M->colind, (const uint64_t*)M->tiles, M->nblocks, M->nblockrows, by_row, out); break;
    }
  };
  launch(A, false, S);
  launch(B, true, S + n);
  GBO_LAUNCH_CHECK();
  k_dot_u64<<<256, 256, 0, st>>>(S, S + n, n, (unsigned long long*)d_sum);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

}  // extern "C"
