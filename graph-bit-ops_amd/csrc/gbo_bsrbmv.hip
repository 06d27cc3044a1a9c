// bsrbmv: bit-BSR x vector products on gfx950.
//
// Reference kernels: bmv{8,16,32,64}_sparse(_new) (count), bmv32_sparse_bin_
// masked_v4 (BFS), bmv32_sparse_full_minplus (SSSP), bmv32_sparse_full (PR),
// bmv32_sparse_full_cc_new (CC) in algorithm/bsrbmv.cu.  The reference maps
// one warp to one block row (lane = tile row) and so idles on power-law
// rows; here every wavefront takes a fixed-size CHUNK of consecutive tiles
// (a tile-balanced split, independent of the row structure) and walks the
// block rows that chunk touches.  Lane l of the wave is row (l % bs) of tile
// group (l / bs): a bs=64 tile is one coalesced 512-B load per wave, a bs=32
// step loads two tiles, bs=16 four, bs=8 eight.  Rows that a chunk boundary
// cuts ("split rows") are combined with integer atomics (count, BFS, min,
// min-plus: the result is order independent) or, for the FP32 sum, written
// as partials and added in chunk order by a fix-up kernel so the result is
// bitwise reproducible.
#include <type_traits>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {

// Tiles per group and round of the row-group min / sum loops: their words
// and column indices are loaded together, then one set bit of each is
// gathered per pass.  Eight measured faster than four for min and sum at bs
// 32/64 (R-MAT s22: 3-10 %; bs 32 sum 3.00 -> 2.69 ms once the round's
// loads were unconditional) and far slower for the weighted min-plus loop
// (bs 32 4.9 -> 6.8 ms), which keeps four.
template <int BS, int OP>
constexpr int kValueTiles = (BS >= 32) ? 8 : 4;

template <int NU>
__device__ __forceinline__ bool any_word(const uint64_t* wd) {
  uint64_t o = 0;
#pragma unroll
  for (int u = 0; u < NU; ++u) o |= wd[u];
  return o != 0;
}

// Identity written for a block row with no tiles (lane-per-row, see k_spmv).
template <int BS, int OP>
__device__ __forceinline__ void write_empty_row(const SpmvArgs& a, int32_t R) {
  if constexpr (OP == OP_BFS) {
    ((word_t<BS>*)a.out)[R] = (word_t<BS>)0;
  } else {
    for (int k = 0; k < BS; ++k) {
      const int64_t i = (int64_t)R * BS + k;
      if (i >= a.nrows) break;
      if constexpr (OP == OP_COUNT) {
        if (!a.accumulate) ((int32_t*)a.out)[i] = 0;
      } else if constexpr (OP == OP_MINPLUS) {
        ((uint32_t*)a.out)[i] = ((const uint32_t*)a.xval)[i];
      } else if constexpr (OP == OP_SUM) {
        ((float*)a.out)[i] = a.bias;
      } else if constexpr (OP == OP_MIN) {
        ((int32_t*)a.out)[i] = INT32_MAX;
      }
    }
  }
}

// Tile row word for the value products (min-plus, sum, min).  At bs 32/64
// the tiles are streamed once and far outweigh the gathered vector, so they
// are loaded nontemporally to leave L2 to the gathers (measured 2-8 % faster
// there; slightly slower at bs 8/16, which keep plain loads).
template <int BS>
__device__ __forceinline__ word_t<BS> ld_value_tile(const word_t<BS>* p) {
  if constexpr (BS >= 32) return __builtin_nontemporal_load(p);
  else return *p;
}

// Per-tile (or per-block-column) read at a tile index that is wave-uniform
// at bs 64 (one tile per wave instruction): a scalar load there, a vector
// load at the smaller block sizes, where a wave spans several tiles.
template <int BS, typename T>
__device__ __forceinline__ T ld_tile_meta(const T* p, int64_t i) {
  if constexpr (BS == 64) return ld_uniform(p, i);
  else return p[i];
}

template <int BS, int OP>
__global__ __launch_bounds__(256) void k_spmv(const SpmvArgs a) {
  using W = word_t<BS>;
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  // at bs 64 the group index is the constant 0, so tile indices stay
  // wave-uniform for the scalar loads below
  const int g = BS == 64 ? 0 : lane / BS, r = lane % BS;
  const int c = wave_index();
  if (c >= a.nchunks) return;  // wave-uniform
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const W* __restrict__ tiles = (const W*)a.tiles;
  const int32_t* __restrict__ colind = a.colind;
  int64_t added = 0;  // BFS

  // Walk the chunk's block rows 64 at a time: lane l looks at row R0+l.
  // Empty rows are written by their own lane; non-empty rows are then
  // processed one after another by the whole wave.
  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    if (belong && rel == rsl) write_empty_row<BS, OP>(a, Rl);
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int32_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      const int64_t i = (int64_t)R * BS + r;

      if constexpr (OP == OP_COUNT && BS == 64) {
        // One tile per wave instruction, so the tile index, its column index
        // and its x word are wave-uniform: read through constant-address-
        // space pointers they become scalar loads (one s_load_dwordx8 for
        // eight column indices, one s_load_dwordx2 per x word) and the vector
        // memory path carries only the tile stream.  R-MAT s22: 5.66 ->
        // 4.19 ms (32 VGPRs; sixteen tiles in flight were no faster).
        const W* __restrict__ xb = (const W*)a.xbits;
        int acc = 0;
        int64_t t = s;
        for (; t + 7 < te; t += 8) {
          int32_t cc[8];
          W ww[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            cc[u] = ld_uniform(colind, t + u);
            ww[u] = __builtin_nontemporal_load(&tiles[(t + u) * BS + r]);
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) acc += popcw((W)(ww[u] & ld_uniform(xb, cc[u])));
        }
        for (; t < te; ++t)
          acc += popcw((W)(__builtin_nontemporal_load(&tiles[t * BS + r]) & ld_uniform(xb, ld_uniform(colind, t))));
        if (i < a.nrows) {
          int32_t* y = (int32_t*)a.out;
          if (full) {
            if (a.accumulate) y[i] += acc;
            else y[i] = acc;
          } else if (acc) {
            atomicAdd(&y[i], acc);
          }
        }
      } else if constexpr (OP == OP_COUNT) {
        const W* __restrict__ xb = (const W*)a.xbits;
        // Tiles are streamed once: nontemporal loads keep L2 for x.  Eight
        // tiles per lane are in flight before the dependent x gathers.
        int acc = 0;
        int64_t t = s + g;
        for (; t + 7 * G < te; t += 8 * G) {
          int32_t cc[8];
          W ww[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            cc[u] = colind[t + u * G];
            ww[u] = __builtin_nontemporal_load(&tiles[(t + u * G) * BS + r]);
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) acc += popcw((W)(ww[u] & xb[cc[u]]));
        }
        for (; t < te; t += G) acc += popcw((W)(__builtin_nontemporal_load(&tiles[t * BS + r]) & xb[colind[t]]));
        for (int off = BS; off < kWave; off <<= 1) acc += __shfl_xor(acc, off);
        if (g == 0 && i < a.nrows) {
          int32_t* y = (int32_t*)a.out;
          if (full) {
            if (a.accumulate) y[i] += acc;
            else y[i] = acc;
          } else if (acc) {
            atomicAdd(&y[i], acc);
          }
        }
      } else if constexpr (OP == OP_BFS) {
        const W* __restrict__ fb = (const W*)a.xbits;
        W* nx = (W*)a.out;
        const W vword = ((const W*)a.visited)[R];
        const int64_t valid = (int64_t)a.nrows - (int64_t)R * BS;
        const uint64_t pad = valid >= BS ? 0ull : (low_mask<BS>() >> valid);
        if ((((uint64_t)vword | pad) & low_mask<BS>()) == low_mask<BS>()) {
          if (full && lane == 0) nx[R] = (W)0;  // every row of the block already visited
          continue;
        }
        // Tile words and column indices of eight tiles are loaded together,
        // then their frontier words: two dependent loads per eight tiles.
        // With a sparse frontier the column indices go first, then the
        // frontier words, and only the tiles whose frontier word is nonzero
        // are read (group-uniform).
        // At bs 64 the column indices and frontier words are wave-uniform
        // and are read with scalar loads (as in the count).
        auto col_at = [&](int64_t k) -> int32_t { return ld_tile_meta<BS>(colind, k); };
        auto front_at = [&](int32_t cx) -> W { return ld_tile_meta<BS>(fb, cx); };
        int hit = 0;
        int64_t t = s + g;
        if (a.sparse_frontier) {
          for (; t < te; t += 8 * G) {
            int32_t cc[8];
            W fw[8];
            // column and frontier-word loads unconditional (slots past the
            // row read tile s's) so each stage's loads are in flight together
#pragma unroll
            for (int u = 0; u < 8; ++u) cc[u] = col_at(t + u * G < te ? t + u * G : s);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const W f = front_at(cc[u]);
              fw[u] = t + u * G < te ? f : (W)0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (fw[u]) hit |= (W)(tiles[(t + u * G) * BS + r] & fw[u]) != 0;
          }
        }
        for (; t + 7 * G < te; t += 8 * G) {
          int32_t cc[8];
          W ww[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            cc[u] = col_at(t + u * G);
            ww[u] = tiles[(t + u * G) * BS + r];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) hit |= (W)(ww[u] & front_at(cc[u])) != 0;
        }
        for (; t < te; t += G) hit |= (W)(tiles[t * BS + r] & front_at(col_at(t))) != 0;
        const uint64_t bal = fold_groups<BS>(__ballot(hit));  // row r: any of its lane groups hit
        const uint64_t word = brev_low<BS>(bal) & ~(uint64_t)vword & low_mask<BS>();
        if (lane == 0) {
          if (full) {
            nx[R] = (W)word;
            added += popcw(word);
          } else if (word) {
            const uint64_t old = atomic_or_word<BS>(nx, R, word);
            added += popcw(word & ~old);
          }
        }
      } else if constexpr (OP == OP_MINPLUS) {
        const uint32_t* __restrict__ dist = (const uint32_t*)a.xval;
        uint32_t acc = kInfU32;
        // Four tiles per round: their words, column indices and weight
        // ranks are formed first, then one set bit of each tile is relaxed
        // per pass, so the distance gathers of a pass are independent.
        for (int64_t t = s + g; t < te; t += 4 * G) {
          int32_t cc[4];
          uint64_t wd[4];
          int rank[4];
          int64_t wo[4];  // first weight of each tile
          // three stages, each over the four tiles so that their loads are in
          // flight together: column indices (slots past the row load its
          // first tile, s < te), activity bits, then the words and weight
          // offsets of the active tiles -- an inactive source block column's
          // tile is not read (group-uniform)
          bool act[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) cc[u] = ld_tile_meta<BS>(colind, t + u * G < te ? t + u * G : s);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            // cc[u] is a valid column
            const uint64_t aw = a.active ? ld_tile_meta<BS>(a.active, cc[u] >> 6) : ~0ull;
            act[u] = t + u * G < te && ((aw >> (cc[u] & 63)) & 1ull);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            wd[u] = 0;
            wo[u] = 0;
            if (act[u]) {
              wd[u] = (uint64_t)ld_value_tile<BS>(&tiles[(t + u * G) * BS + r]);
              if (a.w) wo[u] = ld_tile_meta<BS>(a.woff, t + u * G);
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            // index of this row's last set bit among the tile's weights:
            // rows r' < r come first in tile bit order
            const int pc = popcw(wd[u]);
            int incl = pc;
            if (a.w) {
              if (!__ballot(pc > 1)) {
                // at most one set bit per row (the common case at bs 32/64):
                // count the earlier rows of this lane group that have one
                const uint64_t grp_rows = low_mask<BS>() << (g * BS);
                incl += __popcll(__ballot(pc != 0) & grp_rows & ((1ull << lane) - 1ull));
              } else {
                for (int off = 1; off < BS; off <<= 1) {
                  const int o = __shfl_up(incl, off, BS);
                  if (r >= off) incl += o;
                }
              }
            }
            rank[u] = incl - 1;  // set bits are taken from the last column back
          }
          while (wd[0] | wd[1] | wd[2] | wd[3]) {
            uint32_t dv[4], wt[4];  // the gathers of a pass are issued before any is used
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              dv[u] = kInfU32;
              wt[u] = 0u;
              if (wd[u]) {
                const int p = __builtin_ctzll(wd[u]);
                dv[u] = dist[(int64_t)cc[u] * BS + (BS - 1 - p)];
                wt[u] = a.w ? a.w[wo[u] + rank[u]] : 1u;
                wd[u] &= wd[u] - 1;
                --rank[u];
              }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const uint32_t cand = (dv[u] > kInfU32 - wt[u]) ? kInfU32 : dv[u] + wt[u];
              acc = cand < acc ? cand : acc;
            }
          }
        }
        for (int off = BS; off < kWave; off <<= 1) {
          const uint32_t o = __shfl_xor(acc, off);
          acc = o < acc ? o : acc;
        }
        if (g == 0 && i < a.nrows) {
          uint32_t* out = (uint32_t*)a.out;
          if (full) {
            const uint32_t d = dist[i];
            out[i] = acc < d ? acc : d;
          } else if (acc != kInfU32) {
            atomicMin(&out[i], acc);
          }
        }
      } else if constexpr (OP == OP_SUM) {
        // FP64 accumulation of the FP32 terms: a row sum (a hub row adds
        // >10^5 terms) keeps far more than FP32 resolution, in a fixed order
        // (lane order, then the butterfly), so results are reproducible.
        // Four tiles per round; one set bit of each per pass, so the x
        // gathers of a pass are independent.
        const float* __restrict__ x = (const float*)a.xval;
        double acc = 0.0;
        constexpr int NU = kValueTiles<BS, OP>;
        for (int64_t t = s + g; t < te; t += NU * G) {
          int32_t cc[NU];
          uint64_t wd[NU];
#pragma unroll
          for (int u = 0; u < NU; ++u) {
            // slots past the row load its first tile (s < te) and are then
            // emptied: the loads are unconditional, so all of them are in
            // flight together
            const int64_t tu = t + u * G;
            const int64_t tc = tu < te ? tu : s;
            cc[u] = ld_tile_meta<BS>(colind, tc);
            const uint64_t wv = (uint64_t)ld_value_tile<BS>(&tiles[tc * BS + r]);
            wd[u] = tu < te ? wv : 0ull;
          }
          while (any_word<NU>(wd)) {
            float v[NU];
#pragma unroll
            for (int u = 0; u < NU; ++u) {
              v[u] = 0.0f;
              if (wd[u]) {
                const int p = __builtin_ctzll(wd[u]);
                v[u] = x[(int64_t)cc[u] * BS + (BS - 1 - p)];
                wd[u] &= wd[u] - 1;
              }
            }
#pragma unroll
            for (int u = 0; u < NU; ++u) acc += (double)v[u];
          }
        }
        for (int off = BS; off < kWave; off <<= 1) acc += __shfl_xor(acc, off);
        if (g == 0 && i < a.nrows) {
          if (full) {
            ((float*)a.out)[i] = (float)((double)a.scale * acc + (double)a.bias);
          } else {
            const int slot = (rs < cb) ? 0 : 1;
            a.partial[((int64_t)c * 2 + slot) * BS + r] = acc;
          }
        }
      } else if constexpr (OP == OP_MIN) {
        const int32_t* __restrict__ x = (const int32_t*)a.xval;
        int32_t acc = INT32_MAX;
        constexpr int NU = kValueTiles<BS, OP>;
        for (int64_t t = s + g; t < te; t += NU * G) {
          int32_t cc[NU];
          uint64_t wd[NU];
#pragma unroll
          for (int u = 0; u < NU; ++u) {
            // slots past the row load its first tile (s < te) and are then
            // emptied: the loads are unconditional, so all of them are in
            // flight together
            const int64_t tu = t + u * G;
            const int64_t tc = tu < te ? tu : s;
            cc[u] = ld_tile_meta<BS>(colind, tc);
            const uint64_t wv = (uint64_t)ld_value_tile<BS>(&tiles[tc * BS + r]);
            wd[u] = tu < te ? wv : 0ull;
          }
          while (any_word<NU>(wd)) {
            int32_t v[NU];  // the NU gathers of a pass are issued before any is used
#pragma unroll
            for (int u = 0; u < NU; ++u) {
              v[u] = INT32_MAX;
              if (wd[u]) {
                const int p = __builtin_ctzll(wd[u]);
                v[u] = x[(int64_t)cc[u] * BS + (BS - 1 - p)];
                wd[u] &= wd[u] - 1;
              }
            }
#pragma unroll
            for (int u = 0; u < NU; ++u) acc = v[u] < acc ? v[u] : acc;
          }
        }
        for (int off = BS; off < kWave; off <<= 1) {
          const int32_t o = __shfl_xor(acc, off);
          acc = o < acc ? o : acc;
        }
        if (g == 0 && i < a.nrows) {
          int32_t* y = (int32_t*)a.out;
          if (full) y[i] = acc;
          else if (acc != INT32_MAX) atomicMin(&y[i], acc);
        }
      }
    }
    const int nb = __popcll(bmask);
    if (nb < kWave) break;
    R0 += kWave;
  }
  if constexpr (OP == OP_BFS) {
    if (lane == 0 && added) atomicAdd(&a.count_slots[c & (kCountSlots - 1)], (unsigned long long)added);
  }
}

// Lane-per-tile variant for bs 8 and 16 (COUNT and BFS).  A bs=8 tile is
// one uint64 (8 row bytes), a bs=16 tile four (4 row shorts each), so each
// lane streams whole tiles with 8- or 16-byte loads and a wave step covers
// 64 tiles instead of 64/bs.  Per-row popcounts are computed SWAR-style on
// the 64-bit words (x replicated into every byte/short) and kept in packed
// 16-bit fields; at the end of a block row a butterfly reduce-scatter leaves
// row r's total in lanes r*(64/bs) .. r*(64/bs)+64/bs-1.
template <int BS>
__device__ __forceinline__ void popc_fields(const uint64_t* w, uint64_t x, uint64_t* m) {
  // m[j] = per-row popcounts of (w[j] & x-replicated), one row per 16-bit
  // field (bs 16) or per byte (bs 8)
#pragma unroll
  for (int j = 0; j < BS * BS / 64; ++j) {
    uint64_t v;
    if constexpr (BS == 8) v = w[j] & (x * 0x0101010101010101ull);
    else v = w[j] & (x * 0x0001000100010001ull);
    v = v - ((v >> 1) & 0x5555555555555555ull);
    v = (v & 0x3333333333333333ull) + ((v >> 2) & 0x3333333333333333ull);
    v = (v + (v >> 4)) & 0x0F0F0F0F0F0F0F0Full;
    if constexpr (BS == 16) v = (v + (v >> 8)) & 0x00FF00FF00FF00FFull;
    m[j] = v;
  }
}

// Butterfly reduce-scatter of per-lane row partials v[0..BS) (v is
// overwritten): every lane of group l / (64/BS) gets the wave total of row
// l / (64/BS).
template <int BS>
__device__ __forceinline__ uint32_t reduce_scatter_rows(uint32_t* v) {
  constexpr int LPR = kWave / BS;
  const int lane = lane_id();
#pragma unroll
  for (int lvl = 0, nv = BS; nv > 1; ++lvl, nv >>= 1) {
    const int step = 32 >> lvl;
    const bool upper = (lane & step) != 0;
#pragma unroll
    for (int k = 0; k < nv / 2; ++k) {
      const uint32_t send = upper ? v[k] : v[k + nv / 2];
      const uint32_t keep = upper ? v[k + nv / 2] : v[k];
      v[k] = keep + (uint32_t)__shfl_xor((int)send, step);
    }
  }
  uint32_t tot = v[0];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) tot += (uint32_t)__shfl_xor((int)tot, off);
  return tot;
}

template <int BS>
__device__ __forceinline__ void load_tile(const void* tiles, int64_t t, uint64_t* w) {
  if constexpr (BS == 8) {
    w[0] = __builtin_nontemporal_load((const uint64_t*)tiles + t);
  } else {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* p = (const u32x4*)tiles + t * 2;
    const u32x4 q0 = __builtin_nontemporal_load(p), q1 = __builtin_nontemporal_load(p + 1);
    w[0] = ((uint64_t)q0.y << 32) | q0.x;
    w[1] = ((uint64_t)q0.w << 32) | q0.z;
    w[2] = ((uint64_t)q1.y << 32) | q1.x;
    w[3] = ((uint64_t)q1.w << 32) | q1.z;
  }
}

template <int BS, int OP>
__global__ __launch_bounds__(256) void k_spmv_lpt(const SpmvArgs a) {
  using W = word_t<BS>;
  constexpr int NW = BS * BS / 64;  // 64-bit words per tile
  constexpr int NA = (BS == 8) ? 2 : 4;  // packed 16-bit-field accumulators
  constexpr int LPR = kWave / BS;        // lanes per row after the reduce-scatter
  const int lane = lane_id();
  const int c = wave_index();
  if (c >= a.nchunks) return;
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const int32_t* __restrict__ colind = a.colind;
  const W* __restrict__ xb = (const W*)a.xbits;
  int64_t added = 0;

  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    if (belong && rel == rsl) write_empty_row<BS, OP>(a, Rl);
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int32_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);

      if constexpr (OP == OP_COUNT) {
        uint32_t v[BS];
#pragma unroll
        for (int k = 0; k < BS; ++k) v[k] = 0;
        uint64_t acc[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) acc[k] = 0;
        auto flush = [&]() {
          // field f of acc[k]: bs 8 -> row 2f + k; bs 16 -> row 4k + f
#pragma unroll
          for (int k = 0; k < NA; ++k) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
              const uint32_t val = (uint32_t)(acc[k] >> (16 * f)) & 0xFFFFu;
              if constexpr (BS == 8) v[2 * f + k] += val;
              else v[4 * k + f] += val;
            }
            acc[k] = 0;
          }
        };
        auto add_tile = [&](const uint64_t* w, uint64_t x) {
          uint64_t m[NW];
          popc_fields<BS>(w, x, m);
          if constexpr (BS == 8) {
            acc[0] += m[0] & 0x00FF00FF00FF00FFull;
            acc[1] += (m[0] >> 8) & 0x00FF00FF00FF00FFull;
          } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] += m[k];
          }
        };
        int since = 0;  // tiles added since the last flush (fields stay < 2^16)
        int64_t t = s + lane;
        for (; t + 3 * kWave < te; t += 4 * kWave) {
          uint64_t w0[NW], w1[NW], w2[NW], w3[NW];
          const int32_t c0 = colind[t], c1 = colind[t + kWave], c2 = colind[t + 2 * kWave], c3 = colind[t + 3 * kWave];
          load_tile<BS>(a.tiles, t, w0);
          load_tile<BS>(a.tiles, t + kWave, w1);
          load_tile<BS>(a.tiles, t + 2 * kWave, w2);
          load_tile<BS>(a.tiles, t + 3 * kWave, w3);
          const uint64_t x0 = xb[c0], x1 = xb[c1], x2 = xb[c2], x3 = xb[c3];
          add_tile(w0, x0);
          add_tile(w1, x1);
          add_tile(w2, x2);
          add_tile(w3, x3);
          since += 4;
          if (since >= 2048) {
            flush();
            since = 0;
          }
        }
        for (; t < te; t += kWave) {
          uint64_t w0[NW];
          load_tile<BS>(a.tiles, t, w0);
          add_tile(w0, xb[colind[t]]);
        }
        flush();
        const uint32_t tot = reduce_scatter_rows<BS>(v);
        const int row = lane / LPR;
        const int64_t i = (int64_t)R * BS + row;
        if ((lane % LPR) == 0 && i < a.nrows) {
          int32_t* y = (int32_t*)a.out;
          if (full) {
            if (a.accumulate) y[i] += (int32_t)tot;
            else y[i] = (int32_t)tot;
          } else if (tot) {
            atomicAdd(&y[i], (int32_t)tot);
          }
        }
      } else if constexpr (OP == OP_BFS) {
        W* nx = (W*)a.out;
        const W vword = ((const W*)a.visited)[R];
        const int64_t valid = (int64_t)a.nrows - (int64_t)R * BS;
        const uint64_t pad = valid >= BS ? 0ull : (low_mask<BS>() >> valid);
        if ((((uint64_t)vword | pad) & low_mask<BS>()) == low_mask<BS>()) {
          if (full && lane == 0) nx[R] = (W)0;
          continue;
        }
        uint64_t hit[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) hit[k] = 0;
        // tiles and column indices of four steps are loaded before their
        // frontier words
        auto or_tile = [&](const uint64_t* w, uint64_t f) {
          const uint64_t xr = (BS == 8) ? f * 0x0101010101010101ull : f * 0x0001000100010001ull;
#pragma unroll
          for (int k = 0; k < NW; ++k) hit[k] |= w[k] & xr;
        };
        int64_t t = s + lane;
        for (; t + 3 * kWave < te; t += 4 * kWave) {
          uint64_t w[4][NW];
          int32_t cc[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            cc[u] = colind[t + u * kWave];
            load_tile<BS>(a.tiles, t + u * kWave, w[u]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) or_tile(w[u], (uint64_t)xb[cc[u]]);
        }
        for (; t < te; t += kWave) {
          uint64_t w[NW];
          load_tile<BS>(a.tiles, t, w);
          or_tile(w, (uint64_t)xb[colind[t]]);
        }
#pragma unroll
        for (int k = 0; k < NW; ++k)
          hit[k] = wave_or64(hit[k]);
        // row r hit <=> its byte (bs 8) / short (bs 16) of hit is non-zero
        uint64_t word = 0;
#pragma unroll
        for (int r = 0; r < BS; ++r) {
          const uint64_t field = (BS == 8) ? (hit[0] >> (8 * r)) & 0xFFull : (hit[r / 4] >> (16 * (r % 4))) & 0xFFFFull;
          if (field) word |= 1ull << (BS - 1 - r);
        }
        word &= ~(uint64_t)vword & low_mask<BS>();
        if (lane == 0) {
          if (full) {
            nx[R] = (W)word;
            added += popcw(word);
          } else if (word) {
            const uint64_t old = atomic_or_word<BS>(nx, R, word);
            added += popcw(word & ~old);
          }
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
  if constexpr (OP == OP_BFS) {
    if (lane == 0 && added) atomicAdd(&a.count_slots[c & (kCountSlots - 1)], (unsigned long long)added);
  }
}

// Streamed lane-per-tile COUNT kernel for bs 8 and 16.  The chunk's tiles
// are read as one contiguous stream whatever rows they belong to: lane l
// owns TPL tiles of every wave step, in vector groups of 2 consecutive
// tiles (a 16-byte load for bs=8, four for bs=16, and an 8-byte load of
// their column indices), and the loads of step k+1 are issued before step k
// is reduced, so a wave keeps a whole step in flight across row boundaries
// and the first step is issued before the row walk starts.  Per-row popcounts are formed
// SWAR-style in 16-bit fields; the block rows are walked with the 64-row
// rowptr window of k_spmv, and a row's fields are reduced across the wave
// only when the row ends.
template <int BS>
struct StreamSlots {
  static constexpr int TPL = (BS == 8) ? 4 : 2;  // tiles per lane per step
  static constexpr int GT = 2;                   // consecutive tiles per vector group
  static constexpr int NG = TPL / GT;            // vector groups per lane per step
  static constexpr int NWT = BS * BS / 64;       // 64-bit words per tile
  uint64_t w[TPL * NWT];
  int32_t col[TPL];
  // Tile index of slot u in the step starting at tile ts.  Group q of the
  // wave covers 64*GT consecutive tiles, GT per lane, so each vector load
  // instruction reads one contiguous block (1 KB of bs=8 tiles).
  static __device__ __forceinline__ int64_t tile(int64_t ts, int lane, int u) {
    return ts + (int64_t)(u / GT) * (kWave * GT) + (int64_t)lane * GT + (u % GT);
  }
};

// Loads the step starting at tile ts (a multiple of TPL).  A vector group
// with no tile in [lo, hi) is not read; a group that runs past the last
// tile is read tile by tile.  Slots that are not read get column 0 and zero
// words.
template <int BS>
__device__ __forceinline__ void stream_load(const SpmvArgs& a, int64_t ts, int lane, int64_t lo, int64_t hi,
                                            StreamSlots<BS>& s) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef int i32x2 __attribute__((ext_vector_type(2)));
  using S = StreamSlots<BS>;
  constexpr int GT = S::GT, NWT = S::NWT;
#pragma unroll
  for (int q = 0; q < S::NG; ++q) {
    const int64_t g0 = S::tile(ts, lane, q * GT);
    const bool touch = (g0 + GT > lo) && (g0 < hi);
    if (touch && g0 + GT <= a.T) {
      const u32x4* tp = (const u32x4*)((const uint64_t*)a.tiles + g0 * NWT);
#pragma unroll
      for (int k = 0; k < GT * NWT / 2; ++k) {
        const u32x4 v = __builtin_nontemporal_load(tp + k);
        s.w[q * GT * NWT + 2 * k] = ((uint64_t)v.y << 32) | v.x;
        s.w[q * GT * NWT + 2 * k + 1] = ((uint64_t)v.w << 32) | v.z;
      }
      const i32x2 cv = __builtin_nontemporal_load((const i32x2*)(a.colind + g0));
      s.col[q * GT] = cv.x;
      s.col[q * GT + 1] = cv.y;
    } else {
#pragma unroll
      for (int e = 0; e < GT; ++e) {
        const int64_t t = g0 + e;
        const bool ok = touch && t < a.T;
#pragma unroll
        for (int k = 0; k < NWT; ++k) s.w[(q * GT + e) * NWT + k] = ok ? ((const uint64_t*)a.tiles)[t * NWT + k] : 0ull;
        s.col[q * GT + e] = ok ? a.colind[t] : 0;
      }
    }
  }
}

// Rows W0 .. W0+63 of the rowptr window, lane l holding row W0 + l.
__device__ __forceinline__ void window_load(const SpmvArgs& a, int32_t W0, int lane, int32_t& rsl, int32_t& rel) {
  const int32_t Rl = W0 + lane;
  rsl = 0;
  rel = 0;
  if (Rl < a.nbr) {
    rsl = a.rowptr[Rl];
    rel = a.rowptr[Rl + 1];
  }
}

// One chunk of the streamed count: `nxt` holds the chunk's first step
// (already issued) and rsl/rel the rowptr window starting at W0 = plan[c].
template <int BS>
__device__ __forceinline__ void stream_chunk(const SpmvArgs& a, int c, StreamSlots<BS>& nxt, int32_t W0,
                                             int32_t rsl, int32_t rel) {
  using W = word_t<BS>;
  using S = StreamSlots<BS>;
  constexpr int TPL = S::TPL, NWT = S::NWT;
  constexpr int NA = (BS == 8) ? 2 : 4;  // packed 16-bit-field accumulators
  constexpr int LPR = kWave / BS;        // lanes per row after the reduce-scatter
  constexpr int64_t STEP = (int64_t)kWave * TPL;
  const int lane = lane_id();
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const W* __restrict__ xb = (const W*)a.xbits;
  int32_t* __restrict__ y = (int32_t*)a.out;
  const int64_t base = cb - cb % TPL;

  uint64_t work = 0;
  bool more = false;
  auto apply_window = [&]() {
    const int32_t Rl = W0 + lane;
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    more = (__ballot(belong) == ~0ull);
    if (belong && rel == rsl) write_empty_row<BS, OP_COUNT>(a, Rl);
    work = __ballot(belong && rel > rsl);
  };
  int32_t R = 0;
  int64_t rs = 0, re = 0;
  auto next_row = [&]() -> bool {
    while (!work) {
      if (!more) return false;
      W0 += kWave;
      window_load(a, W0, lane, rsl, rel);
      apply_window();
    }
    const int j = __builtin_ctzll(work);
    work &= work - 1;
    R = W0 + j;
    rs = __builtin_amdgcn_readlane(rsl, j);
    re = __builtin_amdgcn_readlane(rel, j);
    return true;
  };

  uint32_t v[BS];
  uint64_t acc[NA];
#pragma unroll
  for (int k = 0; k < BS; ++k) v[k] = 0;
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0;
  auto flush = [&]() {
    // field f of acc[k]: bs 8 -> row 2f + k; bs 16 -> row 4k + f
#pragma unroll
    for (int k = 0; k < NA; ++k) {
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const uint32_t val = (uint32_t)(acc[k] >> (16 * f)) & 0xFFFFu;
        if constexpr (BS == 8) v[2 * f + k] += val;
        else v[4 * k + f] += val;
      }
      acc[k] = 0;
    }
  };
  auto finish_row = [&]() {
    flush();
    const uint32_t tot = reduce_scatter_rows<BS>(v);
    const int64_t i = (int64_t)R * BS + lane / LPR;
    if ((lane % LPR) == 0 && i < a.nrows) {
      if (rs >= cb && re <= ce) {
        if (a.accumulate) y[i] += (int32_t)tot;
        else y[i] = (int32_t)tot;
      } else if (tot) {
        atomicAdd(&y[i], (int32_t)tot);
      }
    }
#pragma unroll
    for (int k = 0; k < BS; ++k) v[k] = 0;
  };

  apply_window();
  bool open = next_row();  // a row is being accumulated
  int since = 0;           // steps added since the last flush (fields stay < 2^16)
  for (int64_t ts = base; open && ts < ce; ts += STEP) {
    const S cur = nxt;
    if (ts + STEP < ce) stream_load<BS>(a, ts + STEP, lane, cb, ce, nxt);
    // per-slot per-row popcounts: bytes (bs 8) or 16-bit fields (bs 16)
    uint64_t m[TPL][NWT];
#pragma unroll
    for (int u = 0; u < TPL; ++u) popc_fields<BS>(&cur.w[u * NWT], (uint64_t)xb[cur.col[u]], m[u]);
    const int64_t tend = min(ts + STEP, ce);
    while (true) {
      const int64_t lo = rs > cb ? rs : cb, hi = re < tend ? re : tend;
      uint64_t sum[NWT];
#pragma unroll
      for (int k = 0; k < NWT; ++k) sum[k] = 0;
#pragma unroll
      for (int u = 0; u < TPL; ++u) {
        const int64_t tu = S::tile(ts, lane, u);
        const bool in = (tu >= lo) && (tu < hi);
#pragma unroll
        for (int k = 0; k < NWT; ++k) sum[k] += in ? m[u][k] : 0ull;
      }
      // bs 8: at most TPL*8 = 32 per byte, so the byte sums do not carry
      if constexpr (BS == 8) {
        acc[0] += sum[0] & 0x00FF00FF00FF00FFull;
        acc[1] += (sum[0] >> 8) & 0x00FF00FF00FF00FFull;
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += sum[k];
      }
      if (re > tend) break;  // the row goes on past this step
      finish_row();
      since = 0;
      open = next_row();
      if (!open || rs >= tend) break;
    }
    if (++since >= 1024) {  // <= 32 per field per step
      flush();
      since = 0;
    }
  }
  if (open) finish_row();  // a row that continues past the chunk: atomic partial
}

// First tile of chunk c's first step (a multiple of the tiles per lane).
template <int BS>
__device__ __forceinline__ int64_t stream_base(const SpmvArgs& a, int c) {
  const int64_t cb = (int64_t)c * a.chunk;
  return cb - cb % StreamSlots<BS>::TPL;
}

template <int BS>
__device__ __forceinline__ void stream_first(const SpmvArgs& a, int c, int lane, StreamSlots<BS>& s) {
  const int64_t cb = (int64_t)c * a.chunk;
  stream_load<BS>(a, stream_base<BS>(a, c), lane, cb, min(cb + (int64_t)a.chunk, a.T), s);
}

// Variant 2: one wavefront per chunk.  The first step's loads are issued
// before the plan entry and the rowptr window.
template <int BS>
__global__ __launch_bounds__(256) void k_count_stream(const SpmvArgs a) {
  const int lane = lane_id();
  const int c = wave_index();
  if (c >= a.nchunks) return;  // wave-uniform
  StreamSlots<BS> nxt;
  stream_first<BS>(a, c, lane, nxt);
  const int32_t W0 = a.plan[c];
  int32_t rsl, rel;
  window_load(a, W0, lane, rsl, rel);
  stream_chunk<BS>(a, c, nxt, W0, rsl, rel);
}

// Variant 3, bs 8 only: the streamed count with 32-bit chunk-relative
// indices and byte-sliced popcounts.  Per step each lane holds 4 tiles
// (slots u = 2q + e are tiles ps + 128q + 2*lane + e: 16-byte loads, each
// wave instruction reading one contiguous block).  The 4 x 8 row bytes are
// transposed with v_perm so that word r holds row r of the 4 tiles; ANDed
// with the 4 gathered x bytes packed the same way, one v_bcnt per row adds
// 4 tiles.  A step that lies inside one block row takes this path
// unmasked; at row boundaries the slots of other rows are zeroed first.
// Row totals are reduced with permlane swaps and DPP (no LDS).
__device__ __forceinline__ void transpose_rows4(uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3, uint32_t* r) {
  // v_perm_b32(S0, S1, sel): selector byte k picks byte sel_k of {S0:S1}
  // (0-3 = S1, 4-7 = S0)
  const uint32_t u0 = __builtin_amdgcn_perm(t1, t0, 0x05010400u);  // t0.b0 t1.b0 t0.b1 t1.b1
  const uint32_t u1 = __builtin_amdgcn_perm(t1, t0, 0x07030602u);  // t0.b2 t1.b2 t0.b3 t1.b3
  const uint32_t u2 = __builtin_amdgcn_perm(t3, t2, 0x05010400u);
  const uint32_t u3 = __builtin_amdgcn_perm(t3, t2, 0x07030602u);
  r[0] = __builtin_amdgcn_perm(u2, u0, 0x05040100u);  // t0.b0 t1.b0 t2.b0 t3.b0
  r[1] = __builtin_amdgcn_perm(u2, u0, 0x07060302u);
  r[2] = __builtin_amdgcn_perm(u3, u1, 0x05040100u);
  r[3] = __builtin_amdgcn_perm(u3, u1, 0x07060302u);
}

// Wave total of row lane / 8 for per-lane row partials v[0..8), in every
// lane of that row's group of 8 (v is overwritten).  Levels: lanes l and
// l^32 (permlane32 swap), l and l^16 (permlane16 swap), the two halves of
// each 16-lane row (DPP row_mirror), then within 8 lanes (DPP).
__device__ __forceinline__ uint32_t reduce_rows8(uint32_t* v) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // lanes < 32 end with row k, lanes >= 32 with row k + 4
    const auto p = __builtin_amdgcn_permlane32_swap(v[k], v[k + 4], false, false);
    v[k] = p[0] + p[1];
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    // even 16-lane rows end with k, odd ones with k + 2
    const auto p = __builtin_amdgcn_permlane16_swap(v[k], v[k + 2], false, false);
    v[k] = p[0] + p[1];
  }
  const bool up = (lane & 8) != 0;
  const uint32_t keep = up ? v[1] : v[0];
  const uint32_t send = up ? v[0] : v[1];
  uint32_t t = keep + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x140, 0xF, 0xF, false);  // row_mirror
  t += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0x141, 0xF, 0xF, false);  // row_half_mirror
  t += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  t += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  return t;
}

// NQ vector groups of 2 consecutive tiles per lane per step: slots u = 2q + e
// are tiles ps + 128q + 2*lane + e (one 16-byte tile load and one 8-byte
// colind load per group; each wave instruction reads one contiguous block),
// a step is 128*NQ tiles.
template <int NQ>
struct Slots8 {
  uint32_t lo[2 * NQ], hi[2 * NQ];  // rows 0-3 / 4-7 of each tile
  int32_t col[2 * NQ];
};

// Loads the step at chunk-relative position ps; positions are relative to
// the first tile `tl`/`cl` point at, `pend` = end of the chunk's tiles in
// that frame (groups past it are not read; positions before the chunk start
// are read and masked later).
template <int NQ>
__device__ __forceinline__ void load8(const uint64_t* __restrict__ tl, const int32_t* __restrict__ cl, int32_t ps,
                                      int32_t pend, int lane, Slots8<NQ>& s) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef int i32x2 __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int32_t g = ps + q * 128 + 2 * lane;
    if (g + 2 <= pend) {
      const u32x4 v = __builtin_nontemporal_load((const u32x4*)(tl + g));
      const i32x2 cv = __builtin_nontemporal_load((const i32x2*)(cl + g));
      s.lo[2 * q] = v.x;
      s.hi[2 * q] = v.y;
      s.lo[2 * q + 1] = v.z;
      s.hi[2 * q + 1] = v.w;
      s.col[2 * q] = cv.x;
      s.col[2 * q + 1] = cv.y;
    } else {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const bool ok = g + e < pend;
        const uint64_t w = ok ? tl[g + e] : 0ull;
        s.lo[2 * q + e] = (uint32_t)w;
        s.hi[2 * q + e] = (uint32_t)(w >> 32);
        s.col[2 * q + e] = ok ? cl[g + e] : 0;
      }
    }
  }
}

// Steps are processed one after another, without a prefetch buffer: with
// the default chunk of one step, a wave issues all of its tile and colind
// loads before the plan entry and the rowptr window, and the registers a
// second buffer would take go to occupancy instead.
template <int NQ>
__global__ __launch_bounds__(256) void k_count8(const SpmvArgs a) {
  constexpr int32_t STEP = 128 * NQ;
  const int lane = lane_id();
  const int c = wave_index();
  if (c >= a.nchunks) return;  // wave-uniform
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  // frame: positions relative to b0 (a multiple of 4); tiles and rowptr are
  // int32-indexed (T < 2^31), so every position fits 32 bits
  const int32_t b0 = (int32_t)(cb & ~(int64_t)3);
  const int32_t pbeg = (int32_t)(cb - b0), pend = (int32_t)(ce - b0);
  const uint64_t* __restrict__ tl = (const uint64_t*)a.tiles + b0;
  const int32_t* __restrict__ cl = a.colind + b0;
  const uint8_t* __restrict__ xb = (const uint8_t*)a.xbits;
  int32_t* __restrict__ y = (int32_t*)a.out;

  Slots8<NQ> cur;
  load8<NQ>(tl, cl, 0, pend, lane, cur);

  int32_t W0 = a.plan[c];
  int32_t rsl, rel;
  window_load(a, W0, lane, rsl, rel);
  uint64_t work = 0;
  bool more = false;
  auto apply_window = [&]() {
    const int32_t Rl = W0 + lane;
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    more = (__ballot(belong) == ~0ull);
    if (belong && rel == rsl) write_empty_row<8, OP_COUNT>(a, Rl);
    work = __ballot(belong && rel > rsl);
  };
  int32_t R = 0, rs = 0, re = 0;  // current row; [rs, re) clamped to the chunk, frame positions
  bool full = false;
  auto next_row = [&]() -> bool {
    while (!work) {
      if (!more) return false;
      W0 += kWave;
      window_load(a, W0, lane, rsl, rel);
      apply_window();
    }
    const int j = __builtin_ctzll(work);
    work &= work - 1;
    R = W0 + j;
    const int32_t s0 = __builtin_amdgcn_readlane(rsl, j) - b0, e0 = __builtin_amdgcn_readlane(rel, j) - b0;
    full = (s0 >= pbeg) && (e0 <= pend);
    rs = s0 > pbeg ? s0 : pbeg;
    re = e0 < pend ? e0 : pend;
    return true;
  };

  uint32_t cnt[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cnt[k] = 0;
  auto finish_row = [&]() {
    const uint32_t tot = reduce_rows8(cnt);
    const int64_t i = (int64_t)R * 8 + (lane >> 3);
    if ((lane & 7) == 0 && i < a.nrows) {
      if (full) {
        if (a.accumulate) y[i] += (int32_t)tot;
        else y[i] = (int32_t)tot;
      } else if (tot) {
        atomicAdd(&y[i], (int32_t)tot);
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) cnt[k] = 0;
  };
  // adds groups of 4 slots to the current row: byte u of word r of the
  // transposed rows is row r of slot u, byte u of xpack is slot u's x word
  auto add_rows = [&](const uint32_t* lo4, const uint32_t* hi4, uint32_t xpack) {
    uint32_t r[8];
    transpose_rows4(lo4[0], lo4[1], lo4[2], lo4[3], r);
    transpose_rows4(hi4[0], hi4[1], hi4[2], hi4[3], r + 4);
#pragma unroll
    for (int k = 0; k < 8; ++k) cnt[k] += (uint32_t)__builtin_popcount(r[k] & xpack);
  };

  apply_window();
  bool open = next_row();  // a row is being accumulated
  for (int32_t ps = 0; open && ps < pend; ps += STEP) {
    if (ps > 0) load8<NQ>(tl, cl, ps, pend, lane, cur);
    uint32_t xpack[NQ / 2];
#pragma unroll
    for (int h = 0; h < NQ / 2; ++h) {
      const uint32_t x0 = xb[(uint32_t)cur.col[4 * h]], x1 = xb[(uint32_t)cur.col[4 * h + 1]];
      const uint32_t x2 = xb[(uint32_t)cur.col[4 * h + 2]], x3 = xb[(uint32_t)cur.col[4 * h + 3]];
      xpack[h] = x0 | (x1 << 8) | (x2 << 16) | (x3 << 24);
    }
    const int32_t tend = ps + STEP < pend ? ps + STEP : pend;
    while (true) {
      const int32_t lo = rs > ps ? rs : ps, hi = re < tend ? re : tend;
      if (lo == ps && hi == ps + STEP) {
#pragma unroll
        for (int h = 0; h < NQ / 2; ++h) add_rows(cur.lo + 4 * h, cur.hi + 4 * h, xpack[h]);  // whole step in this row
      } else {
#pragma unroll
        for (int h = 0; h < NQ / 2; ++h) {
          uint32_t lo4[4], hi4[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int v = 4 * h + u;
            const int32_t p = ps + (v >> 1) * 128 + 2 * lane + (v & 1);
            const bool in = (p >= lo) && (p < hi);
            lo4[u] = in ? cur.lo[v] : 0u;
            hi4[u] = in ? cur.hi[v] : 0u;
          }
          add_rows(lo4, hi4, xpack[h]);
        }
      }
      if (re > tend) break;  // the row goes on past this step
      finish_row();
      open = next_row();
      if (!open || rs >= tend) break;
    }
  }
  if (open) finish_row();  // a row that continues past the chunk: atomic partial
}

// Row l / 4 * 4 + l % 4 of a block row whose lane j (j < 16) holds the totals
// of rows 4j..4j+3 in c[0..3]: lane l reads c[l % 4] of lane l / 4.  The
// reads are unconditional (every lane takes part in the shuffles).
__device__ __forceinline__ uint32_t row_of_quad(const uint32_t* c, int lane) {
  const int src = (lane >> 2) & 15;
  const uint32_t v0 = (uint32_t)__shfl((int)c[0], src), v1 = (uint32_t)__shfl((int)c[1], src);
  const uint32_t v2 = (uint32_t)__shfl((int)c[2], src), v3 = (uint32_t)__shfl((int)c[3], src);
  const int k = lane & 3;
  return k == 0 ? v0 : k == 1 ? v1 : k == 2 ? v2 : v3;
}

// bs 16 count (plan variant 3) with four lanes per tile: lane l takes tile
// slot l / 4 of a step (16 tiles, one 512-byte contiguous read per wave
// instruction) and rows 4q..4q+3 (q = l % 4) of it as one 8-byte word; x is
// replicated into the four 16-bit fields and each field's popcount added to
// the lane's row counter, which are summed over the 16 slots at the end of a
// block row.  The row totals are moved to lanes 0..15 before they are
// written: four strided store or atomic instructions per block row had
// tripled the written bytes (split rows' atomics are not coalesced).
// Against the row-group kernel (16 lanes of 2 bytes per tile): R-MAT s22
// 0.79-0.80 ms against 0.94-0.96, same counts (0.88 before the transposed
// write); two lanes per tile (8 counters per lane) took 0.996.
constexpr int kCount16Step = 4;  // tiles per lane in flight
__global__ __launch_bounds__(256) void k_count16q(const SpmvArgs a) {
  constexpr int BS = 16, NU = kCount16Step, SL = kWave / 4;  // tile slots per step
  const int lane = lane_id();
  const int slot = lane >> 2, q = lane & 3;
  const int c = wave_index();
  if (c >= a.nchunks) return;  // wave-uniform
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const uint64_t* __restrict__ tw = (const uint64_t*)a.tiles;  // 4 words per tile
  const uint16_t* __restrict__ xb = (const uint16_t*)a.xbits;
  const int32_t* __restrict__ colind = a.colind;
  int32_t* __restrict__ y = (int32_t*)a.out;
  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    if (belong && rel == rsl) write_empty_row<BS, OP_COUNT>(a, Rl);
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int32_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      uint32_t cnt[4] = {0u, 0u, 0u, 0u};
      for (int64_t t0 = s; t0 < te; t0 += (int64_t)SL * NU) {
        uint64_t w[NU];
        int32_t cc[NU];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const int64_t t = t0 + (int64_t)u * SL + slot;
          const bool ok = t < te;
          const int64_t tc = ok ? t : s;  // in range: the row's first tile
          cc[u] = colind[tc];
          const uint64_t wv = __builtin_nontemporal_load(&tw[tc * 4 + q]);
          w[u] = ok ? wv : 0ull;
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const uint64_t xx = (uint64_t)xb[(uint32_t)cc[u]] * 0x0001000100010001ull;
          const uint64_t m = w[u] & xx;
          cnt[0] += (uint32_t)__popc((uint32_t)m & 0xFFFFu);
          cnt[1] += (uint32_t)__popc((uint32_t)m >> 16);
          cnt[2] += (uint32_t)__popc((uint32_t)(m >> 32) & 0xFFFFu);
          cnt[3] += (uint32_t)__popc((uint32_t)(m >> 48));
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        for (int off = 4; off < kWave; off <<= 1) cnt[k] += (uint32_t)__shfl_xor((int)cnt[k], off);
      // every lane with q = l % 4 now holds rows 4q..4q+3; lane l < 16 takes
      // row l from lane l / 4, so one store (or atomic) instruction covers
      // the block row's 64 contiguous bytes
      const int32_t v = (int32_t)row_of_quad(cnt, lane);
      const int64_t i = (int64_t)R * BS + lane;
      if (lane < BS && i < a.nrows) {
        if (full) {
          if (a.accumulate) y[i] += v;
          else y[i] = v;
        } else if (v) {
          atomicAdd(&y[i], v);
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// bs 32 count with eight lanes per tile: lane l reads 16 bytes (rows
// 4p..4p+3, p = l % 8) of tile slot l / 8, so one wave instruction reads
// eight whole tiles (1 KB contiguous) and one column-index load and one x
// gather instruction serve them (the row-group kernel needs one of each per
// two tiles); row totals are written as in k_count16q.  R-MAT s22: 1.50 ms
// against 1.64 with two tiles per lane in flight (1.52-1.54 with four, 1.53-
// 1.55 with eight).
constexpr int kCount32Step = 2;  // tiles per lane in flight
__global__ __launch_bounds__(256) void k_count32q(const SpmvArgs a) {
  constexpr int BS = 32, NU = kCount32Step, SL = kWave / 8;  // tile slots per step
  const int lane = lane_id();
  const int slot = lane >> 3, p = lane & 7;
  const int c = wave_index();
  if (c >= a.nchunks) return;  // wave-uniform
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4* __restrict__ tw = (const u32x4*)a.tiles;  // 8 x 16 bytes per tile
  const uint32_t* __restrict__ xb = (const uint32_t*)a.xbits;
  const int32_t* __restrict__ colind = a.colind;
  int32_t* __restrict__ y = (int32_t*)a.out;
  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    if (belong && rel == rsl) write_empty_row<BS, OP_COUNT>(a, Rl);
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int32_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      uint32_t cnt[4] = {0u, 0u, 0u, 0u};
      for (int64_t t0 = s; t0 < te; t0 += (int64_t)SL * NU) {
        u32x4 w[NU];
        int32_t cc[NU];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const int64_t t = t0 + (int64_t)u * SL + slot;
          const bool ok = t < te;
          const int64_t tc = ok ? t : s;  // in range: the row's first tile
          cc[u] = colind[tc];
          const u32x4 wv = __builtin_nontemporal_load(&tw[tc * 8 + p]);
          w[u] = ok ? wv : (u32x4){0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const uint32_t x = xb[(uint32_t)cc[u]];
          cnt[0] += (uint32_t)__popc(w[u].x & x);
          cnt[1] += (uint32_t)__popc(w[u].y & x);
          cnt[2] += (uint32_t)__popc(w[u].z & x);
          cnt[3] += (uint32_t)__popc(w[u].w & x);
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        for (int off = 8; off < kWave; off <<= 1) cnt[k] += (uint32_t)__shfl_xor((int)cnt[k], off);
      // every lane with p = l % 8 now holds rows 4p..4p+3; lane l < 32 takes
      // row l from lane l / 4 (one contiguous store or atomic instruction)
      const int32_t v = (int32_t)row_of_quad(cnt, lane);
      const int64_t i = (int64_t)R * BS + lane;
      if (lane < BS && i < a.nrows) {
        if (full) {
          if (a.accumulate) y[i] += v;
          else y[i] = v;
        } else if (v) {
          atomicAdd(&y[i], v);
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// Value products (min / min-plus / FP32 sum) with four rows per lane: lane l
// holds rows 4p..4p+3 (p = l % (bs/4)) of tile slot l / (bs/4), so a wave
// instruction loads 64/(bs/4) whole tiles (bs 16: 16, 32: 8, 64: 4) and a gather
// pass issues one gather instruction per four tiles per lane instead of one
// per tile slot of the row-group kernel.  Per pass each lane takes one set
// bit of each of its NU tiles (rows from the last to the first, each row's
// bits from the last column back, so a weighted tile's weight ranks count
// down from the lane's last bit); the four row accumulators are summed over
// the tile slots at the end of a block row (FP64 butterfly for the sum: a
// fixed order) and moved to one lane per row for the write.
// Used at bs 32 (R-MAT s22 against the row-group kernel: weighted min-plus
// 4.51 -> 2.83 ms, unit 3.88 -> 2.89 ms, sum 2.70 -> 2.57 ms, min 2.36 ->
// 2.29 ms).  Tiles per lane in flight: four for min; two for min-plus and
// the sum, whose extra registers (weights, FP64 accumulators) cost waves at
// four (weighted min-plus 3.85, sum 2.77 ms) and whose single tile lost
// overlap (3.37, 3.08 ms).  At bs 16 the set-bit list kernel is faster for
// every op (min 1.05 against 1.49 ms).  At bs 64 (sixteen lanes per tile)
// only the weighted min-plus gains (6.77 -> 5.60 ms); min, the sum and the
// unit-weight min-plus are faster on the row-group kernel, whose per-tile
// reads are scalar there.
template <int BS, int OP>
__global__ __launch_bounds__(256) void k_value_q(const SpmvArgs a) {
  static_assert(BS == 16 || BS == 32 || BS == 64, "four rows per lane");
  using RW = typename std::conditional<BS == 64, uint64_t, uint32_t>::type;  // row word
  static_assert(OP == OP_MIN || OP == OP_MINPLUS || OP == OP_SUM, "min / min-plus / sum only");
  constexpr bool PLUS = OP == OP_MINPLUS;
  constexpr bool SUM = OP == OP_SUM;
  using Acc = typename std::conditional<SUM, double, uint32_t>::type;
  constexpr int LPT = BS / 4, SL = kWave / LPT, NU = OP == OP_MIN ? 4 : 2;
  constexpr uint32_t kFlip = 0x80000000u;
  const Acc ident = SUM ? (Acc)0 : (Acc)0xFFFFFFFFu;
  const int lane = lane_id();
  const int slot = lane / LPT, p = lane % LPT;
  const int c = wave_index();
  if (c >= a.nchunks) return;  // wave-uniform
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const int32_t* __restrict__ colind = a.colind;
  const uint32_t* __restrict__ xv = (const uint32_t*)a.xval;
  const bool weighted = PLUS && a.w != nullptr;
  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    if (belong && rel == rsl) write_empty_row<BS, OP>(a, Rl);
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int32_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      Acc acc[4] = {ident, ident, ident, ident};
      for (int64_t t0 = s; t0 < te; t0 += (int64_t)SL * NU) {
        int32_t cc[NU];
        bool act[NU];
        RW rw[NU][4];
        int rank[NU];
        int64_t wo[NU];
        // three stages over the NU tiles so each stage's loads are in flight
        // together: column indices (slots past the row read its first tile),
        // activity bits (min-plus), then the words and weight offsets of the
        // active tiles
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const int64_t t = t0 + (int64_t)u * SL + slot;
          cc[u] = colind[t < te ? t : s];
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const int64_t t = t0 + (int64_t)u * SL + slot;
          bool on = t < te;
          if constexpr (PLUS) {
            if (a.active) on = on && ((a.active[cc[u] >> 6] >> (cc[u] & 63)) & 1ull);
          }
          act[u] = on;
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const int64_t t = t0 + (int64_t)u * SL + slot;
          const int64_t tc = t < te ? t : s;
          RW r4[4] = {0, 0, 0, 0};
          wo[u] = 0;
          if (!PLUS || act[u]) {
            if constexpr (BS == 64) {
              const uint64_t* tp = (const uint64_t*)a.tiles + tc * 64 + 4 * p;
              r4[0] = tp[0]; r4[1] = tp[1]; r4[2] = tp[2]; r4[3] = tp[3];
            } else if constexpr (BS == 32) {
              typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
              const u32x4 v = ((const u32x4*)a.tiles)[tc * 8 + p];
              r4[0] = v.x; r4[1] = v.y; r4[2] = v.z; r4[3] = v.w;
            } else {
              const uint64_t v = ((const uint64_t*)a.tiles)[tc * 4 + p];
              r4[0] = (uint32_t)v & 0xFFFFu;
              r4[1] = (uint32_t)v >> 16;
              r4[2] = (uint32_t)(v >> 32) & 0xFFFFu;
              r4[3] = (uint32_t)(v >> 48);
            }
            if (weighted) wo[u] = a.woff[tc];
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) rw[u][k] = act[u] ? r4[k] : (RW)0;
        }
        if (weighted) {
          // weight rank of the lane's last set bit: set bits of the tile's
          // earlier rows (lanes p' < p of the slot) plus the lane's own
#pragma unroll
          for (int u = 0; u < NU; ++u) {
            const int pc = popcw(rw[u][0]) + popcw(rw[u][1]) + popcw(rw[u][2]) + popcw(rw[u][3]);
            int incl = pc;
#pragma unroll
            for (int off = 1; off < LPT; off <<= 1) {
              const int o = __shfl_up(incl, off, LPT);
              if (p >= off) incl += o;
            }
            rank[u] = incl - 1;
          }
        }
        while (true) {
          bool any = false;
#pragma unroll
          for (int u = 0; u < NU; ++u) any |= (rw[u][0] | rw[u][1] | rw[u][2] | rw[u][3]) != (RW)0;
          if (!__ballot(any)) break;  // wave-uniform
          // one set bit of each tile; all gathers of the pass are issued
          // before any is used
          uint32_t gv[NU], wt[NU];
          int kk[NU];
          bool has[NU];
#pragma unroll
          for (int u = 0; u < NU; ++u) {
            const RW w0 = rw[u][0], w1 = rw[u][1], w2 = rw[u][2], w3 = rw[u][3];
            const int k = w3 ? 3 : w2 ? 2 : w1 ? 1 : 0;
            const RW wk = k == 3 ? w3 : k == 2 ? w2 : k == 1 ? w1 : w0;
            has[u] = wk != (RW)0;
            kk[u] = k;
            gv[u] = 0u;
            wt[u] = 1u;
            if (has[u]) {
              const int b = BS == 64 ? __builtin_ctzll((uint64_t)wk) : __builtin_ctz((uint32_t)wk);
              gv[u] = xv[(int64_t)cc[u] * BS + (BS - 1 - b)];
              if (weighted) wt[u] = a.w[wo[u] + rank[u]];
              const RW nw = wk & (wk - (RW)1);
#pragma unroll
              for (int q = 0; q < 4; ++q) rw[u][q] = q == k ? nw : rw[u][q];
              --rank[u];
            }
          }
#pragma unroll
          for (int u = 0; u < NU; ++u) {
            if (!has[u]) continue;
            Acc v;
            if constexpr (SUM) v = (double)__uint_as_float(gv[u]);
            else if constexpr (PLUS) v = (gv[u] > kInfU32 - wt[u]) ? kInfU32 : gv[u] + wt[u];
            else v = gv[u] ^ kFlip;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if (q == kk[u]) {
                if constexpr (SUM) acc[q] += v;
                else acc[q] = v < acc[q] ? v : acc[q];
              }
            }
          }
        }
      }
      // totals over the tile slots, then row l of the block row to lane l
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        for (int off = LPT; off < kWave; off <<= 1) {
          const Acc o = __shfl_xor(acc[k], off);
          if constexpr (SUM) acc[k] += o;
          else acc[k] = o < acc[k] ? o : acc[k];
        }
      }
      const int src = (lane >> 2) & (LPT - 1);
      Acc v0 = __shfl(acc[0], src), v1 = __shfl(acc[1], src), v2 = __shfl(acc[2], src), v3 = __shfl(acc[3], src);
      const int kq = lane & 3;
      const Acc m = kq == 0 ? v0 : kq == 1 ? v1 : kq == 2 ? v2 : v3;
      const int64_t i = (int64_t)R * BS + lane;
      if (lane < BS && i < a.nrows) {
        if constexpr (SUM) {
          if (full) ((float*)a.out)[i] = (float)((double)a.scale * m + (double)a.bias);
          else a.partial[((int64_t)c * 2 + (rs < cb ? 0 : 1)) * BS + lane] = m;
        } else if constexpr (PLUS) {
          uint32_t* out = (uint32_t*)a.out;
          if (full) {
            const uint32_t dd = xv[i];
            out[i] = m < dd ? m : dd;
          } else if (m != kInfU32) {
            atomicMin(&out[i], m);
          }
        } else {
          int32_t* y = (int32_t*)a.out;
          const int32_t val = (int32_t)(m ^ kFlip);
          if (full) y[i] = val;
          else if (val != INT32_MAX) atomicMin(&y[i], val);
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// bs 8/16 min / min-plus, lane per tile (plan variant 3).  A wave takes its
// chunk 64 block rows (a window) at a time and the window's tiles 64 per
// round (one per lane: more tiles per lane made the kernel register-bound
// and slower), lane per tile across row boundaries; a 6-step shuffle search
// over the window's row ends gives each tile's row.  A wave prefix sum of
// the tiles' popcounts lists every set bit (target vertex, source vertex,
// weight index) in a wave-private LDS list, so the value gathers are spread
// evenly over the lanes (four per lane, issued together) however the bits
// fall in the tiles, instead of one dependent pass per set bit of each
// lane's row word.  Each gathered candidate is folded into its vertex with
// an LDS atomic min (signed values as x ^ 0x80000000, an order-preserving
// map to unsigned); rows wholly inside the chunk are stored, split rows
// combined with global atomics after the prefill, as in k_spmv.  The FP32
// sum folds the same way into FP64 accumulators with LDS atomic adds: the
// accumulators are wave-private, so no other wave interleaves its adds and
// the one wave's adds run in program order, each instruction's same-address
// lanes serialised by the LDS in a fixed order; split rows leave FP64
// partials to k_sum_fixup, as in k_spmv.
constexpr int kValCap = 4 * kWave;  // list entries per wave and pass

template <int BS, int OP>
__global__ __launch_bounds__(256) void k_value_lpt(const SpmvArgs a) {
  static_assert(BS == 8 || BS == 16, "lane-per-tile tiles of 8 or 32 bytes");
  static_assert(OP == OP_MIN || OP == OP_MINPLUS || OP == OP_SUM, "min / min-plus / sum only");
  constexpr bool PLUS = OP == OP_MINPLUS;
  constexpr bool SUM = OP == OP_SUM;
  using Acc = typename std::conditional<SUM, double, uint32_t>::type;
  constexpr int NWD = BS == 8 ? 1 : 4;  // 64-bit words per tile
  constexpr uint32_t kFlip = 0x80000000u;
  // block rows per window: the window's accumulators are WIN * BS values in
  // LDS per wave.  At bs 16 a 64-row window made the sum's workgroup hold
  // 38 KB of LDS (4 waves per SIMD); 32 rows take it to 22.5 KB (7 waves):
  // R-MAT s22 sum 1.38 -> 1.09 ms, weighted min-plus 1.47 -> 1.40 ms.  At
  // bs 8 smaller windows were no faster.
  constexpr int WIN = BS == 16 ? 32 : kWave;
  __shared__ uint32_t s_idx[4][kValCap];
  __shared__ uint32_t s_widx[4][PLUS ? kValCap : 1];
  __shared__ uint16_t s_dst[4][kValCap];
  __shared__ Acc s_acc[4][WIN * BS];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int c = wave_index();
  if (c >= a.nchunks) return;  // wave-uniform; no workgroup barriers below
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const int32_t* __restrict__ colind = a.colind;
  const uint32_t* __restrict__ xv = (const uint32_t*)a.xval;
  const bool weighted = PLUS && a.w != nullptr;
  uint32_t* idx = s_idx[wv];
  uint32_t* widx = s_widx[wv];
  uint16_t* dst = s_dst[wv];
  Acc* acc = s_acc[wv];

  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = lane < WIN && (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const int nb = __popcll(__ballot(belong));  // rows R0 .. R0+nb-1 (a prefix)
    if (nb == 0) break;
    if (belong && rel == rsl) write_empty_row<BS, OP>(a, Rl);
    const int32_t relx = belong ? rel : INT32_MAX;
    const int64_t w0 = max((int64_t)__builtin_amdgcn_readlane(rsl, 0), cb);
    const int64_t w1 = min((int64_t)__builtin_amdgcn_readlane(rel, nb - 1), ce);
    for (int k = lane; k < WIN * BS; k += kWave) acc[k] = SUM ? (Acc)0 : (Acc)0xFFFFFFFFu;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int64_t t0 = w0; t0 < w1; t0 += kWave) {
      const int64_t t = t0 + lane;
      uint64_t w[NWD];
#pragma unroll
      for (int q = 0; q < NWD; ++q) w[q] = 0;
      uint32_t cv = 0, wb = 0;
      if (t < w1) {
        const int32_t col = colind[t];
        cv = (uint32_t)col * BS;
        // inactive source block column (min-plus): the tile and its weight
        // offset are not read
        if (!(PLUS && a.active) || ((a.active[col >> 6] >> (col & 63)) & 1ull)) {
          if constexpr (BS == 8) {
            w[0] = ((const uint64_t*)a.tiles)[t];
          } else {  // word q holds rows 4q..4q+3 as 16-bit fields
            const uint4* tp = (const uint4*)a.tiles + 2 * t;
            const uint4 lo = tp[0], hi = tp[1];
            w[0] = lo.x | ((uint64_t)lo.y << 32);
            w[1] = lo.z | ((uint64_t)lo.w << 32);
            w[2] = hi.x | ((uint64_t)hi.y << 32);
            w[3] = hi.z | ((uint64_t)hi.w << 32);
          }
          if (weighted) wb = (uint32_t)a.woff[t];
        }
      }
      // window row of tile t = the number of window rows that end at or
      // before it (relx is nondecreasing over the lanes).  Usually few rows
      // end inside the round: the rows ended before it are one ballot, and
      // each row end inside it one lane read; with more than 8, a 6-step
      // shuffle search (all lanes take part).
      const int64_t tc = t < w1 ? t : w1 - 1;
      const uint64_t inner = __ballot((int64_t)relx > t0 && (int64_t)relx < t0 + kWave);
      int j = 0;
      if (__popcll(inner) <= 8) {
        j = __popcll(__ballot((int64_t)relx <= t0));
        for (uint64_t m = inner; m; m &= m - 1) j += tc >= (int64_t)__builtin_amdgcn_readlane(relx, __builtin_ctzll(m));
      } else {
#pragma unroll
        for (int step = kWave / 2; step >= 1; step >>= 1)
          if ((int64_t)__shfl(relx, j + step - 1) <= tc) j += step;
      }
      const int jr = j * BS;
      int cnt = 0;
#pragma unroll
      for (int q = 0; q < NWD; ++q) cnt += __popcll(w[q]);
      const int incl = wave_incl_scan(cnt);  // inclusive prefix sum of the popcounts over the lanes
      const int excl = incl - cnt;
      const int tot = __builtin_amdgcn_readlane(incl, kWave - 1);
      for (int base = 0; base < tot; base += kValCap) {
        int e = excl;
        uint32_t before = 0;  // set bits in the tile's earlier words
#pragma unroll
        for (int q = 0; q < NWD; ++q) {
          uint64_t ww = w[q];
          while (ww) {
            const int b = __builtin_ctzll(ww);
            ww &= ww - 1;
            if (e >= base && e < base + kValCap) {
              // row and bit of the row word (column bs-1-bit) of bit b of word q
              const int fs = BS == 8 ? 3 : 4;        // log2 of the row field width
              const int row = (BS == 8 ? 0 : 4 * q) + (b >> fs);
              const int bit = b & (BS - 1);
              idx[e - base] = cv + (uint32_t)(BS - 1 - bit);
              dst[e - base] = (uint16_t)(jr + row);
              if (weighted) {
                // weights are in tile order: rows above first, then this
                // row's columns left to right
                const int sh = (b >> fs) << fs;
                const uint32_t below = before + (uint32_t)__popcll(w[q] & ((1ull << sh) - 1ull));
                const uint32_t rw = (uint32_t)((w[q] >> sh) & ((1ull << BS) - 1ull));
                widx[e - base] = wb + below + (uint32_t)__popc(rw >> (bit + 1));
              }
            }
            ++e;
          }
          before += (uint32_t)__popcll(w[q]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int nl = tot - base < kValCap ? tot - base : kValCap;
        // All gathers of the pass are issued before any is used: a slot
        // holding list entries (nl is wave-uniform) loads in every lane,
        // lanes past the list reading entry 0 (a pass has nl >= 1); their
        // values are left out when folding.  (Loads and uses inside one
        // per-lane branch made the compiler wait for each gather in turn.)
        constexpr int NQ = kValCap / kWave;
        uint32_t xr[NQ], wr[NQ];
        int d[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          xr[q] = 0;
          wr[q] = 1u;
          d[q] = 0;
          if (q * kWave < nl) {  // wave-uniform
            const int e2 = q * kWave + lane;
            const int ee = e2 < nl ? e2 : 0;
            d[q] = dst[ee];
            xr[q] = xv[idx[ee]];
            if constexpr (PLUS) {
              if (weighted) wr[q] = a.w[widx[ee]];
            }
          }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (q * kWave + lane < nl) {
            if constexpr (SUM) {
              atomicAdd(&acc[d[q]], (double)__uint_as_float(xr[q]));
            } else if constexpr (PLUS) {
              atomicMin(&acc[d[q]], (xr[q] > kInfU32 - wr[q]) ? kInfU32 : xr[q] + wr[q]);
            } else {
              atomicMin(&acc[d[q]], xr[q] ^ kFlip);
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // results of the window's rows that have tiles: lane per (row, tile row)
    for (int p0 = 0; p0 < nb * BS; p0 += kWave) {
      const int p = p0 + lane;
      const int j = p / BS;
      const int32_t rs = __shfl(rsl, j & (kWave - 1)), re = __shfl(rel, j & (kWave - 1));
      const int64_t i = (int64_t)(R0 + j) * BS + (p % BS);
      if (p < nb * BS && re > rs && i < a.nrows) {
        const bool full = (rs >= cb) && (re <= ce);
        const Acc m = acc[p];
        if constexpr (SUM) {
          if (full) ((float*)a.out)[i] = (float)((double)a.scale * m + (double)a.bias);
          else a.partial[((int64_t)c * 2 + (rs < cb ? 0 : 1)) * BS + (p % BS)] = m;
        } else if constexpr (PLUS) {
          uint32_t* out = (uint32_t*)a.out;
          if (full) {
            const uint32_t dd = xv[i];
            out[i] = m < dd ? m : dd;
          } else if (m != kInfU32) {
            atomicMin(&out[i], m);
          }
        } else {
          int32_t* y = (int32_t*)a.out;
          const int32_t val = (int32_t)(m ^ kFlip);
          if (full) y[i] = val;
          else if (val != INT32_MAX) atomicMin(&y[i], val);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (nb < WIN) break;
    R0 += WIN;
  }
}

// Initializes the outputs of split rows (rows that begin before the chunk
// that lists them) to the identity of the combining atomic.
template <int BS, int OP>
__global__ __launch_bounds__(256) void k_prefill_split(const SpmvArgs a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = idx / BS + 1;
  const int r = (int)(idx % BS);
  if (c >= a.nchunks) return;
  const int32_t R = a.plan[c];
  if ((int64_t)a.rowptr[R] >= c * a.chunk) return;
  const int64_t i = (int64_t)R * BS + r;
  if constexpr (OP == OP_COUNT) {
    if (i < a.nrows) ((int32_t*)a.out)[i] = 0;
  } else if constexpr (OP == OP_BFS) {
    if (r == 0) ((word_t<BS>*)a.out)[R] = 0;
  } else if constexpr (OP == OP_MINPLUS) {
    if (i < a.nrows) ((uint32_t*)a.out)[i] = ((const uint32_t*)a.xval)[i];
  } else if constexpr (OP == OP_MIN) {
    if (i < a.nrows) ((int32_t*)a.out)[i] = INT32_MAX;
  }
}

// FP32 sum: the chunk that holds the first tile of a split row adds its own
// partial and those of the following chunks in chunk order.
template <int BS>
__global__ __launch_bounds__(256) void k_sum_fixup(const SpmvArgs a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = idx / BS;
  const int r = (int)(idx % BS);
  if (c + 1 >= a.nchunks) return;
  const int64_t nb = (c + 1) * a.chunk;
  const int32_t R = a.plan[c + 1];
  const int64_t rs = a.rowptr[R], re = a.rowptr[R + 1];
  if (!(rs < nb) || rs < c * a.chunk) return;  // not split here, or not the owner
  // chunks c+1 .. ce-1 hold the rest of the row (a hub row spans hundreds):
  // their partials are loaded 16 at a time, so the chain costs one memory
  // latency per 16 chunks, and added in chunk order
  const int64_t ce = min((int64_t)a.nchunks, (re + a.chunk - 1) / a.chunk);
  double s = a.partial[(c * 2 + 1) * BS + r];
  int64_t c2 = c + 1;
  for (; c2 + 16 <= ce; c2 += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = a.partial[((c2 + u) * 2 + 0) * BS + r];
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
  }
  for (; c2 < ce; ++c2) s += a.partial[(c2 * 2 + 0) * BS + r];
  const int64_t i = (int64_t)R * BS + r;
  if (i < a.nrows) ((float*)a.out)[i] = (float)((double)a.scale * s + (double)a.bias);
}

// Dense BFS step, second half: next[R] = rows of block row R whose
// frontier-neighbour count is nonzero, minus the visited ones; new bits are
// counted into the slots.
template <int BS>
__global__ __launch_bounds__(256) void k_bfs_pack(const int32_t* __restrict__ cnt, const word_t<BS>* __restrict__ vis,
                                                  word_t<BS>* __restrict__ next, int32_t nbr, int32_t nrows,
                                                  unsigned long long* __restrict__ slots) {
  const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t word = 0;
  if (R < nbr) {
    const int4* p = (const int4*)(cnt + R * BS);
#pragma unroll
    for (int h = 0; h < BS / 4; ++h) {
      const int4 v = p[h];
      const int32_t c4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = 4 * h + k;
        if (R * BS + r < nrows && c4[k] > 0) word |= 1ull << (BS - 1 - r);
      }
    }
    word &= ~(uint64_t)vis[R];
    next[R] = (word_t<BS>)word;
  }
  uint64_t added = (uint64_t)__popcll(word);
  for (int off = 1; off < kWave; off <<= 1) added += __shfl_xor(added, off);
  if (lane_id() == 0 && added) atomicAdd(&slots[(R / kWave) & (kCountSlots - 1)], (unsigned long long)added);
}

template <int BS, int OP>
static int spmv_launch_t(const SpmvArgs& a, hipStream_t st) {
  // the streamed and multi-row-per-lane kernels (variants 2 and 3, the bs
  // 32 count / min / min-plus) read 16- or 8-byte vectors of tiles and column
  // indices: they need 16-byte-aligned arrays, otherwise the row-group kernel
  // runs.
  const bool aligned = ((((uintptr_t)a.tiles) | ((uintptr_t)a.colind)) & 15) == 0;
  // variant 3: bs 8 count is k_count8, bs 16 count k_count16q
  const bool lean8 = (BS == 8 && OP == OP_COUNT && a.variant == 3 && aligned);
  const bool quad16 = (BS == 16 && OP == OP_COUNT && a.variant == 3 && aligned);
  const bool stream = (BS <= 16 && OP == OP_COUNT && a.variant == 2 && aligned);
  // variant 3 at bs 8 is the lean count kernel; its BFS step is the
  // lane-per-tile kernel (faster than the row-group one at bs 8)
  const bool lpt = ((BS <= 16 && (OP == OP_COUNT || OP == OP_BFS) && a.variant == 1) ||
                    (BS == 8 && OP == OP_BFS && a.variant == 3)) &&
                   !(OP == OP_BFS && a.sparse_frontier);
  if constexpr (OP == OP_BFS) {
    if (dense_bfs_step(BS, a.variant) && a.bfs_counts && aligned && !a.sparse_frontier) {
      // dense step: the count kernel (k_count8 / k_count16q / k_count32q /
      // k_spmv<64>) gives each row's number of frontier neighbours (every
      // tile is read; no visited-row skip)
      SpmvArgs c = a;
      c.out = a.bfs_counts;
      c.accumulate = 0;
      int rc = spmv_launch_t<BS, OP_COUNT>(c, st);
      if (rc) return rc;
      k_bfs_pack<BS><<<(unsigned)(((int64_t)a.nbr + 255) / 256), 256, 0, st>>>(
          a.bfs_counts, (const word_t<BS>*)a.visited, (word_t<BS>*)a.out, a.nbr, a.nrows, a.count_slots);
      GBO_LAUNCH_CHECK();
      return GBO_OK;
    }
  }
  const bool need_prefill = OP == OP_BFS || OP == OP_MINPLUS || OP == OP_MIN || (OP == OP_COUNT && !a.accumulate);
  if (need_prefill && a.nchunks > 1) {
    const int64_t n = (int64_t)(a.nchunks - 1) * BS;
    k_prefill_split<BS, OP><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a);
    GBO_LAUNCH_CHECK();
  }
  const int64_t blocks = ((int64_t)a.nchunks + 3) / 4;
  if constexpr (BS <= 16 && OP == OP_COUNT) {
    if (quad16) {
      k_count16q<<<(unsigned)blocks, 256, 0, st>>>(a);
    } else if (lean8) {
      // steps of 512 tiles for chunks of 512 or more, else 256
      if constexpr (BS == 8) {
        if (a.chunk >= 512) k_count8<4><<<(unsigned)blocks, 256, 0, st>>>(a);
        else k_count8<2><<<(unsigned)blocks, 256, 0, st>>>(a);
      }
    } else if (stream) {
      k_count_stream<BS><<<(unsigned)blocks, 256, 0, st>>>(a);
    } else if (lpt) {
      k_spmv_lpt<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    } else {
      k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    }
  } else if constexpr (BS <= 16 && OP == OP_BFS) {
    if (lpt) k_spmv_lpt<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    else k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
  } else if constexpr (BS <= 16 && (OP == OP_MIN || OP == OP_MINPLUS || OP == OP_SUM)) {
    // variant 3 (the bs 8/16 default): the set-bit list kernel
    if (a.variant == 3 && aligned) k_value_lpt<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    else k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
  } else if constexpr (BS == 32 && (OP == OP_MIN || OP == OP_MINPLUS || OP == OP_SUM)) {
    if (aligned) k_value_q<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    else k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
  } else if constexpr (BS == 64 && OP == OP_MINPLUS) {
    if (aligned && a.w) k_value_q<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    else k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
  } else if constexpr (BS == 32 && OP == OP_COUNT) {
    if (aligned) k_count32q<<<(unsigned)blocks, 256, 0, st>>>(a);
    else k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
  } else {
    k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
  }
  GBO_LAUNCH_CHECK();
  if (OP == OP_SUM && a.nchunks > 1) {
    const int64_t n = (int64_t)(a.nchunks - 1) * BS;
    k_sum_fixup<BS><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a);
    GBO_LAUNCH_CHECK();
  }
  return GBO_OK;
}

template <int OP>
static int spmv_dispatch_bs(int bs, const SpmvArgs& a, hipStream_t st) {
  switch (bs) {
    case 8: return spmv_launch_t<8, OP>(a, st);
    case 16: return spmv_launch_t<16, OP>(a, st);
    case 32: return spmv_launch_t<32, OP>(a, st);
    case 64: return spmv_launch_t<64, OP>(a, st);
  }
  return set_error(GBO_EUNSUPPORTED, "block size %d not supported", bs);
}

int spmv_launch(int bs, int op, const SpmvArgs& a, hipStream_t st) {
  switch (op) {
    case OP_COUNT: return spmv_dispatch_bs<OP_COUNT>(bs, a, st);
    case OP_BFS: return spmv_dispatch_bs<OP_BFS>(bs, a, st);
    case OP_MINPLUS: return spmv_dispatch_bs<OP_MINPLUS>(bs, a, st);
    case OP_SUM: return spmv_dispatch_bs<OP_SUM>(bs, a, st);
    case OP_MIN: return spmv_dispatch_bs<OP_MIN>(bs, a, st);
  }
  return set_error(GBO_EINVAL, "unknown op %d", op);
}

static SpmvArgs base_args(const gbo_bsr* A, const PlanView& pv) {
  SpmvArgs a{};
  a.rowptr = A->rowptr;
  a.colind = A->colind;
  a.tiles = A->tiles;
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.nbr = A->nblockrows;
  a.T = A->nblocks;
  a.plan = pv.plan;
  a.chunk = pv.chunk;
  a.nchunks = pv.nchunks;
  a.variant = pv.variant;
  a.scale = 1.0f;
  a.bias = 0.0f;
  return a;
}

// One block: *out = sum of the slots; the slots are cleared for the next use.
__global__ __launch_bounds__(256) void k_count_fold(unsigned long long* slots, int64_t* out) {
  __shared__ unsigned long long sh[256];
  unsigned long long s = 0;
  for (int i = threadIdx.x; i < kCountSlots; i += 256) {
    s += slots[i];
    slots[i] = 0ull;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = (int64_t)sh[0];
}

int count_fold(unsigned long long* slots, int64_t* d_count, hipStream_t st) {
  k_count_fold<<<1, 256, 0, st>>>(slots, d_count);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

// Runs one operation with a plan acquired (and released) around it; `fill`
// sets the operands and may take scratch from the plan view.
template <class Fill>
static int run_op(const gbo_bsr* A, int op, hipStream_t st, Fill fill) {
  PlanGuard g(st);
  int rc = acquire_plan(A, &g.pv, st);
  if (rc) return rc;
  SpmvArgs a = base_args(A, g.pv);
  rc = fill(a, g.pv);
  if (rc == GBO_OK) rc = spmv_launch(A->bs, op, a, st);
  return rc;
}

}  // namespace gbo

using namespace gbo;

extern "C" {

int gbo_bsrbmv_count(const gbo_bsr* A, const void* d_xbits, int32_t* d_y, int accumulate, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_count");
  if (rc) return rc;
  if (A->nrows > 0 && (!d_y || !d_xbits)) return set_error(GBO_EINVAL, "gbo_bsrbmv_count: null vector");
  return run_op(A, OP_COUNT, stream, [&](SpmvArgs& a, const PlanView&) -> int {
    a.xbits = d_xbits;
    a.out = d_y;
    a.accumulate = accumulate ? 1 : 0;
    return GBO_OK;
  });
}

// The frontier-size slots, and at bs 8 under variant 3 the per-row counts
// of the dense step, come from the plan's scratch; the plan is held until
// the fold of the slots is on the stream.
int gbo_bsrbmv_bfs_step(const gbo_bsr* A, const void* d_frontier, const void* d_visited, void* d_next,
                        int64_t* d_count, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_bfs_step");
  if (rc) return rc;
  if (!d_frontier || !d_visited || !d_next || !d_count) return set_error(GBO_EINVAL, "gbo_bsrbmv_bfs_step: null argument");
  PlanGuard g(stream);
  rc = acquire_plan(A, &g.pv, stream);
  if (rc) return rc;
  SpmvArgs a = base_args(A, g.pv);
  a.xbits = d_frontier;
  a.visited = d_visited;
  a.out = d_next;
  a.count_slots = g.pv.slots;
  if (dense_bfs_step(A->bs, a.variant)) a.bfs_counts = g.pv.bfs_counts;
  rc = spmv_launch(A->bs, OP_BFS, a, stream);
  if (rc == GBO_OK) rc = count_fold(g.pv.slots, d_count, stream);
  return rc;
}

int gbo_bsrbmv_minplus(const gbo_bsr* A, const uint32_t* d_w, const int64_t* d_woff, const uint32_t* d_dist,
                       uint32_t* d_out, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_minplus");
  if (rc) return rc;
  if (A->nrows != A->ncols) return set_error(GBO_EINVAL, "gbo_bsrbmv_minplus: matrix must be square");
  if (A->nrows > 0 && (!d_dist || !d_out)) return set_error(GBO_EINVAL, "gbo_bsrbmv_minplus: null vector");
  if ((d_w == nullptr) != (d_woff == nullptr)) return set_error(GBO_EINVAL, "gbo_bsrbmv_minplus: weights need offsets");
  return run_op(A, OP_MINPLUS, stream, [&](SpmvArgs& a, const PlanView&) -> int {
    a.w = d_w;
    a.woff = d_woff;
    a.xval = d_dist;
    a.out = d_out;
    return GBO_OK;
  });
}

// The FP64 split-row partials come from the plan's scratch.
int gbo_bsrbmv_sum_f32(const gbo_bsr* A, const float* d_x, float* d_y, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_sum_f32");
  if (rc) return rc;
  if (A->nrows > 0 && (!d_x || !d_y)) return set_error(GBO_EINVAL, "gbo_bsrbmv_sum_f32: null vector");
  return run_op(A, OP_SUM, stream, [&](SpmvArgs& a, const PlanView& pv) -> int {
    a.xval = d_x;
    a.out = d_y;
    a.partial = pv.partial;
    return GBO_OK;
  });
}

int gbo_bsrbmv_min_i32(const gbo_bsr* A, const int32_t* d_x, int32_t* d_y, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_min_i32");
  if (rc) return rc;
  if (A->nrows > 0 && (!d_x || !d_y)) return set_error(GBO_EINVAL, "gbo_bsrbmv_min_i32: null vector");
  return run_op(A, OP_MIN, stream, [&](SpmvArgs& a, const PlanView&) -> int {
    a.xval = d_x;
    a.out = d_y;
    return GBO_OK;
  });
}

}  // extern "C"
