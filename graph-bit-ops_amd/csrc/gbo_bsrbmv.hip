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
#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {

// Neumaier's compensated addition: s + c carries the sum of all added terms
// with the rounding error of each addition accumulated in c.
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float t = s + v;
  if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
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

template <int BS, int OP>
__global__ __launch_bounds__(256) void k_spmv(const SpmvArgs a) {
  using W = word_t<BS>;
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int g = lane / BS, r = lane % BS;
  const int c = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
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
      const int64_t rs = __shfl(rsl, j), re = __shfl(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      const int64_t i = (int64_t)R * BS + r;

      if constexpr (OP == OP_COUNT) {
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
        int hit = 0;
        for (int64_t t = s + g; t < te; t += G) {
          const W f = fb[colind[t]];
          if (f != 0 && !hit) {  // skip tiles whose frontier word is empty
            if ((W)(tiles[t * BS + r] & f) != 0) hit = 1;
          }
        }
        for (int off = BS; off < kWave; off <<= 1) hit |= __shfl_xor(hit, off);
        const uint64_t bal = __ballot(hit);
        const uint64_t word = brev_low<BS>(bal & low_mask<BS>()) & ~(uint64_t)vword & low_mask<BS>();
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
        for (int64_t t = s + g; t < te; t += G) {
          const int32_t col = colind[t];
          if (a.active && !((a.active[col >> 6] >> (col & 63)) & 1ull)) continue;  // group-uniform
          uint64_t wd = (uint64_t)tiles[t * BS + r];
          const int pc = popcw(wd);
          int rank0 = 0;
          if (a.w) {
            int incl = pc;
            for (int off = 1; off < BS; off <<= 1) {
              const int o = __shfl_up(incl, off, BS);
              if (r >= off) incl += o;
            }
            rank0 = incl - pc;
          }
          int k = 0;
          while (wd) {
            const int p = __builtin_ctzll(wd);
            const uint32_t dv = dist[(int64_t)col * BS + (BS - 1 - p)];
            const uint32_t wt = a.w ? a.w[a.woff[t] + rank0 + (pc - 1 - k)] : 1u;
            const uint32_t cand = (dv > kInfU32 - wt) ? kInfU32 : dv + wt;
            acc = cand < acc ? cand : acc;
            wd &= wd - 1;
            ++k;
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
        // Compensated (Neumaier) FP32 summation: the error of a row sum stays
        // within a few ulps of the result whatever the row length (a hub
        // row adds >10^5 terms).
        const float* __restrict__ x = (const float*)a.xval;
        float acc = 0.0f, comp = 0.0f;
        for (int64_t t = s + g; t < te; t += G) {
          const int32_t col = colind[t];
          uint64_t wd = (uint64_t)tiles[t * BS + r];
          while (wd) {
            const int p = __builtin_ctzll(wd);
            kahan_add(acc, comp, x[(int64_t)col * BS + (BS - 1 - p)]);
            wd &= wd - 1;
          }
        }
        for (int off = BS; off < kWave; off <<= 1) {
          const float os = __shfl_xor(acc, off), oc = __shfl_xor(comp, off);
          kahan_add(acc, comp, os);
          comp += oc;
        }
        if (g == 0 && i < a.nrows) {
          if (full) {
            ((float*)a.out)[i] = a.scale * (acc + comp) + a.bias;
          } else {
            const int slot = (rs < cb) ? 0 : 1;
            a.partial[((int64_t)c * 2 + slot) * BS + r] = acc + comp;
          }
        }
      } else if constexpr (OP == OP_MIN) {
        const int32_t* __restrict__ x = (const int32_t*)a.xval;
        int32_t acc = INT32_MAX;
        for (int64_t t = s + g; t < te; t += G) {
          const int32_t col = colind[t];
          uint64_t wd = (uint64_t)tiles[t * BS + r];
          while (wd) {
            const int p = __builtin_ctzll(wd);
            const int32_t v = x[(int64_t)col * BS + (BS - 1 - p)];
            acc = v < acc ? v : acc;
            wd &= wd - 1;
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
    if (lane == 0 && added) atomicAdd((unsigned long long*)a.count, (unsigned long long)added);
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
  const int c = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
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
      const int64_t rs = __shfl(rsl, j), re = __shfl(rel, j);
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
        // butterfly reduce-scatter: after the loop lane l holds a partial
        // of row l / LPR; then sum the LPR lanes of that row
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
        for (int64_t t = s + lane; t < te; t += kWave) {
          const uint64_t f = xb[colind[t]];
          if (f) {  // tiles whose frontier word is empty are not read
            uint64_t w[NW];
            load_tile<BS>(a.tiles, t, w);
            const uint64_t xr = (BS == 8) ? f * 0x0101010101010101ull : f * 0x0001000100010001ull;
#pragma unroll
            for (int k = 0; k < NW; ++k) hit[k] |= w[k] & xr;
          }
        }
#pragma unroll
        for (int k = 0; k < NW; ++k)
          for (int off = 1; off < kWave; off <<= 1) hit[k] |= (uint64_t)__shfl_xor((long long)hit[k], off);
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
    if (lane == 0 && added) atomicAdd((unsigned long long*)a.count, (unsigned long long)added);
  }
}

// Row outputs that the segmented kernel never writes: empty rows get the
// identity; rows cut by a chunk boundary are initialised for the atomics.
template <int BS, int OP>
__global__ __launch_bounds__(256) void k_prefill_rows(const SpmvArgs a) {
  const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (R >= a.nbr) return;
  const int64_t rs = a.rowptr[R], re = a.rowptr[R + 1];
  const bool empty = rs == re;
  const bool split = !empty && (rs / a.chunk != (re - 1) / a.chunk);
  if (!(empty || split)) return;
  if constexpr (OP == OP_COUNT) {
    if (!a.accumulate)
      for (int r = 0; r < BS && R * BS + r < a.nrows; ++r) ((int32_t*)a.out)[R * BS + r] = 0;
  } else if constexpr (OP == OP_BFS) {
    ((word_t<BS>*)a.out)[R] = (word_t<BS>)0;
  }
}

// Segmented lane-per-tile kernel for bs 8 and 16 (COUNT).  Every wave step
// takes the next 64 tiles of its chunk whatever rows they belong to: lane l
// finds the block row of its tile by a binary search (ds_bpermute) over a
// window of 64 rowptr entries, per-row popcounts are formed SWAR-style on
// the 64-bit tile words, and a segmented inclusive scan over the lanes adds
// the tiles of each row; the last lane of a row segment writes the row (or
// carries it into the next step when the row continues).  Loads of step
// k+1 are issued before step k is reduced.
template <int BS>
__device__ __forceinline__ void unpack_counts(const uint64_t* v, uint32_t* cnt) {
  if constexpr (BS == 8) {
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      cnt[2 * f] = (uint32_t)(v[0] >> (16 * f)) & 0xFFFFu;
      cnt[2 * f + 1] = (uint32_t)(v[1] >> (16 * f)) & 0xFFFFu;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int f = 0; f < 4; ++f) cnt[4 * k + f] = (uint32_t)(v[k] >> (16 * f)) & 0xFFFFu;
  }
}

template <int BS>
__global__ __launch_bounds__(256) void k_count_seg(const SpmvArgs a) {
  using W = word_t<BS>;
  constexpr int NW = BS * BS / 64;      // 64-bit words per tile
  constexpr int NA = (BS == 8) ? 2 : 4;  // packed 16-bit-field sums
  const int lane = lane_id();
  const int c = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  if (c >= a.nchunks) return;
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  if (cb >= ce) return;
  const int32_t* __restrict__ colind = a.colind;
  const W* __restrict__ xb = (const W*)a.xbits;
  int32_t* __restrict__ y = (int32_t*)a.out;

  int32_t R0 = a.plan[c];
  int64_t rsw, wend;
  auto load_window = [&]() {
    rsw = (R0 + lane <= a.nbr) ? (int64_t)a.rowptr[R0 + lane] : INT64_MAX;
    wend = (R0 + kWave <= a.nbr) ? (int64_t)a.rowptr[R0 + kWave] : a.T;
  };
  load_window();

  auto flush_row = [&](int64_t R, const uint32_t* cnt, bool full) {
    for (int r = 0; r < BS; ++r) {
      const int64_t i = R * BS + r;
      if (i >= a.nrows) break;
      if (full) {
        if (a.accumulate) y[i] += (int32_t)cnt[r];
        else y[i] = (int32_t)cnt[r];
      } else if (cnt[r]) {
        atomicAdd(&y[i], (int32_t)cnt[r]);
      }
    }
  };

  uint32_t carry[BS];
#pragma unroll
  for (int r = 0; r < BS; ++r) carry[r] = 0;
  int64_t carry_row = -1;

  int32_t ncol = 0;
  uint64_t nw[NW];
  {
    const int64_t t = cb + lane;
    if (t < ce) {
      ncol = colind[t];
      load_tile<BS>(a.tiles, t, nw);
    }
  }
  for (int64_t t0 = cb; t0 < ce; t0 += kWave) {
    const int64_t t = t0 + lane;
    const bool valid = t < ce;
    const int32_t col = ncol;
    uint64_t w[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = nw[k];
    {
      const int64_t tn = t + kWave;  // prefetch the next step
      if (tn < ce) {
        ncol = colind[tn];
        load_tile<BS>(a.tiles, tn, nw);
      }
    }
    uint64_t m[NA];
    {
      const uint64_t x = valid ? (uint64_t)xb[col] : 0ull;
      uint64_t pf[NW];
      popc_fields<BS>(w, x, pf);
      if constexpr (BS == 8) {
        m[0] = pf[0] & 0x00FF00FF00FF00FFull;
        m[1] = (pf[0] >> 8) & 0x00FF00FF00FF00FFull;
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) m[k] = pf[k];
      }
    }
    bool pending = valid;
    while (__ballot(pending)) {
      const bool inwin = pending && t < wend;
      const uint64_t inmask = __ballot(inwin);
      if (inmask) {
        // last window slot whose row start is <= t
        int pos = 0;
#pragma unroll
        for (int k = 32; k > 0; k >>= 1) {
          const int cand = pos + k;
          const int64_t v = __shfl(rsw, cand & (kWave - 1));
          if (cand < kWave && v <= t) pos = cand;
        }
        const int64_t key = inwin ? (int64_t)R0 + pos : -2 - lane;  // never equals a row or -1
        uint64_t v[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) v[k] = inwin ? m[k] : 0ull;
        // segmented inclusive scan (rows are contiguous in lane order)
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
          const int64_t kl = __shfl_up(key, off);
#pragma unroll
          for (int k = 0; k < NA; ++k) {
            const uint64_t o = (uint64_t)__shfl_up((long long)v[k], off);
            if (lane >= off && kl == key) v[k] += o;
          }
        }
        const int64_t knext = __shfl_down(key, 1);
        const bool seg_end = inwin && (lane == kWave - 1 || knext != key);
        const int64_t rsR = __shfl(rsw, pos);
        const int64_t reR0 = __shfl(rsw, (pos + 1) & (kWave - 1));
        const int64_t reR = (pos == kWave - 1) ? wend : reR0;
        const bool cont = seg_end && (reR > t + 1);  // row goes on past this tile
        const int lastl = 63 - __builtin_clzll(inmask);
        uint32_t cnt[BS];
        unpack_counts<BS>(v, cnt);
        if (key == carry_row) {
#pragma unroll
          for (int r = 0; r < BS; ++r) cnt[r] += carry[r];
        }
        if (seg_end && !cont) flush_row(key, cnt, rsR >= cb && reR <= ce);
        const bool cont_any = __shfl((int)cont, lastl) != 0;
        if (cont_any) {
#pragma unroll
          for (int r = 0; r < BS; ++r) carry[r] = (uint32_t)__shfl((int)cnt[r], lastl);
          carry_row = __shfl(key, lastl);
        } else {
          carry_row = -1;
        }
        pending = pending && !inwin;
      }
      if (__ballot(pending)) {
        R0 += kWave;
        load_window();
      }
    }
  }
  // a row that continues past the chunk: partial, combined atomically
  if (carry_row >= 0 && lane == 0) flush_row(carry_row, carry, false);
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
  float s = a.partial[(c * 2 + 1) * BS + r], comp = 0.0f;
  for (int64_t c2 = c + 1; c2 < a.nchunks && c2 * a.chunk < re; ++c2) kahan_add(s, comp, a.partial[(c2 * 2 + 0) * BS + r]);
  const int64_t i = (int64_t)R * BS + r;
  if (i < a.nrows) ((float*)a.out)[i] = a.scale * (s + comp) + a.bias;
}

template <int BS, int OP>
static int spmv_launch_t(const SpmvArgs& a, hipStream_t st) {
  const bool seg = (BS <= 16 && OP == OP_COUNT && a.variant == 2);
  const bool lpt = (BS <= 16 && (OP == OP_COUNT || OP == OP_BFS) && a.variant >= 1 && !seg);
  const bool need_prefill = !seg && (OP == OP_BFS || OP == OP_MINPLUS || OP == OP_MIN || (OP == OP_COUNT && !a.accumulate));
  if (need_prefill && a.nchunks > 1) {
    const int64_t n = (int64_t)(a.nchunks - 1) * BS;
    k_prefill_split<BS, OP><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a);
    GBO_LAUNCH_CHECK();
  }
  const int64_t blocks = ((int64_t)a.nchunks + 3) / 4;
  if constexpr (BS <= 16 && OP == OP_COUNT) {
    if (seg) {
      k_prefill_rows<BS, OP><<<(unsigned)((a.nbr + 255) / 256), 256, 0, st>>>(a);
      GBO_LAUNCH_CHECK();
      k_count_seg<BS><<<(unsigned)blocks, 256, 0, st>>>(a);
    } else if (lpt) {
      k_spmv_lpt<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    } else {
      k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
    }
  } else if constexpr (BS <= 16 && OP == OP_BFS) {
    if (lpt) k_spmv_lpt<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
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

// Runs one operation with a plan acquired (and released) around it.
template <class Fill>
static int run_op(const gbo_bsr* A, int op, hipStream_t st, Fill fill) {
  PlanView pv;
  int rc = acquire_plan(A, &pv, st);
  if (rc) return rc;
  SpmvArgs a = base_args(A, pv);
  rc = fill(a);
  if (rc == GBO_OK) rc = spmv_launch(A->bs, op, a, st);
  int rc2 = release_plan(&pv, st);
  return rc ? rc : rc2;
}

}  // namespace gbo

using namespace gbo;

extern "C" {

int gbo_bsrbmv_count(const gbo_bsr* A, const void* d_xbits, int32_t* d_y, int accumulate, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_count");
  if (rc) return rc;
  if (A->nrows > 0 && (!d_y || !d_xbits)) return set_error(GBO_EINVAL, "gbo_bsrbmv_count: null vector");
  return run_op(A, OP_COUNT, stream, [&](SpmvArgs& a) -> int {
    a.xbits = d_xbits;
    a.out = d_y;
    a.accumulate = accumulate ? 1 : 0;
    return GBO_OK;
  });
}

int gbo_bsrbmv_bfs_step(const gbo_bsr* A, const void* d_frontier, const void* d_visited, void* d_next,
                        int64_t* d_count, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_bfs_step");
  if (rc) return rc;
  if (!d_frontier || !d_visited || !d_next || !d_count) return set_error(GBO_EINVAL, "gbo_bsrbmv_bfs_step: null argument");
  GBO_HIP_CHECK(hipMemsetAsync(d_count, 0, sizeof(int64_t), stream));
  return run_op(A, OP_BFS, stream, [&](SpmvArgs& a) -> int {
    a.xbits = d_frontier;
    a.visited = d_visited;
    a.out = d_next;
    a.count = d_count;
    return GBO_OK;
  });
}

int gbo_bsrbmv_minplus(const gbo_bsr* A, const uint32_t* d_w, const int64_t* d_woff, const uint32_t* d_dist,
                       uint32_t* d_out, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_minplus");
  if (rc) return rc;
  if (A->nrows != A->ncols) return set_error(GBO_EINVAL, "gbo_bsrbmv_minplus: matrix must be square");
  if (A->nrows > 0 && (!d_dist || !d_out)) return set_error(GBO_EINVAL, "gbo_bsrbmv_minplus: null vector");
  if ((d_w == nullptr) != (d_woff == nullptr)) return set_error(GBO_EINVAL, "gbo_bsrbmv_minplus: weights need offsets");
  return run_op(A, OP_MINPLUS, stream, [&](SpmvArgs& a) -> int {
    a.w = d_w;
    a.woff = d_woff;
    a.xval = d_dist;
    a.out = d_out;
    return GBO_OK;
  });
}

int gbo_bsrbmv_sum_f32(const gbo_bsr* A, const float* d_x, float* d_y, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_sum_f32");
  if (rc) return rc;
  if (A->nrows > 0 && (!d_x || !d_y)) return set_error(GBO_EINVAL, "gbo_bsrbmv_sum_f32: null vector");
  float* partial = nullptr;
  rc = run_op(A, OP_SUM, stream, [&](SpmvArgs& a) -> int {
    GBO_HIP_CHECK(hipMallocAsync((void**)&partial, sizeof(float) * 2 * (size_t)a.nchunks * A->bs, stream));
    a.xval = d_x;
    a.out = d_y;
    a.partial = partial;
    return GBO_OK;
  });
  if (partial) (void)hipFreeAsync(partial, stream);
  return rc;
}

int gbo_bsrbmv_min_i32(const gbo_bsr* A, const int32_t* d_x, int32_t* d_y, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_bsrbmv_min_i32");
  if (rc) return rc;
  if (A->nrows > 0 && (!d_x || !d_y)) return set_error(GBO_EINVAL, "gbo_bsrbmv_min_i32: null vector");
  return run_op(A, OP_MIN, stream, [&](SpmvArgs& a) -> int {
    a.xval = d_x;
    a.out = d_y;
    return GBO_OK;
  });
}

}  // extern "C"
