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
        int acc = 0;
        int64_t t = s + g;
        for (; t + 3 * G < te; t += 4 * G) {
          const int32_t c0 = colind[t], c1 = colind[t + G], c2 = colind[t + 2 * G], c3 = colind[t + 3 * G];
          const W w0 = tiles[t * BS + r], w1 = tiles[(t + G) * BS + r], w2 = tiles[(t + 2 * G) * BS + r],
                  w3 = tiles[(t + 3 * G) * BS + r];
          acc += popcw((W)(w0 & xb[c0])) + popcw((W)(w1 & xb[c1])) + popcw((W)(w2 & xb[c2])) +
                 popcw((W)(w3 & xb[c3]));
        }
        for (; t < te; t += G) acc += popcw((W)(tiles[t * BS + r] & xb[colind[t]]));
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
  const bool need_prefill = (OP == OP_BFS || OP == OP_MINPLUS || OP == OP_MIN || (OP == OP_COUNT && !a.accumulate));
  if (need_prefill && a.nchunks > 1) {
    const int64_t n = (int64_t)(a.nchunks - 1) * BS;
    k_prefill_split<BS, OP><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a);
    GBO_LAUNCH_CHECK();
  }
  const int64_t blocks = ((int64_t)a.nchunks + 3) / 4;
  k_spmv<BS, OP><<<(unsigned)blocks, 256, 0, st>>>(a);
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
