// Many-graph batch: BFS + connected components of hundreds of independent
// graphs in one persistent launch.  No reference counterpart: the reference
// runs one graph per process (algorithm/bfs.cu, algorithm/cc.cu) with a
// host round trip per iteration.
//
// The graphs are packed side by side as one block-diagonal bit-BSR: graph g
// occupies vertices [vbase_g, vbase_g + n_g) with vbase_g a multiple of bs,
// so every tile (and every block row) belongs to exactly one graph.  One
// persistent kernel, sized so that all of its workgroups are resident at
// once, sweeps the block rows of every graph together: BFS levels and FastSV
// iterations advance in lockstep across the batch, separated by grid-wide
// barriers, so a large graph's rows are spread over the whole GPU instead of
// one workgroup.  Per-graph state keeps each graph's own schedule: a graph
// whose frontier emptied (BFS) or that stopped changing (FastSV) skips its
// rows from then on, exactly where a one-graph run would have stopped.  The
// barrier wait is bounded: a wait past the limit sets an abort flag, every
// workgroup leaves, and the host reports an error.  The launch (state reset
// kernel + persistent kernel) is captured once in a hipGraph and replayed.
#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {
// Grid barrier state and the rotating "anything changed" counters.  The
// arrival counter, the generation the waiters poll and the abort flag sit
// on cache lines of their own, so polling does not contend with arrivals.
struct GridSync {
  alignas(256) unsigned int arrive;
  alignas(256) unsigned int gen;
  alignas(256) unsigned int abort;
  alignas(256) unsigned int cnt[3];
  unsigned int join;     // tickets taken (| closed bit)
  unsigned int members;  // member count, published at the close
};
}  // namespace gbo

struct gbo_batch {
  int bs = 0;
  int32_t ngraphs = 0;
  gbo_bsr A{};                 // union matrix
  int64_t nunion = 0;          // padded vertex count
  std::vector<int64_t> out_off;  // per-graph offset in the outputs (unpadded), ngraphs+1
  // device descriptors
  int32_t* d_rowg = nullptr;   // graph of each block row
  int64_t* d_vbase = nullptr;  // first union vertex
  int32_t* d_n = nullptr;      // vertex count
  int64_t* d_out = nullptr;    // output offset
  int32_t* d_src = nullptr;    // BFS sources (graph-local)
  // scratch over the union
  void* d_f0 = nullptr;        // frontier bits, two levels (nblockrows words each)
  void* d_f1 = nullptr;
  void* d_vis = nullptr;
  int32_t* d_lab = nullptr;    // FastSV parent (union ids)
  int32_t* d_gf = nullptr;
  int32_t* d_act = nullptr;    // per graph: BFS level, then FastSV iteration, of its last change (2 x ngraphs)
  gbo::GridSync* d_sync = nullptr;
  // captured launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t* graph_level = nullptr;
  int32_t* graph_label = nullptr;
  int64_t graph_spin_ms = -1;  // barrier wait limit baked into the captured launch
  int64_t spin_ms = 0;         // limit of the last launch (direct or replayed)
};

namespace gbo {

struct BatchArgs {
  const int32_t* rowptr;
  const int32_t* colind;
  const void* tiles;
  int64_t nblockrows;
  int32_t ngraphs;
  const int32_t* rowg;
  const int64_t* vbase;
  const int32_t* n;
  const int64_t* out;
  const int32_t* src;
  void* f0;
  void* f1;
  void* vis;
  int32_t* lab;
  int32_t* gf;
  int32_t* bfs_act;  // last BFS level at which the graph found vertices
  int32_t* cc_act;   // last FastSV iteration that changed the graph
  int32_t* level;
  int32_t* label;
  GridSync* sync;
  int64_t spin_ticks;  // barrier wait limit, wall-clock ticks
  int64_t join_ticks;  // how long membership stays open, wall-clock ticks
};

constexpr int kBatchThreads = 512;
constexpr int kBatchWaves = kBatchThreads / kWave;
constexpr int kBatchSpinMs = 5000;       // barrier wait limit
constexpr int kBatchJoinUs = 50;         // how long membership stays open
// Block rows per wave task: bs 8/16 take a group of rows as one tile range;
// bs 32/64 rows are long enough alone.  Concurrent tasks then cover a
// window of rows (and so of graphs) small enough for their gathers to stay
// in L2.
template <int BS>
constexpr int row_group() { return BS == 8 ? 16 : BS == 16 ? 8 : 1; }
constexpr unsigned int kJoinClosed = 0x80000000u;

__global__ void k_batch_reset(GridSync* s) {
  s->arrive = 0;
  s->gen = 0;
  s->abort = 0;
  s->cnt[0] = s->cnt[1] = s->cnt[2] = 0;
  s->join = 0;
  s->members = 0;
}

__device__ __forceinline__ void set_abort(GridSync* s) {
  __hip_atomic_store(&s->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Membership.  A workgroup joins by taking a ticket; the holder of ticket 0
// closes the membership once every workgroup of the grid has joined or after
// join_ticks, and publishes the member count.  A workgroup that joins after
// the close was not running alongside the members and leaves at once, so the
// grid barrier only ever waits for workgroups that are resident.
__device__ bool grid_join(GridSync* s, int64_t join_ticks, int64_t spin_ticks, int& part, int& nparts) {
  __shared__ int s_part, s_np;
  if (threadIdx.x == 0) {
    int p = -1, np = 0;
    const unsigned int t = __hip_atomic_fetch_add(&s->join, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (!(t & kJoinClosed)) {
      p = (int)t;
      if (t == 0) {
        const int64_t t0 = wall_clock64();
        while (__hip_atomic_load(&s->join, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gridDim.x &&
               wall_clock64() - t0 < join_ticks)
          __builtin_amdgcn_s_sleep(2);
        const unsigned int n =
            __hip_atomic_fetch_or(&s->join, kJoinClosed, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) & ~kJoinClosed;
        __hip_atomic_store(&s->members, n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        np = (int)n;
      } else {
        const int64_t t0 = wall_clock64();
        while ((np = (int)__hip_atomic_load(&s->members, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
          if (wall_clock64() - t0 > spin_ticks) {
            set_abort(s);
            p = -1;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
    }
    s_part = p;
    s_np = np;
  }
  __syncthreads();
  part = s_part;
  nparts = s_np;
  return part >= 0;
}

// Grid-wide barrier over the members, and counted phases: workgroups that
// saw a change add 1 to cnt[ph % 3], the barrier completes, and everyone
// reads the total.  Slot (ph+1) % 3 is cleared by member 0 during phase ph:
// its last readers finished before the barrier that closed phase ph-1.  Every
// wait is bounded; past the limit the abort flag is set and all members
// leave.
struct GridBarrier {
  GridSync* s;
  int64_t spin_ticks;
  unsigned int nparts;
  bool leader;
  unsigned int gen = 0, ph = 0;

  // The arrival publishes this workgroup's writes with a release fence and a
  // relaxed increment; the last arriver acquires them (fence) and releases
  // the next generation; waiters poll with relaxed loads (an acquiring load
  // in the loop would invalidate caches on every poll) and acquire once the
  // generation has moved.  With acquiring polls, the members that finished
  // a phase early kept invalidating the caches of the ones still working:
  // every phase of the batch ran slower (512 ER graphs: 21 ms -> 13 ms at
  // bs 8 with relaxed polls), and a bare barrier cost 73 us instead of 43.
  __device__ bool sync() {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      const unsigned int prev = __hip_atomic_fetch_add(&s->arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (prev == nparts - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(&s->arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&s->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const int64_t t0 = wall_clock64();
        while (__hip_atomic_load(&s->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
          if (__hip_atomic_load(&s->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
          if (wall_clock64() - t0 > spin_ticks) {
            set_abort(s);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      s_ok = __hip_atomic_load(&s->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    }
    ++gen;
    __syncthreads();
    return s_ok != 0;
  }

  __device__ void open(int* s_any) {
    if (threadIdx.x == 0) {
      *s_any = 0;
      if (leader) __hip_atomic_store(&s->cnt[(ph + 1) % 3], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }

  // 1 if any member saw a change, 0 if none, -1 on abort
  __device__ int close(const int* s_any) {
    __shared__ int s_total;
    __syncthreads();
    if (threadIdx.x == 0 && *s_any)
      __hip_atomic_fetch_add(&s->cnt[ph % 3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!sync()) return -1;
    // the barrier's acquire fence already ordered this read after every add
    if (threadIdx.x == 0) s_total = __hip_atomic_load(&s->cnt[ph % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    ++ph;
    __syncthreads();
    return s_total;
  }
};

// Bit k of the result: byte k of m is nonzero.
__device__ __forceinline__ uint32_t nz_bytes(uint64_t m) {
  const uint64_t t = (((m & 0x7f7f7f7f7f7f7f7full) + 0x7f7f7f7f7f7f7f7full) | m) & 0x8080808080808080ull;
  return (uint32_t)(((t >> 7) * 0x0102040810204080ull) >> 56);
}

// Bit j of the result: 16-bit field j of m is nonzero.
__device__ __forceinline__ uint32_t nz_shorts(uint64_t m) {
  const uint64_t t = (((m & 0x7fff7fff7fff7fffull) + 0x7fff7fff7fff7fffull) | m) & 0x8000800080008000ull;
  return (uint32_t)(((t >> 15) * 0x0001000200040008ull) >> 48);
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  for (int off = kWave / 2; off >= 1; off >>= 1) v |= __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ int64_t readlane64(int64_t v, int j) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), j);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Tile words of tile t as little-endian uint64s: bs 8 -> 1 word (byte r =
// row r), bs 16 -> 4 words (word q holds rows 4q..4q+3 as 16-bit fields).
template <int BS>
__device__ __forceinline__ void tile_words(const void* tiles, int64_t t, uint64_t* w) {
  if constexpr (BS == 8) {
    w[0] = ((const uint64_t*)tiles)[t];
  } else {
    const uint4* p = (const uint4*)tiles + 2 * t;
    const uint4 lo = p[0], hi = p[1];
    w[0] = lo.x | ((uint64_t)lo.y << 32);
    w[1] = lo.z | ((uint64_t)lo.w << 32);
    w[2] = hi.x | ((uint64_t)hi.y << 32);
    w[3] = hi.z | ((uint64_t)hi.w << 32);
  }
}

// Pull-BFS over the block rows R0 .. R0+nrow-1 whose bit is set in amask,
// bs 8/16 (RG*bs = 128 vertices).  The rows' tiles are contiguous and are
// taken lane per tile, 64*NS at a time, across row boundaries: column
// indices are one coalesced load per slot, frontier words one gather, and a
// tile is read only if its frontier word is nonzero.  Returns in gm[0..3]
// (wave-uniform) the 128-bit mask of rows with a frontier neighbour, bit
// j*bs + k = row k of block row R0+j.
template <int BS, int RG>
__device__ __forceinline__ void bfs_group_lpt(const BatchArgs& a, const word_t<BS>* fk, int nrow, int32_t l_rs,
                                              int32_t l_re, uint64_t amask, int lane, uint32_t* gm) {
  static_assert(RG * BS == 128, "group mask is 128 bits");
  constexpr int NS = 4;
  constexpr int NWD = BS / 4 > 1 ? BS / 4 : 1;
  const int32_t* __restrict__ colind = a.colind;
#pragma unroll
  for (int q = 0; q < 4; ++q) gm[q] = 0;
  const int32_t gb = __builtin_amdgcn_readlane(l_rs, 0), ge = __builtin_amdgcn_readlane(l_re, nrow - 1);
  int32_t rend[RG - 1];  // row ends (wave-uniform)
#pragma unroll
  for (int j = 0; j < RG - 1; ++j) rend[j] = j < nrow ? __builtin_amdgcn_readlane(l_re, j) : ge;
  for (int32_t t0 = gb; t0 < ge; t0 += NS * kWave) {
    int32_t c[NS];
    int rr[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int32_t t = t0 + u * kWave + lane;
      int r = 0;
#pragma unroll
      for (int j = 0; j < RG - 1; ++j) r += t >= rend[j];
      rr[u] = r;
      c[u] = (t < ge && ((amask >> r) & 1)) ? colind[t] : -1;
    }
    uint32_t f[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u) f[u] = c[u] >= 0 ? (uint32_t)fk[c[u]] : 0u;
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      if (f[u]) {
        uint64_t w[NWD];
        tile_words<BS>(a.tiles, t0 + u * kWave + lane, w);
        uint32_t m;  // bit k: row k of the tile meets the frontier
        if constexpr (BS == 8) {
          m = nz_bytes(w[0] & (0x0101010101010101ull * f[u]));
        } else {
          const uint64_t rep = 0x0001000100010001ull * f[u];
          m = nz_shorts(w[0] & rep) | (nz_shorts(w[1] & rep) << 4) | (nz_shorts(w[2] & rep) << 8) |
              (nz_shorts(w[3] & rep) << 12);
        }
        const int bit = rr[u] * BS;
        const uint32_t v = m << (bit & 31);
#pragma unroll
        for (int q = 0; q < 4; ++q) gm[q] |= (bit >> 5) == q ? v : 0u;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) gm[q] = wave_or(gm[q]);
}

// Pull-BFS over one block row, bs 32/64: lane per tile for the column
// indices and frontier words, then the tiles with a nonzero frontier word
// are read row per lane (64/bs tiles per step, eight tiles in flight).
template <int BS>
__device__ __forceinline__ uint64_t bfs_row_rl(const BatchArgs& a, const word_t<BS>* fk, int32_t rs, int32_t re,
                                               int lane) {
  using W = word_t<BS>;
  constexpr int G = kWave / BS;
  const int grp = lane / BS, r = lane % BS;
  const int32_t* __restrict__ colind = a.colind;
  const W* __restrict__ tiles = (const W*)a.tiles;
  int hit = 0;
  for (int32_t t0 = rs; t0 < re; t0 += kWave) {
    const int32_t t = t0 + lane;
    const W f = t < re ? fk[colind[t]] : (W)0;
    uint64_t act = __ballot(f != 0);
    while (act) {
      constexpr int D = 8 / G;  // steps in flight: eight tiles per wave
      int jj[D];
#pragma unroll
      for (int u = 0; u < D; ++u) {
        int js = -1;
#pragma unroll
        for (int q = 0; q < G; ++q) {
          const int j = act ? __builtin_ctzll(act) : -1;
          act &= act - 1;
          if (q == grp) js = j;
        }
        jj[u] = js;
      }
      W ww[D], ff[D];
#pragma unroll
      for (int u = 0; u < D; ++u) {
        ww[u] = jj[u] >= 0 ? tiles[(int64_t)(t0 + jj[u]) * BS + r] : (W)0;
        ff[u] = __shfl(f, jj[u] >= 0 ? jj[u] : 0);
      }
#pragma unroll
      for (int u = 0; u < D; ++u) hit |= (W)(ww[u] & ff[u]) != 0;
    }
  }
  return fold_groups<BS>(__ballot(hit));  // row r: any of its lane groups hit
}

constexpr int kListCap = 8 * kWave;  // set-bit list entries per wave and pass

// FastSV hooking of vertex u (a graph in the current iteration) given m, the
// minimum of gf over u's neighbours: lab[lab[u]] and lab[u] drop to m, and
// lab[u] to gf[u] as well.  The sweeps that compute m read only gf, which no
// hook writes, so each vertex hooks as soon as its m is known.  Labels only
// decrease, so a stale read of lab[u] is never below its current value and a
// candidate not below it would leave it unchanged.  A padding vertex past its
// graph's last one has no neighbours and lab = gf = itself, so its hook
// changes nothing.
__device__ __forceinline__ void cc_hook(const BatchArgs& a, int64_t u, int32_t m) {
  const int32_t fu = a.lab[u];
  if (m < fu && m < a.lab[fu]) atomicMin(&a.lab[fu], m);  // skip atomics that cannot lower it
  const int32_t gu = a.gf[u];
  const int32_t cand = m < gu ? m : gu;
  if (cand < fu) atomicMin(&a.lab[u], cand);
}

// Hooks each vertex v of the block rows R0 .. R0+nrow-1 whose bit is set in
// amask with the minimum over the neighbours j of v of gf[j], bs 8/16.  The rows'
// tiles are contiguous and are taken lane per tile, 64*NS at a time, across
// row boundaries.  A wave prefix sum of the tiles' popcounts places every set
// bit (vertex, neighbour) in a wave-private LDS list, so the gathers of gf are
// spread evenly over the lanes (eight per lane in flight) however the bits
// fall in the tiles; each gathered value is folded into its vertex's
// accumulator with an LDS atomic min.
template <int BS, int RG>
__device__ __forceinline__ void cc_group_lpt(const BatchArgs& a, int64_t R0, int nrow, int32_t l_rs, int32_t l_re,
                                             uint64_t amask, int lane, uint32_t* s_idx, uint8_t* s_row,
                                             int32_t* s_acc) {
  constexpr int NS = (BS == 8) ? 4 : 2;
  constexpr int NWD = BS / 4 > 1 ? BS / 4 : 1;
  const int32_t* __restrict__ colind = a.colind;
  for (int i = lane; i < RG * BS; i += kWave) s_acc[i] = INT32_MAX;
  const int32_t gb = __builtin_amdgcn_readlane(l_rs, 0), ge = __builtin_amdgcn_readlane(l_re, nrow - 1);
  int32_t rend[RG - 1];  // row ends (wave-uniform)
#pragma unroll
  for (int j = 0; j < RG - 1; ++j) rend[j] = j < nrow ? __builtin_amdgcn_readlane(l_re, j) : ge;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int32_t t0 = gb; t0 < ge; t0 += NS * kWave) {
    uint32_t cb[NS];
    int rr[NS];
    uint64_t w[NS][NWD];
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int32_t t = t0 + u * kWave + lane;
      cb[u] = 0;
      rr[u] = 0;
#pragma unroll
      for (int q = 0; q < NWD; ++q) w[u][q] = 0;
      if (t < ge) {
        cb[u] = (uint32_t)colind[t] * BS;
        tile_words<BS>(a.tiles, t, w[u]);
        int r = 0;
#pragma unroll
        for (int j = 0; j < RG - 1; ++j) r += t >= rend[j];
        rr[u] = r * BS;
        if (!((amask >> r) & 1)) {
#pragma unroll
          for (int q = 0; q < NWD; ++q) w[u][q] = 0;
        }
      }
#pragma unroll
      for (int q = 0; q < NWD; ++q) cnt += __popcll(w[u][q]);
    }
    int incl = cnt;  // inclusive prefix sum of the popcounts over the lanes
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    const int excl = incl - cnt;
    const int tot = __builtin_amdgcn_readlane(incl, kWave - 1);
    for (int base = 0; base < tot; base += kListCap) {
      int idx = excl;
#pragma unroll
      for (int u = 0; u < NS; ++u) {
#pragma unroll
        for (int q = 0; q < NWD; ++q) {
          uint64_t ww = w[u][q];
          while (ww) {
            const int b = __builtin_ctzll(ww);
            ww &= ww - 1;
            if (idx >= base && idx < base + kListCap) {
              // bs 8: byte b/8 = row, bit 7 - col;  bs 16: word q holds rows 4q..4q+3
              const int row = (BS == 8) ? (b >> 3) : (4 * q + (b >> 4));
              const int col = (BS == 8) ? (7 - (b & 7)) : (15 - (b & 15));
              s_idx[idx - base] = cb[u] + col;
              s_row[idx - base] = (uint8_t)(rr[u] + row);
            }
            ++idx;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const int nl = tot - base < kListCap ? tot - base : kListCap;
      int32_t x[kListCap / kWave];
      int rw[kListCap / kWave];
#pragma unroll
      for (int k = 0; k < kListCap / kWave; ++k) {
        const int e = k * kWave + lane;
        x[k] = INT32_MAX;
        rw[k] = 0;
        if (e < nl) {
          rw[k] = s_row[e];
          x[k] = a.gf[s_idx[e]];
        }
      }
#pragma unroll
      for (int k = 0; k < kListCap / kWave; ++k)
        if (k * kWave + lane < nl) atomicMin(&s_acc[rw[k]], x[k]);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int i = lane; i < nrow * BS; i += kWave)
    if ((amask >> (i / BS)) & 1) cc_hook(a, R0 * BS + i, s_acc[i]);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The same for one block row, bs 32/64: row per lane, four tiles per group per
// round, one set bit of each per pass (the pass's gathers issued together).
template <int BS>
__device__ __forceinline__ void cc_row_rl(const BatchArgs& a, int64_t R, int32_t rs, int32_t re, int lane) {
  using W = word_t<BS>;
  constexpr int G = kWave / BS;
  const int grp = lane / BS, r = lane % BS;
  const int32_t* __restrict__ colind = a.colind;
  const W* __restrict__ tiles = (const W*)a.tiles;
  int32_t acc = INT32_MAX;
  for (int64_t t = rs + grp; t < re; t += 4 * G) {
    int32_t cc[4];
    uint64_t wd[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t tu = t + u * G;
      if constexpr (BS == 32) {
        // unconditional loads (slots past the row read tile rs, then
        // emptied) so the four are in flight together: 512 ER graphs 39 ->
        // 31 ms; at bs 64 the conditional loads measured faster (43 vs 46 ms)
        const int64_t tc = tu < re ? tu : rs;
        cc[u] = colind[tc];
        const uint64_t wv = (uint64_t)tiles[tc * BS + r];
        wd[u] = tu < re ? wv : 0ull;
      } else {
        cc[u] = tu < re ? colind[tu] : 0;
        wd[u] = tu < re ? (uint64_t)tiles[tu * BS + r] : 0ull;
      }
    }
    while (wd[0] | wd[1] | wd[2] | wd[3]) {
      int32_t x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        x[u] = INT32_MAX;
        if (wd[u]) {
          const int p = __builtin_ctzll(wd[u]);
          x[u] = a.gf[(int64_t)cc[u] * BS + (BS - 1 - p)];
          wd[u] &= wd[u] - 1;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = x[u] < acc ? x[u] : acc;
    }
  }
  for (int off = BS; off < kWave; off <<= 1) {
    const int32_t o = __shfl_xor(acc, off);
    acc = o < acc ? o : acc;
  }
  if (grp == 0) cc_hook(a, R * BS + r, acc);
}

template <int BS>
__global__ __launch_bounds__(kBatchThreads) void k_batch(const BatchArgs a) {
  using W = word_t<BS>;
  int part = 0, nparts = 0;
  if (!grid_join(a.sync, a.join_ticks, a.spin_ticks, part, nparts)) return;  // workgroup-uniform
  part = __builtin_amdgcn_readfirstlane(part);
  nparts = __builtin_amdgcn_readfirstlane(nparts);
  constexpr int RG = row_group<BS>();
  const uint64_t kLow = low_mask<BS>();
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  const int64_t gw = (int64_t)part * kBatchWaves + wave;
  const int64_t nw = (int64_t)nparts * kBatchWaves;
  const int64_t gt = (int64_t)part * kBatchThreads + threadIdx.x;
  const int64_t nt = (int64_t)nparts * kBatchThreads;
  const int64_t nbr = a.nblockrows, nu = nbr * BS;
  const int64_t ngroups = (nbr + RG - 1) / RG;
  W* f0 = (W*)a.f0;
  W* f1 = (W*)a.f1;
  W* vis = (W*)a.vis;
  __shared__ int s_any;
  // wave-private set-bit lists and accumulators of the bs 8/16 CC sweep
  constexpr int kListWaves = BS <= 16 ? kBatchWaves : 1;
  __shared__ uint32_t s_idx[kListWaves][kListCap];
  __shared__ uint8_t s_row[kListWaves][kListCap];
  __shared__ int32_t s_acc[kListWaves][RG * BS];
  GridBarrier bar{a.sync, a.spin_ticks, (unsigned int)nparts, part == 0};

  // ------------------------------------------------------------ BFS
  for (int64_t R = gt; R < nbr; R += nt) {
    const int g = a.rowg[R];
    const int64_t s = a.vbase[g] + a.src[g];
    const W w = (R == s / BS) ? (W)(1ull << (BS - 1 - (s % BS))) : (W)0;
    f0[R] = w;
    vis[R] = w;
  }
  for (int64_t u = gt; u < nu; u += nt) {
    const int g = a.rowg[u / BS];
    const int64_t v = u - a.vbase[g];
    if (v < a.n[g]) a.level[a.out[g] + v] = (v == a.src[g]) ? 0 : -1;
  }
  for (int64_t g = gt; g < a.ngraphs; g += nt) {
    a.bfs_act[g] = 0;
    a.cc_act[g] = 0;
  }
  if (!bar.sync()) return;
  for (int32_t k = 0;; ++k) {
    // frontier of level k in fk; the vertices found at level k+1 go to fn
    const W* fk = (k & 1) ? f1 : f0;
    W* fn = (k & 1) ? f0 : f1;
    bar.open(&s_any);
    int marked = -1;  // graph whose marker this wave already raised to k+1
    for (int64_t q = gw; q < ngroups; q += nw) {
      // lane j < nrow describes block row R0 + j
      const int64_t R0 = q * RG;
      const int nrow = nbr - R0 < RG ? (int)(nbr - R0) : RG;
      int32_t l_rs = 0, l_re = 0, l_g = 0;
      int64_t l_vw = 0, l_ob = 0;
      bool sweep = false;
      if (lane < nrow) {
        const int64_t R = R0 + lane;
        l_rs = a.rowptr[R];
        l_re = a.rowptr[R + 1];
        l_g = a.rowg[R];
        l_vw = (int64_t)(uint64_t)vis[R];
        const int64_t vb = a.vbase[l_g];
        const int64_t valid = vb + a.n[l_g] - R * BS;
        const uint64_t pad = valid >= BS ? 0ull : (kLow >> valid);
        l_ob = a.out[l_g] - vb;
        // skip: the graph's frontier is empty, or every vertex of the block
        // row is visited (rows past the graph's last vertex count as visited)
        sweep = a.bfs_act[l_g] >= k && (((uint64_t)l_vw | pad) & kLow) != kLow && l_re > l_rs;
        if (!sweep) fn[R] = (W)0;
      }
      uint64_t todo = __ballot(sweep);
      if constexpr (BS <= 16) {
        if (!todo) continue;
        uint32_t gm[4];
        bfs_group_lpt<BS, RG>(a, fk, nrow, l_rs, l_re, todo, lane, gm);
        // lane j < nrow: block row R0 + j
        uint64_t word = 0;
        if (lane < nrow && ((todo >> lane) & 1)) {
          const int bit = lane * BS;
          uint32_t gq = gm[0];
#pragma unroll
          for (int q = 1; q < 4; ++q) gq = (bit >> 5) == q ? gm[q] : gq;
          const uint64_t rows = (gq >> (bit & 31)) & kLow;
          word = brev_low<BS>(rows) & ~(uint64_t)l_vw & kLow;
          fn[R0 + lane] = (W)word;
          if (word) vis[R0 + lane] = (W)((uint64_t)l_vw | word);
        }
        uint64_t pend = __ballot(word != 0);
        if (!pend) continue;
        if (lane == 0) s_any = 1;
        for (uint64_t mk = pend; mk;) {  // each graph's marker once
          const int gl = __builtin_amdgcn_readlane(l_g, __builtin_ctzll(mk));
          mk &= ~__ballot(word != 0 && l_g == gl);
          if (lane == 0 && gl != marked && a.bfs_act[gl] < k + 1) atomicMax(&a.bfs_act[gl], k + 1);
          marked = gl;
        }
        for (int p = lane; p < nrow * BS; p += kWave) {  // levels of the new vertices
          const int j = p / BS, kk = p % BS;
          const uint64_t wj = __shfl(word, j);
          const int64_t obj = __shfl(l_ob, j);
          if ((wj >> (BS - 1 - kk)) & 1) a.level[obj + (R0 + j) * BS + kk] = k + 1;
        }
        continue;
      }
      while (todo) {
        const int j = __builtin_ctzll(todo);
        todo &= todo - 1;
        const int64_t R = R0 + j;
        const int32_t rs = __builtin_amdgcn_readlane(l_rs, j), re = __builtin_amdgcn_readlane(l_re, j);
        const uint64_t vw = (uint64_t)readlane64(l_vw, j);
        const uint64_t rows = bfs_row_rl<BS>(a, fk, rs, re, lane);
        const uint64_t word = brev_low<BS>(rows) & ~vw & kLow;
        if (lane == 0) fn[R] = (W)word;
        if (word) {  // wave-uniform
          const int g = __builtin_amdgcn_readlane(l_g, j);
          if (lane == 0) {
            vis[R] = (W)(vw | word);
            if (g != marked && a.bfs_act[g] < k + 1) atomicMax(&a.bfs_act[g], k + 1);
            s_any = 1;
          }
          marked = g;
          const int64_t ob = readlane64(l_ob, j);
          if (lane < BS && ((word >> (BS - 1 - lane)) & 1)) a.level[ob + R * BS + lane] = k + 1;
        }
      }
    }
    const int more = bar.close(&s_any);
    if (more < 0) return;
    if (!more) break;
  }

  // ------------------------------------------------------------- CC (FastSV)
  for (int64_t u = gt; u < nu; u += nt) {
    a.lab[u] = (int32_t)u;
    a.gf[u] = (int32_t)u;
  }
  if (!bar.sync()) return;
  for (int32_t it = 0;; ++it) {
    // a graph takes part in iteration it if it changed in iteration it-1
    for (int64_t q = gw; q < ngroups; q += nw) {
      const int64_t R0 = q * RG;
      const int nrow = nbr - R0 < RG ? (int)(nbr - R0) : RG;
      int32_t l_rs = 0, l_re = 0;
      bool sweep = false;
      if (lane < nrow) {
        const int64_t R = R0 + lane;
        l_rs = a.rowptr[R];
        l_re = a.rowptr[R + 1];
        sweep = a.cc_act[a.rowg[R]] >= it;
      }
      uint64_t todo = __ballot(sweep);
      if constexpr (BS <= 16) {
        if (todo) cc_group_lpt<BS, RG>(a, R0, nrow, l_rs, l_re, todo, lane, s_idx[wave], s_row[wave], s_acc[wave]);
      } else {
        while (todo) {
          const int j = __builtin_ctzll(todo);
          todo &= todo - 1;
          const int32_t rs = __builtin_amdgcn_readlane(l_rs, j), re = __builtin_amdgcn_readlane(l_re, j);
          cc_row_rl<BS>(a, R0 + j, rs, re, lane);
        }
      }
    }
    if (!bar.sync()) return;
    bar.open(&s_any);
    for (int64_t u = gt; u < nu; u += nt) {  // grandparents
      const int g = a.rowg[u / BS];
      bool changed = false;
      if (a.cc_act[g] >= it && u - a.vbase[g] < a.n[g]) {
        const int32_t g2 = a.lab[a.lab[u]];
        if (g2 != a.gf[u]) {
          a.gf[u] = g2;
          changed = true;
        }
      }
      // mark each changed graph once per wave: its lowest changed lane does
      uint64_t pend = __ballot(changed);
      if (pend && lane == 0) s_any = 1;
      while (pend) {
        const int gl = __builtin_amdgcn_readlane(g, __builtin_ctzll(pend));
        pend &= ~__ballot(changed && g == gl);
        if (lane == 0 && a.cc_act[gl] < it + 1) atomicMax(&a.cc_act[gl], it + 1);
      }
    }
    const int more = bar.close(&s_any);
    if (more < 0) return;
    if (!more) break;
  }
  while (true) {  // pointer jumping to stars
    bar.open(&s_any);
    for (int64_t u = gt; u < nu; u += nt) {
      const int32_t p = a.lab[u], gp = a.lab[p];
      if (gp != p) {
        a.lab[u] = gp;
        s_any = 1;
      }
    }
    const int more = bar.close(&s_any);
    if (more < 0) return;
    if (!more) break;
  }
  for (int64_t u = gt; u < nu; u += nt) {
    const int g = a.rowg[u / BS];
    const int64_t v = u - a.vbase[g];
    if (v < a.n[g]) a.label[a.out[g] + v] = (int32_t)(a.lab[u] - a.vbase[g]);
  }
}

}  // namespace gbo

using namespace gbo;

static void batch_release(gbo_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  gbo_bsr_free(&b->A);
  for (void* p : {(void*)b->d_rowg, (void*)b->d_vbase, (void*)b->d_n, (void*)b->d_out, (void*)b->d_src, b->d_f0,
                  b->d_f1, b->d_vis, (void*)b->d_lab, (void*)b->d_gf, (void*)b->d_act,
                  (void*)b->d_sync})
    if (p) (void)hipFree(p);
  delete b;
}

template <class T>
static int upload(T** d, const std::vector<T>& h) {
  GBO_HIP_CHECK(hipMalloc((void**)d, sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) GBO_HIP_CHECK(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return GBO_OK;
}

// Launch-time constants: the grid is every workgroup that fits on the GPU at
// once (the grid barrier needs all of them resident), at most one per
// kBatchWaves block rows.
template <int BS>
static int batch_kernel_launch(const BatchArgs& a, hipStream_t st) {
  int dev = 0, cus = 0, per_cu = 0;
  GBO_HIP_CHECK(hipGetDevice(&dev));
  GBO_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  GBO_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batch<BS>, kBatchThreads, 0));
  if (per_cu < 1) return set_error(GBO_EDEVICE, "gbo_batch_bfs_cc: kernel does not fit on a CU");
  constexpr int RG = row_group<BS>();
  const int64_t want = ((a.nblockrows + RG - 1) / RG + kBatchWaves - 1) / kBatchWaves;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)per_cu * cus));
  k_batch<BS><<<grid, kBatchThreads, 0, st>>>(a);
  return GBO_OK;
}

// Barrier wait limit in ms: kBatchSpinMs, or GBO_BATCH_SPIN_MS from the
// environment (0 makes every wait expire at once: the error path can be
// exercised without anything hanging).
static int64_t batch_spin_ms() {
  const char* e = std::getenv("GBO_BATCH_SPIN_MS");
  if (!e || !*e) return kBatchSpinMs;
  const long long v = std::atoll(e);
  return v < 0 ? 0 : (int64_t)v;
}

static int batch_launch(gbo_batch* b, int32_t* d_level, int32_t* d_label, int64_t spin_ms, hipStream_t st) {
  BatchArgs a{};
  a.rowptr = b->A.rowptr;
  a.colind = b->A.colind;
  a.tiles = b->A.tiles;
  a.nblockrows = b->A.nblockrows;
  a.ngraphs = b->ngraphs;
  a.rowg = b->d_rowg;
  a.vbase = b->d_vbase;
  a.n = b->d_n;
  a.out = b->d_out;
  a.src = b->d_src;
  a.f0 = b->d_f0;
  a.f1 = b->d_f1;
  a.vis = b->d_vis;
  a.lab = b->d_lab;
  a.gf = b->d_gf;
  a.bfs_act = b->d_act;
  a.cc_act = b->d_act + b->ngraphs;
  a.level = d_level;
  a.label = d_label;
  a.sync = b->d_sync;
  int dev = 0, khz = 0;
  GBO_HIP_CHECK(hipGetDevice(&dev));
  GBO_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  a.spin_ticks = (int64_t)std::max(khz, 1) * spin_ms;
  a.join_ticks = (int64_t)std::max(khz, 1) * kBatchJoinUs / 1000;
  k_batch_reset<<<1, 1, 0, st>>>(b->d_sync);
  int rc = GBO_OK;
  switch (b->bs) {
    case 8: rc = batch_kernel_launch<8>(a, st); break;
    case 16: rc = batch_kernel_launch<16>(a, st); break;
    case 32: rc = batch_kernel_launch<32>(a, st); break;
    case 64: rc = batch_kernel_launch<64>(a, st); break;
  }
  if (rc) return rc;
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

// After the stream has drained: a grid barrier that timed out means the
// results are incomplete.
static int batch_check(gbo_batch* b) {
  unsigned int abort = 0;
  GBO_HIP_CHECK(hipMemcpy(&abort, &b->d_sync->abort, sizeof(abort), hipMemcpyDeviceToHost));
  if (abort)
    return set_error(GBO_EDEVICE, "gbo_batch_bfs_cc: grid barrier wait exceeded %lld ms",
                     (long long)b->spin_ms);
  return GBO_OK;
}

extern "C" {

int gbo_batch_create(int bs, int32_t ngraphs, const int32_t* h_nrows, const int32_t* const* h_rowptr,
                     const int32_t* const* h_colind, gbo_batch** out, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  if (!out) return set_error(GBO_EINVAL, "gbo_batch_create: null output");
  *out = nullptr;
  if (!(bs == 8 || bs == 16 || bs == 32 || bs == 64)) return set_error(GBO_EUNSUPPORTED, "gbo_batch_create: block size");
  if (ngraphs < 1 || !h_nrows || !h_rowptr || !h_colind) return set_error(GBO_EINVAL, "gbo_batch_create: bad argument");
  gbo_batch* b = new gbo_batch();
  b->bs = bs;
  b->ngraphs = ngraphs;
  std::vector<int64_t> vbase(ngraphs), outo(ngraphs + 1, 0);
  std::vector<int32_t> blk0(ngraphs), blk1(ngraphs), nn(ngraphs);
  int64_t nu = 0, nnz = 0;
  for (int32_t g = 0; g < ngraphs; ++g) {
    if (h_nrows[g] < 1 || !h_rowptr[g]) {
      batch_release(b);
      return set_error(GBO_EINVAL, "gbo_batch_create: graph %d is empty", g);
    }
    vbase[g] = nu;
    nn[g] = h_nrows[g];
    blk0[g] = (int32_t)(nu / bs);
    nu += ((int64_t)h_nrows[g] + bs - 1) / bs * bs;
    blk1[g] = (int32_t)(nu / bs);
    outo[g + 1] = outo[g] + h_nrows[g];
    nnz += h_rowptr[g][h_nrows[g]];
  }
  if (nu >= INT32_MAX || nnz >= INT32_MAX) {
    batch_release(b);
    return set_error(GBO_ERANGE, "gbo_batch_create: batch too large for int32 indices");
  }
  // union CSR
  std::vector<int32_t> rp((size_t)nu + 1, 0), ci((size_t)nnz);
  int64_t k = 0;
  for (int32_t g = 0; g < ngraphs; ++g) {
    for (int32_t v = 0; v < h_nrows[g]; ++v) {
      const int64_t u = vbase[g] + v;
      rp[u] = (int32_t)k;
      for (int32_t q = h_rowptr[g][v]; q < h_rowptr[g][v + 1]; ++q) {
        const int32_t c = h_colind[g][q];
        if (c < 0 || c >= h_nrows[g]) {
          batch_release(b);
          return set_error(GBO_EINVAL, "gbo_batch_create: graph %d column out of range", g);
        }
        ci[k++] = (int32_t)(vbase[g] + c);
      }
    }
    for (int64_t u = vbase[g] + h_nrows[g]; u < (g + 1 < ngraphs ? vbase[g + 1] : nu); ++u) rp[u] = (int32_t)k;
  }
  rp[nu] = (int32_t)k;
  int rc = gbo_bsr_create(bs, 0, rp.data(), ci.data(), (int32_t)nu, (int32_t)nu, &b->A, stream);
  if (rc) {
    batch_release(b);
    return rc;
  }
  // graph of each block row
  std::vector<int32_t> rowg((size_t)b->A.nblockrows);
  for (int32_t g = 0; g < ngraphs; ++g) std::fill(rowg.begin() + blk0[g], rowg.begin() + blk1[g], g);
  b->nunion = nu;
  b->out_off = outo;
  std::vector<int64_t> oo(outo.begin(), outo.end() - 1);
  if ((rc = upload(&b->d_rowg, rowg)) || (rc = upload(&b->d_vbase, vbase)) || (rc = upload(&b->d_n, nn)) ||
      (rc = upload(&b->d_out, oo))) {
    batch_release(b);
    return rc;
  }
  const size_t wb = (size_t)bs / 8 * ((size_t)b->A.nblockrows + 1);
  hipError_t e = hipSuccess;
  e = e ? e : hipMalloc(&b->d_f0, wb);
  e = e ? e : hipMalloc(&b->d_f1, wb);
  e = e ? e : hipMalloc(&b->d_vis, wb);
  e = e ? e : hipMalloc((void**)&b->d_lab, sizeof(int32_t) * nu);
  e = e ? e : hipMalloc((void**)&b->d_gf, sizeof(int32_t) * nu);
  e = e ? e : hipMalloc((void**)&b->d_act, sizeof(int32_t) * 2 * ngraphs);
  e = e ? e : hipMalloc((void**)&b->d_sync, sizeof(GridSync));
  e = e ? e : hipMalloc((void**)&b->d_src, sizeof(int32_t) * ngraphs);
  if (e != hipSuccess) {
    batch_release(b);
    return set_error(GBO_ENOMEM, "gbo_batch_create: scratch allocation failed");
  }
  *out = b;
  return GBO_OK;
  GBO_TRY_END
}

int gbo_batch_offsets(const gbo_batch* b, int64_t* h_offsets) {
  if (!b || !h_offsets) return set_error(GBO_EINVAL, "gbo_batch_offsets: null argument");
  std::copy(b->out_off.begin(), b->out_off.end(), h_offsets);
  return GBO_OK;
}

int gbo_batch_info(const gbo_batch* b, int32_t* bs, int64_t* nblocks, int64_t* nvertices_padded) {
  if (!b) return set_error(GBO_EINVAL, "gbo_batch_info: null batch");
  if (bs) *bs = b->bs;
  if (nblocks) *nblocks = b->A.nblocks;
  if (nvertices_padded) *nvertices_padded = b->nunion;
  return GBO_OK;
}

int gbo_batch_bfs_cc(gbo_batch* b, const int32_t* h_sources, int32_t* d_level, int32_t* d_label, int use_graph,
                     gbo_stream_t stream) {
  if (!b || !h_sources || !d_level || !d_label) return set_error(GBO_EINVAL, "gbo_batch_bfs_cc: null argument");
  for (int32_t g = 0; g < b->ngraphs; ++g) {
    const int64_t n = b->out_off[g + 1] - b->out_off[g];
    if (h_sources[g] < 0 || h_sources[g] >= n)
      return set_error(GBO_EINVAL, "gbo_batch_bfs_cc: source of graph %d out of range", g);
  }
  hipStream_t st = stream;
  GBO_HIP_CHECK(hipMemcpyAsync(b->d_src, h_sources, sizeof(int32_t) * b->ngraphs, hipMemcpyHostToDevice, st));
  const int64_t spin_ms = batch_spin_ms();
  if (!use_graph) {
    int rc = batch_launch(b, d_level, d_label, spin_ms, st);
    if (rc) return rc;
    b->spin_ms = spin_ms;
    GBO_HIP_CHECK(hipStreamSynchronize(st));
    return batch_check(b);
  }
  // the outputs and the wait limit are baked into the captured launch:
  // capture again when either differs
  if (b->exec && (b->graph_level != d_level || b->graph_label != d_label || b->graph_spin_ms != spin_ms)) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  if (!b->exec) {
    // capture on a private stream (a capture cannot start on the null stream)
    hipStream_t cs;
    GBO_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    GBO_HIP_CHECK(hipStreamSynchronize(st));
    GBO_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = batch_launch(b, d_level, d_label, spin_ms, cs);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &gr);
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(GBO_EDEVICE, "gbo_batch_bfs_cc: capture failed: %s", hipGetErrorString(e));
    b->graph = gr;
    GBO_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_level = d_level;
    b->graph_label = d_label;
    b->graph_spin_ms = spin_ms;
  }
  b->spin_ms = b->graph_spin_ms;
  GBO_HIP_CHECK(hipGraphLaunch(b->exec, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  return batch_check(b);
}

void gbo_batch_free(gbo_batch* b) { batch_release(b); }

}  // extern "C"
