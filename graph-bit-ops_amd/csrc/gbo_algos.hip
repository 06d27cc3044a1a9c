// Graph algorithms on the bit-BSR engine: BFS (algorithm/bfs.cu), SSSP
// (algorithm/sssp.cu), PageRank (algorithm/pr.cu), connected components
// (algorithm/cc.cu, FastSV) and triangle counting (bsrbmm/tc.cu).
//
// The reference drivers launch five to nine kernels per iteration, fill
// vectors with single-thread <<<1,1>>> kernels inside the timed loop and
// copy a convergence scalar to the host each iteration.  Here every
// iteration is one bsrbmv launch plus one fused element-wise kernel (update
// of frontier/visited/levels, distances, scores or labels together with the
// convergence reduction); the host reads one 8-byte scalar per iteration.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {

template <class T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <class T>
static int fill(T* p, int64_t n, T v, hipStream_t st) {
  if (n <= 0) return GBO_OK;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  k_fill<T><<<(unsigned)blocks, 256, 0, st>>>(p, n, v);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

// ------------------------------------------------------------------ BFS
template <int BS>
__global__ __launch_bounds__(256) void k_bfs_init(word_t<BS>* f, word_t<BS>* vis, int32_t* level, int32_t nbr,
                                                  int32_t n, int32_t source) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbr) {
    const word_t<BS> w = (i == source / BS) ? (word_t<BS>)(1ull << (BS - 1 - (source % BS))) : (word_t<BS>)0;
    f[i] = w;
    vis[i] = w;
  }
  if (i < n) level[i] = (i == source) ? 0 : -1;
}

// visited |= next; level[v] = lvl for each newly discovered v.
template <int BS>
__global__ __launch_bounds__(256) void k_bfs_update(const word_t<BS>* __restrict__ next, word_t<BS>* __restrict__ vis,
                                                    int32_t* __restrict__ level, int32_t nbr, int32_t n, int32_t lvl) {
  const int64_t J = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= nbr) return;
  uint64_t w = (uint64_t)next[J];
  if (!w) return;
  vis[J] = (word_t<BS>)((uint64_t)vis[J] | w);
  while (w) {
    const int p = __builtin_ctzll(w);
    w &= w - 1;
    const int64_t v = J * BS + (BS - 1 - p);
    if (v < n) level[v] = lvl;
  }
}

// ---- push steps (direction-optimizing BFS) over A, the out-edge matrix.
// Frontier block rows are cut into segments of at most kPushSeg tiles so that
// a hub row is spread over many wavefronts.
constexpr int32_t kPushSeg = 4096;

template <int BS>
__global__ __launch_bounds__(256) void k_push_segs(const word_t<BS>* __restrict__ f, const int32_t* __restrict__ rowptr,
                                                   int32_t nbr, int32_t* __restrict__ segcnt) {
  const int64_t J = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (J > nbr) return;
  segcnt[J] = (J < nbr && f[J] != 0) ? (rowptr[J + 1] - rowptr[J] + kPushSeg - 1) / kPushSeg : 0;
}

// nx[J'] |= the unvisited out-neighbours of the frontier rows of every tile
// (J, J') of a frontier block row J; segments are taken grid-stride.
template <int BS>
__global__ __launch_bounds__(256) void k_push(const word_t<BS>* __restrict__ f, const word_t<BS>* __restrict__ vis,
                                              word_t<BS>* __restrict__ nx, const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ colind, const word_t<BS>* __restrict__ tiles,
                                              const int32_t* __restrict__ segoff, int32_t nbr) {
  const int lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t S = segoff[nbr];
  for (int64_t sg = wave_index(); sg < S; sg += nw) {
    int32_t lo = 0, hi = nbr;  // block row: last J with segoff[J] <= sg
    while (lo < hi) {
      const int32_t mid = lo + (hi - lo + 1) / 2;
      if (segoff[mid] <= sg) lo = mid;
      else hi = mid - 1;
    }
    const int32_t J = lo;
    const int64_t t0 = (int64_t)rowptr[J] + (sg - segoff[J]) * kPushSeg;
    const int64_t t1 = min(t0 + kPushSeg, (int64_t)rowptr[J + 1]);
    const uint64_t fw = (uint64_t)f[J];
    for (int64_t t = t0 + lane; t < t1; t += kWave) {
      uint64_t nb = 0;  // out-neighbours of the frontier rows of this tile
      uint64_t rows = fw;
      while (rows) {
        const int p = __builtin_ctzll(rows);
        rows &= rows - 1;
        nb |= (uint64_t)tiles[t * BS + (BS - 1 - p)];
      }
      const int32_t J2 = colind[t];
      nb &= ~(uint64_t)vis[J2] & low_mask<BS>();
      if (nb) atomic_or_word<BS>(nx, J2, nb);
    }
  }
}

// Set bits of the next-frontier words, summed per wave into the count slots
// (folded by count_fold): one counter for every wave took 85 us on a pushed
// level of 163K vertices (R-MAT s22), 8K waves queueing on one address.
template <int BS>
__global__ __launch_bounds__(256) void k_popc_words(const word_t<BS>* __restrict__ w, int32_t nbr,
                                                    unsigned long long* __restrict__ slots) {
  const int64_t J = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t c = J < nbr ? (uint64_t)popcw(w[J]) : 0ull;
  for (int off = 1; off < kWave; off <<= 1) c += __shfl_xor(c, off);
  if (lane_id() == 0 && c) atomicAdd(&slots[(J / kWave) & (kCountSlots - 1)], (unsigned long long)c);
}

// Scratch of a BFS call supplied by the caller of bfs_t instead of being
// allocated: frontier / next / visited words (3 x nbr words), the frontier-
// size slots (kCountSlots, cleared on entry) and count, and (where
// dense_bfs_step holds and they fit) the dense step's row counts (bs x
// (nbr + 1) int32, else null).
struct BfsScratch {
  void* words;
  unsigned long long* slots;
  int64_t* count;
  int32_t* dense_counts;
};

// Pull BFS over AT; with A (the out-edge matrix, AT itself for an undirected
// graph) levels whose frontier holds at most n/64 vertices are pushed.  The
// frontier-size slots and the dense step's row counts come from the plan's
// scratch; the frontier, next and visited words are this call's -- or, with
// `ext` (pull only), all of them come from there.
template <int BS>
static int bfs_t(const gbo_bsr* A, const gbo_bsr* Aout, int32_t source, int32_t* d_level, int64_t* sizes,
                 int32_t max_levels, int32_t* nlevels, hipStream_t st, const BfsScratch* ext = nullptr) {
  using W = word_t<BS>;
  const int32_t nbr = A->nblockrows, n = A->nrows;
  AsyncAllocs mem(st);
  W* f = nullptr;
  int rc;
  if (ext) {
    f = (W*)ext->words;
  } else if ((rc = mem.alloc(&f, sizeof(W) * (size_t)std::max(nbr, 1) * 3 + 64, "gbo_bfs"))) {
    return rc;
  }
  W* nx = f + std::max(nbr, 1);
  W* vis = nx + std::max(nbr, 1);
  int32_t* segoff = nullptr;
  if (Aout && (rc = mem.alloc(&segoff, sizeof(int32_t) * ((size_t)nbr + 1), "gbo_bfs_do"))) return rc;
  PlanGuard g(st);
  if ((rc = acquire_plan(A, &g.pv, st))) return rc;
  PlanView pv = g.pv;
  if (ext) {
    pv.slots = ext->slots;
    pv.count = ext->count;
    pv.bfs_counts = ext->dense_counts;
    GBO_HIP_CHECK(hipMemsetAsync(pv.slots, 0, sizeof(unsigned long long) * kCountSlots, st));
  }
  int64_t* d_cnt = pv.count;
  const unsigned blocks_n = (unsigned)((std::max<int64_t>(n, nbr) + 255) / 256);
  k_bfs_init<BS><<<blocks_n, 256, 0, st>>>(f, vis, d_level, nbr, n, source);
  GBO_LAUNCH_CHECK();
  SpmvArgs a{};
  a.rowptr = A->rowptr;
  a.colind = A->colind;
  a.tiles = A->tiles;
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.nbr = nbr;
  a.T = A->nblocks;
  a.plan = pv.plan;
  a.chunk = pv.chunk;
  a.nchunks = pv.nchunks;
  a.variant = pv.variant;
  a.count_slots = pv.slots;
  // bs 8/16 variant 3 and bs 32/64: levels while fewer than half the
  // vertices are visited take the dense step (count kernel + pack); later
  // ones skip visited rows
  int32_t* dense_counts = dense_bfs_step(BS, pv.variant) ? pv.bfs_counts : nullptr;
  int64_t visited = 1;
  int32_t lvl = 0;
  int64_t fsize = 1;  // vertices in the current frontier
  if (sizes && max_levels > 0) sizes[0] = 1;
  while (true) {
    // both step forms end with count_fold, which writes d_cnt
    if (Aout && fsize <= std::max<int64_t>(n / 64, 1)) {
      // push: only the frontier rows' tiles of A are read
      GBO_HIP_CHECK(hipMemsetAsync(nx, 0, sizeof(W) * (size_t)std::max(nbr, 1), st));
      k_push_segs<BS><<<(unsigned)(((int64_t)nbr + 1 + 255) / 256), 256, 0, st>>>(f, Aout->rowptr, nbr, segoff);
      GBO_LAUNCH_CHECK();
      if ((rc = scan_exclusive<int32_t, int32_t>(segoff, segoff, (int64_t)nbr + 1, st))) return rc;
      k_push<BS><<<2048, 256, 0, st>>>(f, vis, nx, Aout->rowptr, Aout->colind, (const W*)Aout->tiles, segoff, nbr);
      GBO_LAUNCH_CHECK();
      k_popc_words<BS><<<(unsigned)((nbr + 255) / 256), 256, 0, st>>>(nx, nbr, pv.slots);
      GBO_LAUNCH_CHECK();
      if ((rc = count_fold(pv.slots, d_cnt, st))) return rc;
    } else {
      a.xbits = f;
      a.visited = vis;
      a.out = nx;
      a.bfs_counts = (dense_counts && visited < (int64_t)n / 2) ? dense_counts : nullptr;
      // a frontier of f vertices leaves at most f*BS/n of the frontier words
      // nonzero: below a quarter, tiles are read only under nonzero words
      // (bs 16-64: R-MAT s22 pull BFS bs 64 23.6 -> 17.6 ms; at bs 8 the dense
      // count step stays faster than the gated row-group kernel)
      a.sparse_frontier = BS >= 16 && fsize * BS <= (int64_t)n / 4;
      if ((rc = spmv_launch(BS, OP_BFS, a, st))) return rc;
      if ((rc = count_fold(pv.slots, d_cnt, st))) return rc;
    }
    // the level update needs no count (an empty next frontier changes
    // nothing), so it is queued before the host waits for the count
    k_bfs_update<BS><<<(unsigned)((nbr + 255) / 256), 256, 0, st>>>(nx, vis, d_level, nbr, n, lvl + 1);
    GBO_LAUNCH_CHECK();
    int64_t cnt = 0;
    GBO_HIP_CHECK(hipMemcpyAsync(&cnt, d_cnt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    GBO_HIP_CHECK(hipStreamSynchronize(st));
    if (cnt == 0) break;
    ++lvl;
    fsize = cnt;
    visited += cnt;
    if (sizes && lvl < max_levels) sizes[lvl] = cnt;
    std::swap(f, nx);
  }
  if (nlevels) *nlevels = lvl + 1;
  return GBO_OK;
}

// ------------------------------------------------------------------ SSSP
// dist = min(dist, cand); the improved vertices form the next frontier: their
// distance goes to `fr` (INF elsewhere) and their block column is marked
// active so the next min-plus product only reads tiles of active columns.
__global__ __launch_bounds__(256) void k_sssp_update(const uint32_t* __restrict__ cand, uint32_t* __restrict__ dist,
                                                     uint32_t* __restrict__ fr, unsigned long long* __restrict__ active,
                                                     int32_t n, int bs, unsigned long long* __restrict__ changed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int improved = 0;
  if (i < n) {
    const uint32_t c = cand[i], d = dist[i];
    if (c < d) {
      dist[i] = c;
      fr[i] = c;
      improved = 1;
    } else {
      fr[i] = kInfU32;
    }
  }
  // block columns of the improved vertices: the 64 vertices of a wave (64-
  // aligned) cover 64/bs consecutive block columns inside one 64-bit word
  const int64_t J = i / bs;
  const uint64_t m = wave_or64(improved ? 1ull << (J & 63) : 0ull);
  if (lane_id() == 0 && m) atomicOr(&active[((i - lane_id()) / bs) >> 6], m);
  wave_flag(improved, changed);
}

__global__ void k_sssp_init(uint32_t* dist, uint32_t* fr, unsigned long long* active, int32_t n, int bs, int32_t s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dist[i] = (i == s) ? 0u : kInfU32;
    fr[i] = (i == s) ? 0u : kInfU32;
  }
  if (i == 0) {
    const int64_t J = s / bs;
    active[J >> 6] = 1ull << (J & 63);
  }
}

// ------------------------------------------------------------------ PageRank
// x[j] = alpha * p[j] / outdeg[j] (0 for vertices without out-edges) -- the
// per-edge factor alpha*f1/nnz of bmv32_sparse_full (algorithm/bsrbmv.cu:
// 1276) computed once per vertex.
__global__ __launch_bounds__(256) void k_pr_scale(const float* __restrict__ p, const int32_t* __restrict__ outdeg,
                                                  float alpha, float* __restrict__ x, int32_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int32_t d = outdeg[i];
    x[i] = d > 0 ? alpha * p[i] / (float)d : 0.0f;
  }
}

// Per-block partial sums of |q-p| in double, and the next iteration's
// x[j] = alpha * q[j] / outdeg[j] (as k_pr_scale) from the same read of q.
// The partials are added in a fixed order by k_sum_partials, so the L1
// change is reproducible.
__global__ __launch_bounds__(256) void k_pr_tail(const float* __restrict__ q, const float* __restrict__ p,
                                                 const int32_t* __restrict__ outdeg, float alpha,
                                                 float* __restrict__ x, int32_t n, double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float qi = q[i];
    s += fabs((double)qi - (double)p[i]);
    const int32_t d = outdeg[i];
    x[i] = d > 0 ? alpha * qi / (float)d : 0.0f;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// *out = the sum of part[0..np) in a fixed order: thread t adds entries
// t, t+256, ... in turn, then a fixed tree over the threads (one block of
// 256 threads).
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int np, double* out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// ------------------------------------------------------------------ CC
// FastSV hooking (Zhang, Azad, Hu; the GraphBLAST algorithm the reference
// follows, algorithm/cc.cu:1018-1083): stochastic hooking f[f[u]] = min(.,
// mngf[u]), aggressive hooking f[u] = min(f[u], mngf[u]) and shortcutting
// f[u] = min(f[u], gf[u]), all as min-accumulating atomics.  Labels only
// decrease, so an atomic whose candidate is not below a value read before it
// would change nothing and is not issued: millions of vertices share a few
// roots, and their atomics on the root's label serialised (0.26 ms a call
// at R-MAT s22).
__global__ __launch_bounds__(256) void k_cc_hook(const int32_t* __restrict__ mngf, const int32_t* __restrict__ gf,
                                                 int32_t* __restrict__ f, int32_t n) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int32_t m = mngf[u];
  const int32_t fu = f[u];
  if (m < fu && m < f[fu]) atomicMin(&f[fu], m);
  const int32_t g = gf[u];
  const int32_t best = m < g ? m : g;
  if (best < fu) atomicMin(&f[u], best);
}

// gf[u] = f[f[u]]; *changed = 1 if any vertex's grandparent changed.
__global__ __launch_bounds__(256) void k_cc_grand(const int32_t* __restrict__ f, int32_t* __restrict__ gf, int32_t n,
                                                  unsigned long long* __restrict__ changed) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int ch = 0;
  if (u < n) {
    const int32_t g = f[f[u]];
    if (g != gf[u]) {
      gf[u] = g;
      ch = 1;
    }
  }
  wave_flag(ch, changed);
}

// f[u] = f[f[u]] until every vertex points at a root (final star shape).
__global__ __launch_bounds__(256) void k_cc_jump(int32_t* __restrict__ f, int32_t n, unsigned long long* changed) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int ch = 0;
  if (u < n) {
    const int32_t p = f[u], gp = f[p];
    if (gp != p) {
      f[u] = gp;
      ch = 1;
    }
  }
  wave_flag(ch, changed);
}

__global__ void k_iota(int32_t* a, int32_t* b, int32_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = (int32_t)i;
    b[i] = (int32_t)i;
  }
}

static int read_u64(unsigned long long* d, uint64_t* h, hipStream_t st) {
  GBO_HIP_CHECK(hipMemcpyAsync(h, d, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  return GBO_OK;
}

static int check_square(const gbo_bsr* A, const char* who) {
  int rc = validate_bsr(A, 0, who);
  if (rc) return rc;
  if (A->nrows != A->ncols) return set_error(GBO_EINVAL, "%s: matrix must be square", who);
  return GBO_OK;
}

// ------------------------------------------------- allocation-free BFS
// gbo_bfs_noalloc's scratch: one fixed array in device memory, held by
// every GPU the library runs on until the process ends.  Host code never
// passes `scratch` itself (in host code that name is only a stand-in); it
// uses the device address hipGetSymbolAddress returns, looked up once per
// device on the first call with that device current.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];  // device address of `scratch`, per device

static int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  GBO_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_error(GBO_EDEVICE, "gbo_bfs_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(GBO_EDEVICE, "gbo_bfs_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[dev];
  return GBO_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Carves the BFS scratch for AT out of the static array; GBO_ERANGE when it
// does not fit.
static int noalloc_layout(const gbo_bsr* AT, unsigned char* base, BfsScratch* s) {
  const size_t wbytes = (size_t)(AT->bs / 8) * (size_t)std::max(AT->nblockrows, 1) * 3 + 64;
  const size_t o_count = align256(sizeof(unsigned long long) * kCountSlots);
  const size_t o_words = o_count + 256;
  const size_t o_dense = o_words + align256(wbytes);
  // the dense step's row counts (4 bytes a vertex) when they fit as well;
  // without them every level takes the step that skips visited rows (same
  // levels)
  const size_t dbytes = sizeof(int32_t) * (size_t)AT->bs * ((size_t)AT->nblockrows + 1);
  const bool dense =
      dense_bfs_step(AT->bs, AT->plan_variant) && o_dense + align256(dbytes) <= kNoAllocScratchBytes;
  const size_t need = o_dense + (dense ? align256(dbytes) : 0);
  if (need > kNoAllocScratchBytes)
    return set_error(GBO_ERANGE,
                     "gbo_bfs_noalloc: %d vertices at bs %d need %zu bytes of scratch, the static array holds %zu",
                     AT->nrows, AT->bs, need, kNoAllocScratchBytes);
  s->slots = (unsigned long long*)base;
  s->count = (int64_t*)(base + o_count);
  s->words = base + o_words;
  s->dense_counts = dense ? (int32_t*)(base + o_dense) : nullptr;
  return GBO_OK;
}

}  // namespace gbo

using namespace gbo;

extern "C" {

int gbo_bfs(const gbo_bsr* AT, int32_t source, int32_t* d_level, int64_t* h_level_sizes, int32_t max_levels,
            int32_t* h_nlevels, gbo_stream_t stream) {
  int rc = check_square(AT, "gbo_bfs");
  if (rc) return rc;
  if (source < 0 || source >= AT->nrows) return set_error(GBO_EINVAL, "gbo_bfs: source %d out of range", source);
  if (!d_level) return set_error(GBO_EINVAL, "gbo_bfs: null level");
  switch (AT->bs) {
    case 8: return bfs_t<8>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
    case 16: return bfs_t<16>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
    case 32: return bfs_t<32>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
    case 64: return bfs_t<64>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
  }
  return set_error(GBO_EUNSUPPORTED, "gbo_bfs: block size");
}

int gbo_bfs_noalloc(const gbo_bsr* AT, int32_t source, int32_t* d_level, int64_t* h_level_sizes,
                    int32_t max_levels, int32_t* h_nlevels, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  int rc = check_square(AT, "gbo_bfs_noalloc");
  if (rc) return rc;
  if (source < 0 || source >= AT->nrows)
    return set_error(GBO_EINVAL, "gbo_bfs_noalloc: source %d out of range", source);
  if (!d_level) return set_error(GBO_EINVAL, "gbo_bfs_noalloc: null level");
  if (!AT->plan || AT->plan_chunk <= 0 || AT->plan_nchunks <= 0)
    return set_error(GBO_EINVAL, "gbo_bfs_noalloc: the matrix has no plan (make one with gbo_bsr_plan first)");
  unsigned char* base = nullptr;
  if ((rc = noalloc_scratch(&base))) return rc;
  BfsScratch s{};
  if ((rc = noalloc_layout(AT, base, &s))) return rc;
  switch (AT->bs) {
    case 8: return bfs_t<8>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream, &s);
    case 16: return bfs_t<16>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream, &s);
    case 32: return bfs_t<32>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream, &s);
    case 64: return bfs_t<64>(AT, nullptr, source, d_level, h_level_sizes, max_levels, h_nlevels, stream, &s);
  }
  return set_error(GBO_EUNSUPPORTED, "gbo_bfs_noalloc: block size");
  GBO_TRY_END
}

int gbo_bfs_do(const gbo_bsr* AT, const gbo_bsr* A, int32_t source, int32_t* d_level, int64_t* h_level_sizes,
               int32_t max_levels, int32_t* h_nlevels, gbo_stream_t stream) {
  int rc = check_square(AT, "gbo_bfs_do");
  if (rc) return rc;
  if (A) {
    rc = check_square(A, "gbo_bfs_do(A)");
    if (rc) return rc;
    if (A->bs != AT->bs || A->nrows != AT->nrows) return set_error(GBO_EINVAL, "gbo_bfs_do: A and AT differ in shape");
  }
  if (source < 0 || source >= AT->nrows) return set_error(GBO_EINVAL, "gbo_bfs_do: source %d out of range", source);
  if (!d_level) return set_error(GBO_EINVAL, "gbo_bfs_do: null level");
  switch (AT->bs) {
    case 8: return bfs_t<8>(AT, A, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
    case 16: return bfs_t<16>(AT, A, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
    case 32: return bfs_t<32>(AT, A, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
    case 64: return bfs_t<64>(AT, A, source, d_level, h_level_sizes, max_levels, h_nlevels, stream);
  }
  return set_error(GBO_EUNSUPPORTED, "gbo_bfs_do: block size");
}

int gbo_sssp(const gbo_bsr* AT, const uint32_t* d_w, const int64_t* d_woff, int32_t source, uint32_t* d_dist,
             int32_t* h_iters, gbo_stream_t stream) {
  int rc = check_square(AT, "gbo_sssp");
  if (rc) return rc;
  const int32_t n = AT->nrows;
  if (source < 0 || source >= n) return set_error(GBO_EINVAL, "gbo_sssp: source %d out of range", source);
  if (!d_dist) return set_error(GBO_EINVAL, "gbo_sssp: null dist");
  if ((d_w == nullptr) != (d_woff == nullptr)) return set_error(GBO_EINVAL, "gbo_sssp: weights need offsets");
  hipStream_t st = stream;
  const int bs = AT->bs;
  const int64_t nact = ((int64_t)AT->nblockcols + 63) / 64;
  AsyncAllocs mem(st);
  uint32_t* fr = nullptr;
  unsigned long long* active = nullptr;
  if ((rc = mem.alloc(&fr, sizeof(uint32_t) * (size_t)std::max(n, 1) * 2, "gbo_sssp"))) return rc;
  if ((rc = mem.alloc(&active, sizeof(unsigned long long) * (size_t)(nact + 1) * 2, "gbo_sssp"))) return rc;
  uint32_t* cand = fr + std::max(n, 1);
  unsigned long long* changed = active + 2 * nact + 1;
  unsigned long long* active2 = active + nact;
  PlanGuard g(st);
  if ((rc = acquire_plan(AT, &g.pv, st))) return rc;
  const PlanView& pv = g.pv;
  GBO_HIP_CHECK(hipMemsetAsync(active, 0, sizeof(unsigned long long) * nact, st));
  const unsigned blocks_n = (unsigned)((n + 255) / 256);
  k_sssp_init<<<blocks_n, 256, 0, st>>>(d_dist, fr, active, n, bs, source);
  GBO_LAUNCH_CHECK();
  SpmvArgs a{};
  a.rowptr = AT->rowptr;
  a.colind = AT->colind;
  a.tiles = AT->tiles;
  a.nrows = n;
  a.ncols = AT->ncols;
  a.nbr = AT->nblockrows;
  a.T = AT->nblocks;
  a.plan = pv.plan;
  a.chunk = pv.chunk;
  a.nchunks = pv.nchunks;
  a.variant = pv.variant;
  a.w = d_w;
  a.woff = d_woff;
  int32_t it = 0;
  unsigned long long *act_cur = active, *act_next = active2;
  while (true) {
    a.xval = fr;
    a.out = cand;
    a.active = (const uint64_t*)act_cur;
    if ((rc = spmv_launch(bs, OP_MINPLUS, a, st))) return rc;
    GBO_HIP_CHECK(hipMemsetAsync(act_next, 0, sizeof(unsigned long long) * nact, st));
    GBO_HIP_CHECK(hipMemsetAsync(changed, 0, sizeof(unsigned long long), st));
    k_sssp_update<<<blocks_n, 256, 0, st>>>(cand, d_dist, fr, act_next, n, bs, changed);
    GBO_LAUNCH_CHECK();
    ++it;
    uint64_t h = 0;
    if ((rc = read_u64(changed, &h, st))) return rc;
    if (h == 0) break;
    if (it > n + 1)  // distances only decrease: cannot happen
      return set_error(GBO_ENOCONV, "gbo_sssp: no convergence after %d iterations", it);
    std::swap(act_cur, act_next);
  }
  if (h_iters) *h_iters = it;
  return GBO_OK;
}

int gbo_pagerank(const gbo_bsr* AT, const int32_t* d_outdeg, float alpha, double tol, int32_t max_iter, float* d_p,
                 int32_t* h_iters, double* h_last_delta, gbo_stream_t stream) {
  int rc = check_square(AT, "gbo_pagerank");
  if (rc) return rc;
  const int32_t n = AT->nrows;
  if (!d_outdeg || !d_p) return set_error(GBO_EINVAL, "gbo_pagerank: null argument");
  if (!(alpha >= 0.0f && alpha < 1.0f) || max_iter < 1) return set_error(GBO_EINVAL, "gbo_pagerank: bad alpha/max_iter");
  if (n == 0) return GBO_OK;
  hipStream_t st = stream;
  const int nparts = 1024;
  AsyncAllocs mem(st);
  float* x = nullptr;
  double* dpart = nullptr;
  if ((rc = mem.alloc(&x, sizeof(float) * (size_t)n * 2, "gbo_pagerank"))) return rc;
  if ((rc = mem.alloc(&dpart, sizeof(double) * (nparts + 1), "gbo_pagerank"))) return rc;
  float* y = x + n;
  PlanGuard g(st);
  if ((rc = acquire_plan(AT, &g.pv, st))) return rc;
  const PlanView& pv = g.pv;
  if ((rc = fill<float>(d_p, n, 1.0f / (float)n, st))) return rc;
  SpmvArgs a{};
  a.rowptr = AT->rowptr;
  a.colind = AT->colind;
  a.tiles = AT->tiles;
  a.nrows = n;
  a.ncols = AT->ncols;
  a.nbr = AT->nblockrows;
  a.T = AT->nblocks;
  a.plan = pv.plan;
  a.chunk = pv.chunk;
  a.nchunks = pv.nchunks;
  a.variant = pv.variant;
  a.partial = pv.partial;  // split-row partials from the plan's scratch
  a.scale = 1.0f;
  a.bias = (1.0f - alpha) / (float)n;
  const unsigned blocks_n = (unsigned)((n + 255) / 256);
  const unsigned lblocks = (unsigned)std::min<int64_t>(nparts, (n + 255) / 256);
  int32_t it = 0;
  double delta = 0.0;
  float* p = d_p;
  float* q = y;  // next iterate
  k_pr_scale<<<blocks_n, 256, 0, st>>>(p, d_outdeg, alpha, x, n);
  GBO_LAUNCH_CHECK();
  while (it < max_iter) {
    a.xval = x;
    a.out = q;
    if ((rc = spmv_launch(AT->bs, OP_SUM, a, st))) return rc;
    k_pr_tail<<<lblocks, 256, 0, st>>>(q, p, d_outdeg, alpha, x, n, dpart);  // x for the next iteration
    k_sum_partials<<<1, 256, 0, st>>>(dpart, (int)lblocks, dpart + nparts);
    GBO_LAUNCH_CHECK();
    GBO_HIP_CHECK(hipMemcpyAsync(&delta, dpart + nparts, sizeof(double), hipMemcpyDeviceToHost, st));
    GBO_HIP_CHECK(hipStreamSynchronize(st));
    ++it;
    std::swap(p, q);
    if (delta < tol) break;
  }
  if (p != d_p) {
    GBO_HIP_CHECK(hipMemcpyAsync(d_p, p, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
    GBO_HIP_CHECK(hipStreamSynchronize(st));
  }
  if (h_iters) *h_iters = it;
  if (h_last_delta) *h_last_delta = delta;
  if (!(delta < tol))
    return set_error(GBO_ENOCONV, "gbo_pagerank: not converged after %d iterations (delta %.3e)", it, delta);
  return GBO_OK;
}

int gbo_cc(const gbo_bsr* A, int32_t* d_label, int32_t* h_iters, gbo_stream_t stream) {
  int rc = check_square(A, "gbo_cc");
  if (rc) return rc;
  const int32_t n = A->nrows;
  if (!d_label) return set_error(GBO_EINVAL, "gbo_cc: null label");
  if (n == 0) return GBO_OK;
  hipStream_t st = stream;
  AsyncAllocs mem(st);
  int32_t* gf = nullptr;
  unsigned long long* changed = nullptr;
  if ((rc = mem.alloc(&gf, sizeof(int32_t) * (size_t)n * 2, "gbo_cc"))) return rc;
  if ((rc = mem.alloc(&changed, sizeof(unsigned long long), "gbo_cc"))) return rc;
  int32_t* mngf = gf + n;
  int32_t* f = d_label;
  const unsigned blocks_n = (unsigned)((n + 255) / 256);
  k_iota<<<blocks_n, 256, 0, st>>>(f, gf, n);
  GBO_LAUNCH_CHECK();
  PlanGuard g(st);
  if ((rc = acquire_plan(A, &g.pv, st))) return rc;
  const PlanView& pv = g.pv;
  SpmvArgs a{};
  a.rowptr = A->rowptr;
  a.colind = A->colind;
  a.tiles = A->tiles;
  a.nrows = n;
  a.ncols = A->ncols;
  a.nbr = A->nblockrows;
  a.T = A->nblocks;
  a.plan = pv.plan;
  a.chunk = pv.chunk;
  a.nchunks = pv.nchunks;
  a.variant = pv.variant;
  int32_t it = 0;
  while (true) {
    a.xval = gf;
    a.out = mngf;
    if ((rc = spmv_launch(A->bs, OP_MIN, a, st))) return rc;
    k_cc_hook<<<blocks_n, 256, 0, st>>>(mngf, gf, f, n);
    GBO_HIP_CHECK(hipMemsetAsync(changed, 0, sizeof(unsigned long long), st));
    k_cc_grand<<<blocks_n, 256, 0, st>>>(f, gf, n, changed);
    GBO_LAUNCH_CHECK();
    ++it;
    uint64_t h = 0;
    if ((rc = read_u64(changed, &h, st))) return rc;
    if (h == 0) break;
    if (it > n + 2) return set_error(GBO_ENOCONV, "gbo_cc: no convergence");
  }
  while (true) {  // make every tree a star
    GBO_HIP_CHECK(hipMemsetAsync(changed, 0, sizeof(unsigned long long), st));
    k_cc_jump<<<blocks_n, 256, 0, st>>>(f, n, changed);
    GBO_LAUNCH_CHECK();
    uint64_t h = 0;
    if ((rc = read_u64(changed, &h, st))) return rc;
    if (h == 0) break;
  }
  if (h_iters) *h_iters = it;
  return GBO_OK;
}

int gbo_tc(int bs, int32_t nrows, const int32_t* h_rp, const int32_t* h_ci, uint64_t* h_count, gbo_stream_t stream) {
  GBO_TRY_BEGIN
  if (!h_rp || !h_count || nrows < 0) return set_error(GBO_EINVAL, "gbo_tc: bad argument");
  int64_t lnnz = 0;
  int32_t *lp = nullptr, *li = nullptr, *tp = nullptr, *ti = nullptr;
  int rc = gbo_csr_tril(nrows, h_rp, h_ci, 1, &lnnz, &lp, &li);
  if (rc) return rc;
  rc = gbo_csr_transpose(nrows, nrows, lp, li, &tp, &ti);
  gbo_bsr L{}, LT{};
  if (rc == GBO_OK) rc = gbo_bsr_create(bs, 0, lp, li, nrows, nrows, &L, stream);
  if (rc == GBO_OK) rc = gbo_bsr_create(bs, 1, tp, ti, nrows, nrows, &LT, stream);
  gbo_free_host(lp);
  gbo_free_host(li);
  gbo_free_host(tp);
  gbo_free_host(ti);
  uint64_t* d = nullptr;
  if (rc == GBO_OK && hipMalloc((void**)&d, sizeof(uint64_t)) != hipSuccess) rc = set_error(GBO_ENOMEM, "gbo_tc: alloc");
  if (rc == GBO_OK) rc = gbo_bsrbmm_tc(&L, &LT, d, stream);
  if (rc == GBO_OK && hipMemcpy(h_count, d, sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess)
    rc = set_error(GBO_EDEVICE, "gbo_tc: copy");
  (void)hipFree(d);
  gbo_bsr_free(&L);
  gbo_bsr_free(&LT);
  return rc;
  GBO_TRY_END
}

}  // extern "C"
