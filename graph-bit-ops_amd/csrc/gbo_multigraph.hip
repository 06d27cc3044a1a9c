// Many-graph batch: BFS + connected components of hundreds of independent
// graphs in one persistent launch.  No reference counterpart: the reference
// runs one graph per process (algorithm/bfs.cu, algorithm/cc.cu) with a
// host round trip per iteration.
//
// The graphs are packed side by side as one block-diagonal bit-BSR: graph g
// occupies vertices [vbase_g, vbase_g + n_g) with vbase_g a multiple of bs,
// so every tile belongs to exactly one graph.  A persistent kernel is
// launched with a few workgroups per CU; each workgroup takes the next graph
// from an atomic queue (largest first) and runs that graph's BFS (pull,
// frontier bit words) and FastSV CC to convergence with workgroup barriers
// only -- graphs terminate independently, without host synchronisation.
// The launch (queue reset kernel + persistent kernel) is captured once in a
// hipGraph and replayed.
#include <algorithm>
#include <numeric>
#include <vector>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

struct gbo_batch {
  int bs = 0;
  int32_t ngraphs = 0;
  gbo_bsr A{};                 // union matrix
  int64_t nunion = 0;          // padded vertex count
  std::vector<int64_t> out_off;  // per-graph offset in the outputs (unpadded), ngraphs+1
  // device descriptors
  int32_t* d_blk0 = nullptr;   // first block row
  int32_t* d_blk1 = nullptr;   // one past the last block row
  int64_t* d_vbase = nullptr;  // first union vertex
  int32_t* d_n = nullptr;      // vertex count
  int64_t* d_out = nullptr;    // output offset
  int32_t* d_order = nullptr;  // processing order (largest first)
  int32_t* d_src = nullptr;    // BFS sources (graph-local)
  // scratch over the union
  void* d_f = nullptr;         // frontier bits (nblockrows words)
  void* d_nx = nullptr;
  void* d_vis = nullptr;
  int32_t* d_lab = nullptr;    // FastSV parent (union ids)
  int32_t* d_gf = nullptr;
  int32_t* d_mngf = nullptr;
  unsigned int* d_queue = nullptr;
  // captured launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t* graph_level = nullptr;
  int32_t* graph_label = nullptr;
};

namespace gbo {

struct BatchArgs {
  const int32_t* rowptr;
  const int32_t* colind;
  const void* tiles;
  int32_t ngraphs;
  const int32_t* blk0;
  const int32_t* blk1;
  const int64_t* vbase;
  const int32_t* n;
  const int64_t* out;
  const int32_t* order;
  const int32_t* src;
  void* f;
  void* nx;
  void* vis;
  int32_t* lab;
  int32_t* gf;
  int32_t* mngf;
  int32_t* level;
  int32_t* label;
  unsigned int* queue;
};

constexpr int kBatchThreads = 512;
constexpr int kBatchWaves = kBatchThreads / kWave;

__global__ void k_batch_reset(unsigned int* q) { *q = 0; }

template <int BS>
__global__ __launch_bounds__(kBatchThreads) void k_batch(const BatchArgs a) {
  using W = word_t<BS>;
  constexpr int G = kWave / BS;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid / kWave;
  const int grp = lane / BS, r = lane % BS;
  const W* __restrict__ tiles = (const W*)a.tiles;
  W* f = (W*)a.f;
  W* nx = (W*)a.nx;
  W* vis = (W*)a.vis;
  __shared__ int s_g;
  __shared__ unsigned long long s_cnt;

  while (true) {
    if (tid == 0) s_g = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int qi = s_g;
    __syncthreads();
    if (qi >= a.ngraphs) break;  // workgroup-uniform
    const int g = a.order[qi];
    const int32_t b0 = a.blk0[g], b1 = a.blk1[g];
    const int64_t vb = a.vbase[g];
    const int32_t n = a.n[g];
    const int64_t oo = a.out[g];
    const int32_t src = a.src[g];

    // ------------------------------------------------------------ BFS
    for (int64_t R = b0 + tid; R < b1; R += kBatchThreads) {
      const int64_t s_union = vb + src;
      const W w = (R == s_union / BS) ? (W)(1ull << (BS - 1 - (s_union % BS))) : (W)0;
      f[R] = w;
      vis[R] = w;
    }
    for (int64_t v = tid; v < n; v += kBatchThreads) a.level[oo + v] = (v == src) ? 0 : -1;
    __syncthreads();
    int32_t lvl = 0;
    while (true) {
      if (tid == 0) s_cnt = 0;
      __syncthreads();
      for (int64_t R = b0 + wave; R < b1; R += kBatchWaves) {  // one wave per block row
        const int64_t rs = a.rowptr[R], re = a.rowptr[R + 1];
        const W vw = vis[R];
        int hit = 0;
        for (int64_t t = rs + grp; t < re; t += G) {
          const W fr = f[a.colind[t]];
          if (fr != 0 && !hit && (W)(tiles[t * BS + r] & fr) != 0) hit = 1;
        }
        for (int off = BS; off < kWave; off <<= 1) hit |= __shfl_xor(hit, off);
        const uint64_t bal = __ballot(hit);
        const uint64_t word = brev_low<BS>(bal & low_mask<BS>()) & ~(uint64_t)vw & low_mask<BS>();
        if (lane == 0) {
          nx[R] = (W)word;
          if (word) atomicAdd(&s_cnt, (unsigned long long)popcw(word));
        }
      }
      __syncthreads();
      const unsigned long long cnt = s_cnt;
      if (cnt == 0) break;  // workgroup-uniform
      ++lvl;
      for (int64_t R = b0 + tid; R < b1; R += kBatchThreads) {
        uint64_t w = (uint64_t)nx[R];
        f[R] = (W)w;
        if (w) {
          vis[R] = (W)((uint64_t)vis[R] | w);
          while (w) {
            const int p = __builtin_ctzll(w);
            w &= w - 1;
            const int64_t v = R * BS + (BS - 1 - p) - vb;
            if (v >= 0 && v < n) a.level[oo + v] = lvl;
          }
        }
      }
      __syncthreads();
    }

    // ------------------------------------------------------------- CC (FastSV)
    for (int64_t v = tid; v < n; v += kBatchThreads) {
      a.lab[vb + v] = (int32_t)(vb + v);
      a.gf[vb + v] = (int32_t)(vb + v);
    }
    __syncthreads();
    while (true) {
      // mngf[i] = min over neighbours of gf (one wave per block row)
      for (int64_t R = b0 + wave; R < b1; R += kBatchWaves) {
        const int64_t rs = a.rowptr[R], re = a.rowptr[R + 1];
        int32_t acc = INT32_MAX;
        for (int64_t t = rs + grp; t < re; t += G) {
          const int32_t col = a.colind[t];
          uint64_t wd = (uint64_t)tiles[t * BS + r];
          while (wd) {
            const int p = __builtin_ctzll(wd);
            wd &= wd - 1;
            const int32_t x = a.gf[(int64_t)col * BS + (BS - 1 - p)];
            acc = x < acc ? x : acc;
          }
        }
        for (int off = BS; off < kWave; off <<= 1) {
          const int32_t o = __shfl_xor(acc, off);
          acc = o < acc ? o : acc;
        }
        const int64_t i = R * BS + r;
        if (grp == 0 && i - vb < n) a.mngf[i] = acc;
      }
      __syncthreads();
      if (tid == 0) s_cnt = 0;
      for (int64_t v = tid; v < n; v += kBatchThreads) {  // hooking + shortcut
        const int64_t u = vb + v;
        const int32_t m = a.mngf[u];
        const int32_t fu = a.lab[u];
        if (m < fu) atomicMin(&a.lab[fu], m);
        const int32_t gu = a.gf[u];
        atomicMin(&a.lab[u], m < gu ? m : gu);
      }
      __syncthreads();
      for (int64_t v = tid; v < n; v += kBatchThreads) {  // grandparents
        const int64_t u = vb + v;
        const int32_t g2 = a.lab[a.lab[u]];
        if (g2 != a.gf[u]) {
          a.gf[u] = g2;
          atomicAdd(&s_cnt, 1ull);
        }
      }
      __syncthreads();
      if (s_cnt == 0) break;  // workgroup-uniform; s_cnt is reset after the next barrier
    }
    while (true) {  // pointer jumping to stars
      __syncthreads();
      if (tid == 0) s_cnt = 0;
      __syncthreads();
      for (int64_t v = tid; v < n; v += kBatchThreads) {
        const int64_t u = vb + v;
        const int32_t p = a.lab[u], gp = a.lab[p];
        if (gp != p) {
          a.lab[u] = gp;
          atomicAdd(&s_cnt, 1ull);
        }
      }
      __syncthreads();
      if (s_cnt == 0) break;
    }
    for (int64_t v = tid; v < n; v += kBatchThreads) a.label[oo + v] = (int32_t)(a.lab[vb + v] - vb);
    __syncthreads();
  }
}

}  // namespace gbo

using namespace gbo;

static void batch_release(gbo_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  gbo_bsr_free(&b->A);
  for (void* p : {(void*)b->d_blk0, (void*)b->d_blk1, (void*)b->d_vbase, (void*)b->d_n, (void*)b->d_out,
                  (void*)b->d_order, (void*)b->d_src, b->d_f, b->d_nx, b->d_vis, (void*)b->d_lab, (void*)b->d_gf,
                  (void*)b->d_mngf, (void*)b->d_queue})
    if (p) (void)hipFree(p);
  delete b;
}

template <class T>
static int upload(T** d, const std::vector<T>& h) {
  GBO_HIP_CHECK(hipMalloc((void**)d, sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) GBO_HIP_CHECK(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return GBO_OK;
}

static int batch_launch(gbo_batch* b, int32_t* d_level, int32_t* d_label, hipStream_t st) {
  BatchArgs a{};
  a.rowptr = b->A.rowptr;
  a.colind = b->A.colind;
  a.tiles = b->A.tiles;
  a.ngraphs = b->ngraphs;
  a.blk0 = b->d_blk0;
  a.blk1 = b->d_blk1;
  a.vbase = b->d_vbase;
  a.n = b->d_n;
  a.out = b->d_out;
  a.order = b->d_order;
  a.src = b->d_src;
  a.f = b->d_f;
  a.nx = b->d_nx;
  a.vis = b->d_vis;
  a.lab = b->d_lab;
  a.gf = b->d_gf;
  a.mngf = b->d_mngf;
  a.level = d_level;
  a.label = d_label;
  a.queue = b->d_queue;
  int dev = 0, cus = 0;
  GBO_HIP_CHECK(hipGetDevice(&dev));
  GBO_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = (unsigned)std::max(1, std::min(b->ngraphs, 2 * cus));
  k_batch_reset<<<1, 1, 0, st>>>(b->d_queue);
  switch (b->bs) {
    case 8: k_batch<8><<<grid, kBatchThreads, 0, st>>>(a); break;
    case 16: k_batch<16><<<grid, kBatchThreads, 0, st>>>(a); break;
    case 32: k_batch<32><<<grid, kBatchThreads, 0, st>>>(a); break;
    case 64: k_batch<64><<<grid, kBatchThreads, 0, st>>>(a); break;
  }
  GBO_LAUNCH_CHECK();
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
  // largest graphs first (longest-processing-time order)
  std::vector<int32_t> brp((size_t)b->A.nblockrows + 1);
  GBO_HIP_CHECK(hipMemcpy(brp.data(), b->A.rowptr, sizeof(int32_t) * brp.size(), hipMemcpyDeviceToHost));
  std::vector<int32_t> order(ngraphs);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    const int64_t tx = (int64_t)brp[blk1[x]] - brp[blk0[x]] + nn[x], ty = (int64_t)brp[blk1[y]] - brp[blk0[y]] + nn[y];
    return tx > ty;
  });
  b->nunion = nu;
  b->out_off = outo;
  std::vector<int64_t> oo(outo.begin(), outo.end() - 1);
  if ((rc = upload(&b->d_blk0, blk0)) || (rc = upload(&b->d_blk1, blk1)) || (rc = upload(&b->d_vbase, vbase)) ||
      (rc = upload(&b->d_n, nn)) || (rc = upload(&b->d_out, oo)) || (rc = upload(&b->d_order, order))) {
    batch_release(b);
    return rc;
  }
  const size_t wb = (size_t)bs / 8 * ((size_t)b->A.nblockrows + 1);
  hipError_t e = hipSuccess;
  e = e ? e : hipMalloc(&b->d_f, wb);
  e = e ? e : hipMalloc(&b->d_nx, wb);
  e = e ? e : hipMalloc(&b->d_vis, wb);
  e = e ? e : hipMalloc((void**)&b->d_lab, sizeof(int32_t) * nu);
  e = e ? e : hipMalloc((void**)&b->d_gf, sizeof(int32_t) * nu);
  e = e ? e : hipMalloc((void**)&b->d_mngf, sizeof(int32_t) * nu);
  e = e ? e : hipMalloc((void**)&b->d_queue, sizeof(unsigned int));
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
  if (!use_graph) {
    int rc = batch_launch(b, d_level, d_label, st);
    if (rc) return rc;
    GBO_HIP_CHECK(hipStreamSynchronize(st));
    return GBO_OK;
  }
  if (b->exec && (b->graph_level != d_level || b->graph_label != d_label)) {
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
    int rc = batch_launch(b, d_level, d_label, cs);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &gr);
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(GBO_EDEVICE, "gbo_batch_bfs_cc: capture failed: %s", hipGetErrorString(e));
    b->graph = gr;
    GBO_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_level = d_level;
    b->graph_label = d_label;
  }
  GBO_HIP_CHECK(hipGraphLaunch(b->exec, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  return GBO_OK;
}

void gbo_batch_free(gbo_batch* b) { batch_release(b); }

}  // extern "C"
