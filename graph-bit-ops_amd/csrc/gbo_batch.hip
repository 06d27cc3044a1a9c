// Batched BFS: K <= 256 sources at once (bsrbmm with a dense bit right-hand
// side).  There is no reference counterpart; it batches the pull BFS of
// algorithm/bfs.cu:508-564 (one bmv32_sparse_bin_masked_v4 sweep per level)
// over K frontiers so that every tile is read once per level for all K
// sources instead of once per source.
//
// Layout: F, Fnext, V are n x KW uint64 words (KW = ceil(K/64)), vertex-
// major, so the K frontier bits of a vertex are one 8..32-byte load.  For a
// tile (I, J) of A^T, row r ORs F[J*bs + c] for every set bit c of its row
// word (bs 16-64: lane r of a tile group; bs 8: the set bits listed per
// tile row, four lanes per entry); rows are masked with ~V and stored (or
// OR-ed atomically for rows cut by a chunk boundary).  A per
// block-column activity mask (any frontier bit among the block's vertices)
// lets the sweep skip tiles whose columns hold no frontier vertex.
#include <algorithm>
#include <vector>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {

struct MsArgs {
  const int32_t* rowptr;
  const int32_t* colind;
  const void* tiles;
  int32_t nrows, nbr;
  int64_t T;
  const int32_t* plan;
  int32_t chunk, nchunks;
  const uint64_t* F;       // n * KW
  const uint64_t* V;       // n * KW
  uint64_t* Fn;            // n * KW (next frontier, masked by ~V)
  const uint64_t* active;  // per block column bit
};

template <int BS, int KW>
__global__ __launch_bounds__(256) void k_msbfs(const MsArgs a) {
  using W = word_t<BS>;
  constexpr int G = kWave / BS;
  const int lane = lane_id();
  const int g = lane / BS, r = lane % BS;
  const int c = wave_index();
  if (c >= a.nchunks) return;
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const W* __restrict__ tiles = (const W*)a.tiles;

  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    if (belong && rel == rsl) {  // empty block row: no frontier reaches it
      for (int k = 0; k < BS; ++k) {
        const int64_t i = (int64_t)Rl * BS + k;
        if (i >= a.nrows) break;
        for (int w = 0; w < KW; ++w) a.Fn[i * KW + w] = 0;
      }
    }
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int32_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      uint64_t acc[KW];
#pragma unroll
      for (int w = 0; w < KW; ++w) acc[w] = 0;
      // Four tiles per group and round: their column indices, then their
      // activity bits, then the words of the active ones (inactive block
      // columns hold no frontier bit: their tiles are not read), then one
      // set bit of each tile per pass with its KW frontier words, so the
      // loads of a stage are in flight together.
      constexpr int U = 4;
      for (int64_t t = s + g; t < te; t += U * G) {
        int32_t cc[U];
        uint64_t wd[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cc[u] = t + u * G < te ? a.colind[t + u * G] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          wd[u] = 0;
          if (cc[u] >= 0 && ((a.active[cc[u] >> 6] >> (cc[u] & 63)) & 1ull))  // group-uniform
            wd[u] = (uint64_t)tiles[(t + u * G) * BS + r];
        }
        while (wd[0] | wd[1] | wd[2] | wd[3]) {
          uint64_t f[U][KW];
#pragma unroll
          for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int w = 0; w < KW; ++w) f[u][w] = 0;
            if (wd[u]) {
              const int p = __builtin_ctzll(wd[u]);
              wd[u] &= wd[u] - 1;
              const int64_t v = (int64_t)cc[u] * BS + (BS - 1 - p);
#pragma unroll
              for (int w = 0; w < KW; ++w) f[u][w] = a.F[v * KW + w];
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int w = 0; w < KW; ++w) acc[w] |= f[u][w];
        }
      }
#pragma unroll
      for (int w = 0; w < KW; ++w)
        for (int off = BS; off < kWave; off <<= 1) acc[w] |= (uint64_t)__shfl_xor((long long)acc[w], off);
      const int64_t i = (int64_t)R * BS + r;
      if (g == 0 && i < a.nrows) {
#pragma unroll
        for (int w = 0; w < KW; ++w) {
          const uint64_t m = acc[w] & ~a.V[i * KW + w];
          if (full) a.Fn[i * KW + w] = m;
          else if (m) atomicOr((unsigned long long*)&a.Fn[i * KW + w], (unsigned long long)m);
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// bs 8 sweep with per-tile-row set-bit lists (the layout of the bs 8 PPR
// SpMM): a 64-tile block's set bits of active block columns are listed per
// tile row in LDS; lane l takes tile row l / 8, entries of parity (l / 4) % 2
// of that row's list, and frontier word l % 4, so one wave instruction
// gathers the 8-byte words of 16 entries (KW <= 4 words each) instead of a
// whole wave serving one set bit.  OR needs no order, so the frontiers are
// those of k_msbfs.  Four entries per lane in flight: 256 sources on R-MAT
// s22 in 16.0 ms against 16.9 with two and 23.7 with k_msbfs.
template <int KW>
__global__ __launch_bounds__(256) void k_msbfs8r(const MsArgs a) {
  constexpr int BS = 8, CAP = 128, U = 4;
  __shared__ uint32_t lists[256 / kWave][BS * CAP];
  const int lane = lane_id();
  const int row = lane >> 3, par = (lane >> 2) & 1, wi = lane & 3;
  const bool wok = wi < KW;
  uint32_t* __restrict__ list = lists[threadIdx.x / kWave];
  const int c = wave_index();
  if (c >= a.nchunks) return;
  const int64_t cb = (int64_t)c * a.chunk;
  const int64_t ce = min(cb + (int64_t)a.chunk, a.T);
  const bool last = (c == a.nchunks - 1);
  const uint64_t* __restrict__ tw = (const uint64_t*)a.tiles;

  for (int32_t R0 = a.plan[c]; R0 < a.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < a.nbr) {
      rsl = a.rowptr[Rl];
      rel = a.rowptr[Rl + 1];
    }
    const bool belong = (Rl < a.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    if (belong && rel == rsl) {  // empty block row: no frontier reaches it
      for (int k = 0; k < BS; ++k) {
        const int64_t i = (int64_t)Rl * BS + k;
        if (i >= a.nrows) break;
        for (int w = 0; w < KW; ++w) a.Fn[i * KW + w] = 0;
      }
    }
    uint64_t work = __ballot(belong && rel > rsl);
    while (work) {
      const int j = __builtin_ctzll(work);
      work &= work - 1;
      const int32_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      uint64_t acc = 0;
      for (int64_t b = s; b < te; b += kWave) {
        const int64_t nt = min((int64_t)kWave, te - b);
        const int32_t col = lane < nt ? a.colind[b + lane] : -1;
        const bool act = col >= 0 && ((a.active[col >> 6] >> (col & 63)) & 1ull);
        const uint64_t word = act ? tw[b + lane] : 0ull;
        const uint32_t vb = act ? (uint32_t)col * BS + (BS - 1) : 0u;
        int maxlen = 0, mylen = 0;
        {
          uint32_t tot[4];
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            const uint32_t qh = (uint32_t)__popc((uint32_t)(word >> (16 * h)) & 0xFFu) |
                                ((uint32_t)__popc((uint32_t)(word >> (16 * h + 8)) & 0xFFu) << 16);
            tot[h] = __builtin_amdgcn_readlane((uint32_t)wave_incl_scan((int)qh), kWave - 1);
          }
#pragma unroll
          for (int r = 0; r < BS; ++r) {
            const int l = (int)((r & 1) ? (tot[r >> 1] >> 16) : (tot[r >> 1] & 0xFFFFu));
            maxlen = l > maxlen ? l : maxlen;
            mylen = (row == r) ? l : mylen;
          }
        }
        for (int pass = 0; pass < maxlen; pass += CAP) {
#pragma unroll
          for (int r = 0; r < BS; ++r) {
            uint32_t byte = (uint32_t)(word >> (8 * r)) & 0xFFu;
            const uint32_t qr = (uint32_t)__popc(byte);
            int k = wave_incl_scan((int)qr) - (int)qr - pass;
            while (byte) {
              const int bit = __builtin_ctz(byte);
              byte &= byte - 1;
              if (k >= 0 && k < CAP) list[r * CAP + k] = vb - (uint32_t)bit;
              ++k;
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          const int lr = min(max(mylen - pass, 0), CAP);
          const int kmax = min(maxlen - pass, CAP);
          for (int k0 = 0; k0 < kmax; k0 += 2 * U) {
            uint64_t f[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int kk = k0 + 2 * u + par;
              ok[u] = kk < lr && wok;
              const uint32_t v = ok[u] ? list[row * CAP + kk] : 0u;  // vertex 0 stands in
              f[u] = a.F[(int64_t)v * KW + (wok ? wi : 0)];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
              if (ok[u]) acc |= f[u];
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
      acc |= (uint64_t)__shfl_xor((long long)acc, 4);  // the two parities
      const int64_t i = (int64_t)R * BS + row;
      if (par == 0 && wok && i < a.nrows) {
        const uint64_t m = acc & ~a.V[i * KW + wi];
        if (full) a.Fn[i * KW + wi] = m;
        else if (m) atomicOr((unsigned long long*)&a.Fn[i * KW + wi], (unsigned long long)m);
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// Zero the next-frontier words of rows cut by chunk boundaries (they are
// OR-ed atomically by several waves).
template <int BS, int KW>
__global__ __launch_bounds__(256) void k_ms_prefill(const MsArgs a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = idx / BS + 1;
  const int r = (int)(idx % BS);
  if (c >= a.nchunks) return;
  const int32_t R = a.plan[c];
  if ((int64_t)a.rowptr[R] >= c * a.chunk) return;
  const int64_t i = (int64_t)R * BS + r;
  if (i < a.nrows)
    for (int w = 0; w < KW; ++w) a.Fn[i * KW + w] = 0;
}

// V |= Fn; level[s][v] = lvl for every new (s, v); next activity mask;
// *count = 1 if any (s, v) is new.
template <int KW>
__global__ __launch_bounds__(256) void k_ms_update(const uint64_t* __restrict__ Fn, uint64_t* __restrict__ V,
                                                   int32_t* __restrict__ level, int32_t n, int32_t nsrc, int bs,
                                                   int32_t lvl, unsigned long long* __restrict__ active,
                                                   unsigned long long* __restrict__ count) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t cnt = 0, act = 0;
  if (v < n) {
    bool any = false;
#pragma unroll
    for (int w = 0; w < KW; ++w) {
      uint64_t m = Fn[v * KW + w];
      if (!m) continue;
      any = true;
      V[v * KW + w] |= m;
      cnt += __popcll(m);
      while (m) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        level[(int64_t)(w * 64 + b) * n + v] = lvl;
      }
    }
    if (any) act = 1ull << ((v / bs) & 63);
  }
  // the 64 vertices of a wave (64-aligned) fall in one word of the block
  // column mask
  const uint64_t m = wave_or64(act);
  if (lane_id() == 0 && m) atomicOr(&active[((v - lane_id()) / bs) >> 6], m);
  wave_flag(cnt != 0, count);
}

template <int KW>
__global__ void k_ms_init(const int32_t* __restrict__ src, int32_t nsrc, uint64_t* F, uint64_t* V, int32_t* level,
                          int32_t n, int bs, unsigned long long* active) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsrc) return;
  const int64_t v = src[s];
  atomicOr((unsigned long long*)&F[v * KW + s / 64], 1ull << (s % 64));
  atomicOr((unsigned long long*)&V[v * KW + s / 64], 1ull << (s % 64));
  level[(int64_t)s * n + v] = 0;
  const int64_t J = v / bs;
  atomicOr(&active[J >> 6], 1ull << (J & 63));
}

template <int BS, int KW>
static int msbfs_t(const gbo_bsr* A, const int32_t* h_src, int32_t nsrc, int32_t* d_level, int32_t* h_nlevels,
                   hipStream_t st) {
  const int32_t n = A->nrows;
  const int64_t words = (int64_t)std::max(n, 1) * KW;
  const int64_t nact = ((int64_t)A->nblockcols + 63) / 64 + 1;
  AsyncAllocs mem(st);
  uint64_t* buf = nullptr;
  int32_t* d_src = nullptr;
  int rc = mem.alloc(&buf, sizeof(uint64_t) * (3 * words + 2 * nact + 1), "gbo_bfs_multisource");
  if (rc || (rc = mem.alloc(&d_src, sizeof(int32_t) * nsrc, "gbo_bfs_multisource"))) return rc;
  uint64_t *F = buf, *Fn = buf + words, *V = buf + 2 * words;
  unsigned long long* act = (unsigned long long*)(buf + 3 * words);
  unsigned long long* act2 = act + nact;
  unsigned long long* cnt = (unsigned long long*)(buf + 3 * words + 2 * nact);
  GBO_HIP_CHECK(hipMemcpyAsync(d_src, h_src, sizeof(int32_t) * nsrc, hipMemcpyHostToDevice, st));
  GBO_HIP_CHECK(hipMemsetAsync(buf, 0, sizeof(uint64_t) * (3 * words + 2 * nact + 1), st));
  GBO_HIP_CHECK(hipMemsetAsync(d_level, 0xFF, sizeof(int32_t) * (size_t)nsrc * n, st));
  k_ms_init<KW><<<(nsrc + 255) / 256, 256, 0, st>>>(d_src, nsrc, F, V, d_level, n, A->bs, act);
  GBO_LAUNCH_CHECK();
  PlanGuard g(st);
  if ((rc = acquire_plan(A, &g.pv, st))) return rc;
  const PlanView& pv = g.pv;
  MsArgs a{};
  a.rowptr = A->rowptr;
  a.colind = A->colind;
  a.tiles = A->tiles;
  a.nrows = n;
  a.nbr = A->nblockrows;
  a.T = A->nblocks;
  a.plan = pv.plan;
  a.chunk = pv.chunk;
  a.nchunks = pv.nchunks;
  int32_t lvl = 0;
  while (true) {
    a.F = F;
    a.V = V;
    a.Fn = Fn;
    a.active = (const uint64_t*)act;
    if (a.nchunks > 1) {
      const int64_t np = (int64_t)(a.nchunks - 1) * BS;
      k_ms_prefill<BS, KW><<<(unsigned)((np + 255) / 256), 256, 0, st>>>(a);
    }
    if (BS == 8)
      k_msbfs8r<KW><<<(unsigned)(((int64_t)a.nchunks + 3) / 4), 256, 0, st>>>(a);
    else
      k_msbfs<BS, KW><<<(unsigned)(((int64_t)a.nchunks + 3) / 4), 256, 0, st>>>(a);
    GBO_LAUNCH_CHECK();
    GBO_HIP_CHECK(hipMemsetAsync(act2, 0, sizeof(unsigned long long) * nact, st));
    GBO_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), st));
    ++lvl;
    k_ms_update<KW><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(Fn, V, d_level, n, nsrc, A->bs, lvl, act2, cnt);
    GBO_LAUNCH_CHECK();
    unsigned long long h = 0;
    GBO_HIP_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
    GBO_HIP_CHECK(hipStreamSynchronize(st));
    if (h == 0) {
      --lvl;
      break;
    }
    std::swap(F, Fn);
    std::swap(act, act2);
  }
  if (h_nlevels) *h_nlevels = lvl + 1;
  return GBO_OK;
}

template <int BS>
static int msbfs_bs(const gbo_bsr* A, const int32_t* src, int32_t nsrc, int32_t* lv, int32_t* nl, hipStream_t st) {
  switch ((nsrc + 63) / 64) {
    case 1: return msbfs_t<BS, 1>(A, src, nsrc, lv, nl, st);
    case 2: return msbfs_t<BS, 2>(A, src, nsrc, lv, nl, st);
    case 3: return msbfs_t<BS, 3>(A, src, nsrc, lv, nl, st);
    default: return msbfs_t<BS, 4>(A, src, nsrc, lv, nl, st);
  }
}

}  // namespace gbo

using namespace gbo;

extern "C" {

int gbo_bfs_multisource(const gbo_bsr* AT, const int32_t* h_sources, int32_t nsrc, int32_t* d_level,
                        int32_t* h_nlevels, gbo_stream_t stream) {
  int rc = validate_bsr(AT, 0, "gbo_bfs_multisource");
  if (rc) return rc;
  if (AT->nrows != AT->ncols) return set_error(GBO_EINVAL, "gbo_bfs_multisource: matrix must be square");
  if (nsrc < 1 || nsrc > 256) return set_error(GBO_EINVAL, "gbo_bfs_multisource: 1..256 sources");
  if (!h_sources || !d_level) return set_error(GBO_EINVAL, "gbo_bfs_multisource: null argument");
  for (int32_t s = 0; s < nsrc; ++s)
    if (h_sources[s] < 0 || h_sources[s] >= AT->nrows)
      return set_error(GBO_EINVAL, "gbo_bfs_multisource: source %d out of range", h_sources[s]);
  switch (AT->bs) {
    case 8: return msbfs_bs<8>(AT, h_sources, nsrc, d_level, h_nlevels, stream);
    case 16: return msbfs_bs<16>(AT, h_sources, nsrc, d_level, h_nlevels, stream);
    case 32: return msbfs_bs<32>(AT, h_sources, nsrc, d_level, h_nlevels, stream);
    case 64: return msbfs_bs<64>(AT, h_sources, nsrc, d_level, h_nlevels, stream);
  }
  return set_error(GBO_EUNSUPPORTED, "gbo_bfs_multisource: block size");
}

}  // extern "C"
