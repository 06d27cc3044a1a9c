// Personalized PageRank over K vectors with a hand-written adjoint.
// No reference counterpart: it generalises the PageRank of algorithm/pr.cu
// (p' = alpha * P^T p + (1-alpha)/n, P = D^-1 A) to K personalization
// vectors s_k and adds the reverse-mode solve.
//
//   forward:  p_k = alpha * P^T p_k + (1 - alpha) * s_k      (pull over A^T)
//   adjoint:  lambda_k = alpha * P lambda_k + w_k            (pull over A)
//   dL/ds_k = (1 - alpha) * lambda_k,
//   dL/dalpha = sum_k lambda_k . (P^T p_k - s_k)
// for L = sum_k w_k . p_k.  The adjoint needs only the converged p (no
// checkpointing of the forward iterates).
//
// The product is a bit-BSR SpMM with a dense K-column right-hand side in
// vertex-major layout X[n][K]: every set bit (r, c) of a tile costs one
// 4K-byte row read X[j][0..K), taken by the wave together.  At bs 8 X
// carries the per-source scale (one elementwise pass).  With whole blocks of
// 64 columns (k_ppr_spmm8r) lane l holds tile row l / 8 and columns
// 8 (l % 8) .. +7: the set bits are listed per tile row and the eight rows'
// lists are walked in step, 32 bytes of a row read per lane.  Otherwise
// (k_ppr_spmm8) lane k owns column k, and the set bits are listed in LDS
// and read back 8 at a time.  Both add each row's terms in the same order,
// so they give the same bits.  At bs 16/32 (k_ppr_spmm) tile
// row words are fetched by lanes r < bs, the rows with bits are found by a
// ballot and set bits are read one at a time.  Each row's partial sum lives
// in register acc[r] (uniform index).  Waves take tile-balanced chunks;
// rows cut by a chunk boundary leave partials that a fix-up kernel adds in
// chunk order, so results are bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gbo_device.hpp"
#include "gbo_ops.hpp"

namespace gbo {

struct PprArgs {
  const int32_t* rowptr;
  const int32_t* colind;
  const void* tiles;
  int32_t nrows, nbr;
  int64_t T;
  const int32_t* plan;
  int32_t chunk, nchunks;
  int32_t K;                 // columns (<= 64 per launch; ld is the row stride)
  int32_t ld;                // row stride of X/Y/Z (total columns)
  const float* X;            // [n][ld] right-hand side (column block offset applied)
  const float* colscale;     // per source vertex factor on X (nullable -> 1)
  const float* rowscale;     // per output row factor (nullable -> 1)
  float a;                   // Y = a * rowscale * (A X) + b * Z
  float b;
  const float* Z;            // nullable
  float* Y;
  float* partial;            // [nchunks][2][bs][64] split-row partials (sum + compensation)
};

// First tile of chunk c (c <= nchunks; chunk nchunks starts at T).
__device__ __forceinline__ int64_t chunk_start(const PprArgs& A, int64_t c) {
  const int64_t b = c * A.chunk;
  return b < A.T ? b : A.T;
}

__device__ __forceinline__ void nadd(float& s, float& c, float v) {
  const float t = s + v;
  if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
}

template <int BS>
__device__ __forceinline__ void ppr_store_row(const PprArgs& A, int64_t R, int lane, const float* acc, const float* cmp,
                                              bool full, int slot, int c) {
  for (int r = 0; r < BS; ++r) {
    const int64_t i = R * BS + r;
    if (i >= A.nrows) break;
    const float v = acc[r] + cmp[r];
    if (full) {
      if (lane < A.K) {
        const float rs = A.rowscale ? A.rowscale[i] : 1.0f;
        float y = A.a * rs * v;
        if (A.Z) y += A.b * A.Z[i * A.ld + lane];
        A.Y[i * A.ld + lane] = y;
      }
    } else {
      A.partial[(((int64_t)c * 2 + slot) * BS + r) * kWave + lane] = v;
    }
  }
}

template <int BS>
__global__ __launch_bounds__(256) void k_ppr_spmm(const PprArgs A) {
  using W = word_t<BS>;
  const int lane = lane_id();
  const int c = wave_index();
  if (c >= A.nchunks) return;
  const int64_t cb = chunk_start(A, c);
  const int64_t ce = chunk_start(A, c + 1);
  const bool last = (c == A.nchunks - 1);
  const W* __restrict__ tiles = (const W*)A.tiles;
  const bool col_ok = lane < A.K;

  for (int32_t R0 = A.plan[c]; R0 < A.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < A.nbr) {
      rsl = A.rowptr[Rl];
      rel = A.rowptr[Rl + 1];
    }
    const bool belong = (Rl < A.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    uint64_t todo = bmask;  // empty rows are written with the epilogue too
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int64_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      float acc[BS], cmp[BS];
#pragma unroll
      for (int r = 0; r < BS; ++r) acc[r] = cmp[r] = 0.0f;
      for (int64_t t = s; t < te; ++t) {
        const int32_t col = A.colind[t];
        const uint64_t word = (lane < BS) ? (uint64_t)tiles[t * BS + lane] : 0ull;
        uint64_t rows = __ballot(word != 0);
        while (rows) {
          const int r = __builtin_ctzll(rows);
          rows &= rows - 1;
          uint64_t wr = (uint64_t)__shfl((long long)word, r);
          while (wr) {
            const int p = __builtin_ctzll(wr);
            wr &= wr - 1;
            const int64_t v = (int64_t)col * BS + (BS - 1 - p);
            float x = col_ok ? A.X[v * A.ld + lane] : 0.0f;
            if (A.colscale) x *= A.colscale[v];
            nadd(acc[r], cmp[r], x);  // r is wave-uniform
          }
        }
      }
      ppr_store_row<BS>(A, R, lane, acc, cmp, full, (rs < cb) ? 0 : 1, c);
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// bs 8 SpMM with batched row gathers.  Lanes are columns, so every set bit
// (r, c) of a tile is one coalesced row read X[v][0..K) for the whole wave.
// The bits are listed in parallel: lane i takes tile b + i of a 64-tile
// block of the row segment, a wave prefix sum of the tiles' popcounts gives
// each bit its slot, and each lane writes (r << 29 | v) for its bits into a
// per-wave LDS list (in tile order, then bit order, so the summation order
// is fixed).  The wave then walks the list 8 entries at a time, issuing the
// 8 row reads together.  Sums are FP64, one register per tile row indexed
// by the wave-uniform row.  X must already carry any per-source scale.
constexpr int kPprCap = 1024;  // list entries per wave and pass

__global__ __launch_bounds__(256) void k_ppr_spmm8(const PprArgs A) {
  constexpr int BS = 8, NB = 8;
  __shared__ uint32_t lists[256 / kWave][kPprCap];
  const int lane = lane_id();
  uint32_t* __restrict__ list = lists[threadIdx.x / kWave];
  const int c = wave_index();
  if (c >= A.nchunks) return;
  const int64_t cb = chunk_start(A, c);
  const int64_t ce = chunk_start(A, c + 1);
  const bool last = (c == A.nchunks - 1);
  const uint64_t* __restrict__ tw = (const uint64_t*)A.tiles;
  const bool col_ok = lane < A.K;

  for (int32_t R0 = A.plan[c]; R0 < A.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < A.nbr) {
      rsl = A.rowptr[Rl];
      rel = A.rowptr[Rl + 1];
    }
    const bool belong = (Rl < A.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    uint64_t todo = bmask;  // empty rows are written with the epilogue too
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int64_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      double acc[BS];
#pragma unroll
      for (int r = 0; r < BS; ++r) acc[r] = 0.0;
      for (int64_t b = s; b < te; b += kWave) {
        const int64_t nt = min((int64_t)kWave, te - b);
        const uint64_t word = lane < nt ? tw[b + lane] : 0ull;
        const uint32_t vb = lane < nt ? (uint32_t)A.colind[b + lane] * BS + (BS - 1) : 0u;
        const int cnt = __popcll(word);
        const int incl = wave_incl_scan(cnt);  // inclusive prefix sum of the popcounts over the lanes
        const int excl = incl - cnt;
        const int tot = __builtin_amdgcn_readlane(incl, kWave - 1);
        for (int base = 0; base < tot; base += kPprCap) {
          // list entries [base, base + kPprCap) of this block
          uint64_t w = word;
          int idx = excl;
          while (w) {
            const int p = __builtin_ctzll(w);
            w &= w - 1;
            if (idx >= base && idx < base + kPprCap)
              list[idx - base] = ((uint32_t)(p >> 3) << 29) | (vb - (uint32_t)(p & 7));  // row p/8, column 7 - p%8
            ++idx;
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          const int nl = min(kPprCap, tot - base);
          for (int i = 0; i < nl; i += NB) {
            uint32_t e[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) e[k] = __builtin_amdgcn_readfirstlane(list[min(i + k, nl - 1)]);
            // the NB row reads are unconditional (e[k] is a valid entry,
            // lanes past K read column 0) so they are in flight together;
            // conditional reads were each waited for in turn
            const int lc = col_ok ? lane : 0;
            float xk[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
              const float xv = A.X[(int64_t)(e[k] & 0x1FFFFFFFu) * A.ld + lc];
              xk[k] = (i + k < nl && col_ok) ? xv : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < NB; ++k)
              if (i + k < nl) acc[e[k] >> 29] += (double)xk[k];
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
      for (int r = 0; r < BS; ++r) {
        const int64_t i = R * BS + r;
        if (i >= A.nrows) break;
        if (full) {
          if (col_ok) {
            const double rsc = A.rowscale ? (double)A.rowscale[i] : 1.0;
            double y = (double)A.a * rsc * acc[r];
            if (A.Z) y += (double)A.b * (double)A.Z[i * A.ld + lane];
            A.Y[i * A.ld + lane] = (float)y;
          }
        } else {
          A.partial[(((int64_t)c * 2 + ((rs < cb) ? 0 : 1)) * BS + r) * kWave + lane] = (float)acc[r];
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// bs 8 SpMM for a block of 64 columns, lane = (tile row lane / 8, columns
// 8 (lane % 8) .. +7).  A 64-tile block's set bits are listed per tile row
// (eight LDS sublists; offsets from wave prefix sums of the rows' popcounts,
// two rows to a 32-bit word as 16-bit fields), and the eight rows' sublists
// are walked in step: each lane reads its row's next entry v and its 32
// bytes of row v of X (two 16-byte loads), and adds them into eight FP64
// registers of its own.  A row's entries keep the order of the flat list
// (tile, then bit), so the sums are those of k_ppr_spmm8, with no
// per-entry scalar work and no indexed register file access.
constexpr int kPprRowCap = 128;  // sublist entries per tile row and pass
// Entries per lane and step: four (eight 16-byte loads in flight per lane)
// measured 8.2 ms per R-MAT s22 K = 64 iteration against 8.4 with two, at
// 4 waves per SIMD (96-110 VGPRs); the flat-list kernel took 12.1.
constexpr int kPprRowStep = 4;

__device__ __forceinline__ uint32_t field16(uint32_t w, int hi) { return hi ? (w >> 16) : (w & 0xFFFFu); }

__global__ __launch_bounds__(256) void k_ppr_spmm8r(const PprArgs A) {
  constexpr int BS = 8, U = kPprRowStep, CAP = kPprRowCap;
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  __shared__ uint32_t lists[256 / kWave][BS * CAP];
  const int lane = lane_id();
  const int row = lane >> 3, cg = lane & 7;
  uint32_t* __restrict__ list = lists[threadIdx.x / kWave];
  const int c = wave_index();
  if (c >= A.nchunks) return;
  const int64_t cb = chunk_start(A, c);
  const int64_t ce = chunk_start(A, c + 1);
  const bool last = (c == A.nchunks - 1);
  const uint64_t* __restrict__ tw = (const uint64_t*)A.tiles;

  for (int32_t R0 = A.plan[c]; R0 < A.nbr;) {
    const int32_t Rl = R0 + lane;
    int32_t rsl = 0, rel = 0;
    if (Rl < A.nbr) {
      rsl = A.rowptr[Rl];
      rel = A.rowptr[Rl + 1];
    }
    const bool belong = (Rl < A.nbr) && (last || (int64_t)rsl < ce);
    const uint64_t bmask = __ballot(belong);
    uint64_t todo = bmask;  // empty rows are written with the epilogue too
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int64_t R = R0 + j;
      const int64_t rs = __builtin_amdgcn_readlane(rsl, j), re = __builtin_amdgcn_readlane(rel, j);
      const int64_t s = rs > cb ? rs : cb;
      const int64_t te = re < ce ? re : ce;
      const bool full = (rs >= cb) && (re <= ce);
      double acc[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m] = 0.0;
      for (int64_t b = s; b < te; b += kWave) {
        const int64_t nt = min((int64_t)kWave, te - b);
        const uint64_t word = lane < nt ? tw[b + lane] : 0ull;
        const uint32_t vb = lane < nt ? (uint32_t)A.colind[b + lane] * BS + (BS - 1) : 0u;
        int maxlen = 0, mylen = 0;
        {
          // row totals of the block (sublist lengths)
          uint32_t tot[4];
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            const uint32_t qh = (uint32_t)__popc((uint32_t)(word >> (16 * h)) & 0xFFu) |
                                ((uint32_t)__popc((uint32_t)(word >> (16 * h + 8)) & 0xFFu) << 16);
            tot[h] = __builtin_amdgcn_readlane((uint32_t)wave_incl_scan((int)qh), kWave - 1);
          }
#pragma unroll
          for (int r = 0; r < BS; ++r) {
            const int l = (int)field16(tot[r >> 1], r & 1);
            maxlen = l > maxlen ? l : maxlen;
            mylen = (row == r) ? l : mylen;
          }
        }
        for (int pass = 0; pass < maxlen; pass += CAP) {
          // this lane's bits of every row whose sublist index is in the pass
          // (offsets recomputed per pass, so that they are not live across
          // the gathers)
#pragma unroll
          for (int r = 0; r < BS; ++r) {
            uint32_t byte = (uint32_t)(word >> (8 * r)) & 0xFFu;
            const uint32_t qr = (uint32_t)__popc(byte);
            int k = wave_incl_scan((int)qr) - (int)qr - pass;
            while (byte) {
              const int bit = __builtin_ctz(byte);  // column 7 - bit
              byte &= byte - 1;
              if (k >= 0 && k < CAP) list[r * CAP + k] = vb - (uint32_t)bit;
              ++k;
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          const int lr = min(max(mylen - pass, 0), CAP);
          const int kmax = min(maxlen - pass, CAP);
          for (int k0 = 0; k0 < kmax; k0 += U) {
            f32x4 xa[U], xb[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              ok[u] = k0 + u < lr;
              const uint32_t v = ok[u] ? list[row * CAP + k0 + u] : 0u;  // row 0 stands in
              const f32x4* px = (const f32x4*)(A.X + (int64_t)v * A.ld + cg * 8);
              xa[u] = px[0];
              xb[u] = px[1];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
              if (ok[u]) {
                acc[0] += (double)xa[u].x;
                acc[1] += (double)xa[u].y;
                acc[2] += (double)xa[u].z;
                acc[3] += (double)xa[u].w;
                acc[4] += (double)xb[u].x;
                acc[5] += (double)xb[u].y;
                acc[6] += (double)xb[u].z;
                acc[7] += (double)xb[u].w;
              }
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
      const int64_t i = R * BS + row;
      if (i < A.nrows) {
        if (full) {
          const double rsc = A.rowscale ? (double)A.rowscale[i] : 1.0;
          f32x4 za = {0.f, 0.f, 0.f, 0.f}, zb = za;
          if (A.Z) {
            const f32x4* pz = (const f32x4*)(A.Z + i * A.ld + cg * 8);
            za = pz[0];
            zb = pz[1];
          }
          const float zz[8] = {za.x, za.y, za.z, za.w, zb.x, zb.y, zb.z, zb.w};
          float y[8];
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            double t = (double)A.a * rsc * acc[m];
            if (A.Z) t += (double)A.b * (double)zz[m];
            y[m] = (float)t;
          }
          f32x4* py = (f32x4*)(A.Y + i * A.ld + cg * 8);
          py[0] = f32x4{y[0], y[1], y[2], y[3]};
          py[1] = f32x4{y[4], y[5], y[6], y[7]};
        } else {
          float* pp = A.partial + (((int64_t)c * 2 + ((rs < cb) ? 0 : 1)) * BS + row) * kWave + cg * 8;
#pragma unroll
          for (int m = 0; m < 8; ++m) pp[m] = (float)acc[m];
        }
      }
    }
    if (__popcll(bmask) < kWave) break;
    R0 += kWave;
  }
}

// Xs[i][k] = X[i][k] * scale[i] over K columns of a row stride ld.
__global__ __launch_bounds__(256) void k_row_scale(const float* __restrict__ X, const float* __restrict__ scale,
                                                   float* __restrict__ Xs, int32_t n, int32_t K, int32_t ld) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * K) return;
  const int64_t i = idx / K, k = idx % K;
  Xs[i * ld + k] = X[i * ld + k] * scale[i];
}

template <int BS>
__global__ __launch_bounds__(256) void k_ppr_fixup(const PprArgs A) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)(idx % kWave);
  const int64_t cr = idx / kWave;  // (chunk, row) pairs
  const int64_t c = cr / BS;
  const int r = (int)(cr % BS);
  if (c + 1 >= A.nchunks) return;
  const int64_t nb = chunk_start(A, c + 1);
  const int32_t R = A.plan[c + 1];
  const int64_t rs = A.rowptr[R], re = A.rowptr[R + 1];
  if (!(rs < nb) || rs < chunk_start(A, c)) return;
  float s = A.partial[((c * 2 + 1) * BS + r) * kWave + lane], cmp = 0.0f;
  for (int64_t c2 = c + 1; c2 < A.nchunks && chunk_start(A, c2) < re; ++c2)
    nadd(s, cmp, A.partial[((c2 * 2 + 0) * BS + r) * kWave + lane]);
  const int64_t i = (int64_t)R * BS + r;
  if (i < A.nrows && lane < A.K) {
    const float rsc = A.rowscale ? A.rowscale[i] : 1.0f;
    float y = A.a * rsc * (s + cmp);
    if (A.Z) y += A.b * A.Z[i * A.ld + lane];
    A.Y[i * A.ld + lane] = y;
  }
}

// Deterministic column-wise reduction over a [n][ld] column block of
// `cols` <= 64 columns; mode 0: sum |Y - P|, mode 1: sum L * (Y - P),
// mode 2: sum |Y|.  Each block writes 64 partials; k_colred_final adds
// them in a fixed order.
constexpr int kColredBlocks = 2048;  // fixed, so the reduction order is the same on every GPU

__global__ __launch_bounds__(256) void k_colred(const float* __restrict__ Y, const float* __restrict__ P,
                                                const float* __restrict__ L, int32_t n, int32_t ld, int32_t cols,
                                                int mode, double* __restrict__ part) {
  const int col = threadIdx.x & 63, rsub = threadIdx.x >> 6;
  // four rows per step into four accumulators (loads of a step in flight
  // together), combined in a fixed order: deterministic for a fixed grid
  double acc4[4] = {0.0, 0.0, 0.0, 0.0};
  if (col < cols) {
    const int64_t S = (int64_t)gridDim.x * 4;
    for (int64_t i0 = (int64_t)blockIdx.x * 4 + rsub; i0 < n; i0 += 4 * S) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t i = i0 + q * S;
        if (i < n) {
          const int64_t o = i * ld + col;
          if (mode == 0) acc4[q] += fabs((double)Y[o] - (double)P[o]);
          else if (mode == 1) acc4[q] += (double)L[o] * ((double)Y[o] - (double)P[o]);
          else acc4[q] += fabs((double)Y[o]);
        }
      }
    }
  }
  const double acc = (acc4[0] + acc4[1]) + (acc4[2] + acc4[3]);
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64)
    part[(int64_t)blockIdx.x * 64 + threadIdx.x] =
        sh[threadIdx.x] + sh[threadIdx.x + 64] + sh[threadIdx.x + 128] + sh[threadIdx.x + 192];
}

// One block of 1024 threads: 16 per column, thread (s, k) adds the partials
// of blocks s, s+16, s+32, ... in order (eight loads in flight), then the
// 16 sums of a column are added in s order -- a fixed order for a fixed nb.
constexpr int kColredFinalThreads = 1024;
__global__ __launch_bounds__(kColredFinalThreads) void k_colred_final(const double* __restrict__ part, int nb,
                                                                      double* out) {
  constexpr int NS = kColredFinalThreads / 64;
  __shared__ double sh[NS][64];
  const int k = threadIdx.x & 63, sidx = threadIdx.x >> 6;
  double s = 0.0;
  int b = sidx;
  for (; b + 7 * NS < nb; b += 8 * NS) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[(int64_t)(b + u * NS) * 64 + k];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; b < nb; b += NS) s += part[(int64_t)b * 64 + k];
  sh[sidx][k] = s;
  __syncthreads();
  if (sidx == 0) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < NS; ++q) t += sh[q][k];
    out[k] = t;
  }
}

__global__ void k_invdeg(const int32_t* __restrict__ deg, float* __restrict__ inv, int32_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[i] = deg[i] > 0 ? 1.0f / (float)deg[i] : 0.0f;
}

__global__ void k_scale_copy(const float* __restrict__ src, float* __restrict__ dst, int64_t n, float a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = a * src[i];
}

// Plans with a chunk sized for the SpMM (one tile costs up to bs*K loads).
// The SpMM's work plan and scratch, released on every return path.
struct PprPlan {
  int32_t* plan = nullptr;
  int32_t chunk = 0, nchunks = 0;
  float* partial = nullptr;
  float* xs = nullptr;  // bs 8: X with the per-source scale applied, [n][K]
  int64_t xs_len = 0;
  hipStream_t st = nullptr;
  explicit PprPlan(hipStream_t s) : st(s) {}
  ~PprPlan() {
    if (plan) (void)hipFreeAsync(plan, st);
    if (partial) (void)hipFreeAsync(partial, st);
    if (xs) (void)hipFreeAsync(xs, st);
  }
  PprPlan(const PprPlan&) = delete;
  PprPlan& operator=(const PprPlan&) = delete;
};

__global__ void k_plan_ppr(const int32_t* __restrict__ rowptr, int32_t nbr, int32_t chunk, int32_t nchunks,
                           int32_t* __restrict__ plan) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t b = (int64_t)c * chunk;
  int32_t lo = 0, hi = nbr;
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if ((int64_t)rowptr[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  int32_t R = lo;
  if ((int64_t)rowptr[R] > b) R -= 1;
  if (R >= nbr) R = nbr > 0 ? nbr - 1 : 0;
  if (R < 0) R = 0;
  plan[c] = R;
}

// k_row_scale for K % 4 == 0 and 16-byte aligned X and Xs (ld == K):
// four columns per thread, n * K < 2^32.
__global__ __launch_bounds__(256) void k_row_scale4(const float* __restrict__ X, const float* __restrict__ scale,
                                                    float* __restrict__ Xs, uint32_t n4, uint32_t K4) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n4) return;
  const uint32_t i = t / K4;
  ((f32x4*)Xs)[t] = ((const f32x4*)X)[t] * scale[i];
}

static int make_ppr_plan(const gbo_bsr* M, PprPlan* P, hipStream_t st) {
  P->chunk = 1024;
  P->nchunks = (int32_t)std::max<int64_t>(1, (M->nblocks + P->chunk - 1) / P->chunk);
  GBO_HIP_CHECK(hipMallocAsync((void**)&P->plan, sizeof(int32_t) * P->nchunks, st));
  GBO_HIP_CHECK(hipMallocAsync((void**)&P->partial, sizeof(float) * 2 * (size_t)P->nchunks * M->bs * kWave, st));
  k_plan_ppr<<<(P->nchunks + 255) / 256, 256, 0, st>>>(M->rowptr, M->nblockrows, P->chunk, P->nchunks, P->plan);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

// Y = a * rowscale * (M (colscale .* X)) + b * Z over all K columns.
static int spmm(const gbo_bsr* M, PprPlan& P, int32_t K, const float* X, const float* colscale,
                const float* rowscale, float a, float b, const float* Z, float* Y, hipStream_t st) {
  if (M->bs == 8 && M->ncols > (1 << 29))
    return set_error(GBO_EUNSUPPORTED, "PPR at block size 8 supports up to 2^29 vertices");
  if (M->bs == 8 && colscale) {
    // the bs 8 kernel reads pre-scaled rows
    const int64_t len = (int64_t)M->ncols * K;
    if (P.xs_len < len) {
      if (P.xs) GBO_HIP_CHECK(hipFreeAsync(P.xs, st));
      P.xs = nullptr;
      GBO_HIP_CHECK(hipMallocAsync((void**)&P.xs, sizeof(float) * std::max<int64_t>(len, 1), st));
      P.xs_len = len;
    }
    if (len > 0) {
      if (K % 4 == 0 && len < ((int64_t)1 << 32) && ((((uintptr_t)X) | ((uintptr_t)P.xs)) & 15) == 0) {
        const uint32_t n4 = (uint32_t)(len / 4);
        k_row_scale4<<<(n4 + 255) / 256, 256, 0, st>>>(X, colscale, P.xs, n4, (uint32_t)(K / 4));
      } else {
        k_row_scale<<<(unsigned)((len + 255) / 256), 256, 0, st>>>(X, colscale, P.xs, M->ncols, K, K);
      }
      GBO_LAUNCH_CHECK();
    }
    X = P.xs;
    colscale = nullptr;
  }
  // bs 8 with whole 64-column blocks and 16-byte aligned rows: the per-row
  // sublist kernel (the same sums as the flat-list one, fewer instructions)
  const bool rows8 = M->bs == 8 && K % kWave == 0 && ((((uintptr_t)X) | ((uintptr_t)Y) | ((uintptr_t)Z)) & 15) == 0;
  for (int32_t k0 = 0; k0 < K; k0 += kWave) {
    PprArgs A{};
    A.rowptr = M->rowptr;
    A.colind = M->colind;
    A.tiles = M->tiles;
    A.nrows = M->nrows;
    A.nbr = M->nblockrows;
    A.T = M->nblocks;
    A.plan = P.plan;
    A.chunk = P.chunk;
    A.nchunks = P.nchunks;
    A.K = std::min(kWave, K - k0);
    A.ld = K;
    A.X = X + k0;
    A.colscale = colscale;
    A.rowscale = rowscale;
    A.a = a;
    A.b = b;
    A.Z = Z ? Z + k0 : nullptr;
    A.Y = Y + k0;
    A.partial = P.partial;
    const unsigned blocks = (unsigned)(((int64_t)P.nchunks + 3) / 4);
    switch (M->bs) {
      case 8: {
        if (rows8) k_ppr_spmm8r<<<blocks, 256, 0, st>>>(A);
        else k_ppr_spmm8<<<blocks, 256, 0, st>>>(A);
        break;
      }
      case 16: k_ppr_spmm<16><<<blocks, 256, 0, st>>>(A); break;
      case 32: k_ppr_spmm<32><<<blocks, 256, 0, st>>>(A); break;
      default: return set_error(GBO_EUNSUPPORTED, "PPR supports block sizes 8, 16, 32");
    }
    GBO_LAUNCH_CHECK();
    if (P.nchunks > 1) {
      const int64_t thr = (int64_t)(P.nchunks - 1) * M->bs * kWave;
      const unsigned fb = (unsigned)((thr + 255) / 256);
      switch (M->bs) {
        case 8: k_ppr_fixup<8><<<fb, 256, 0, st>>>(A); break;
        case 16: k_ppr_fixup<16><<<fb, 256, 0, st>>>(A); break;
        case 32: k_ppr_fixup<32><<<fb, 256, 0, st>>>(A); break;
      }
      GBO_LAUNCH_CHECK();
    }
  }
  return GBO_OK;
}

// Column-wise reduction over [n][K] in column blocks of 64 (deterministic).
static int colreduce(const float* Y, const float* P, const float* L, int32_t n, int32_t K, int mode, double* d_part,
                     double* d_out, double* h_out, hipStream_t st) {
  constexpr int nb = kColredBlocks;
  for (int32_t k0 = 0; k0 < K; k0 += 64) {
    const int32_t cols = std::min(64, K - k0);
    k_colred<<<nb, 256, 0, st>>>(Y + k0, P ? P + k0 : nullptr, L ? L + k0 : nullptr, n, K, cols, mode, d_part);
    GBO_LAUNCH_CHECK();
    k_colred_final<<<1, kColredFinalThreads, 0, st>>>(d_part, nb, d_out);
    GBO_LAUNCH_CHECK();
    GBO_HIP_CHECK(hipMemcpyAsync(h_out + k0, d_out, sizeof(double) * cols, hipMemcpyDeviceToHost, st));
  }
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  return GBO_OK;
}

// max over columns of delta[k] / norm[k] (delta[k] itself where norm[k] is 0)
static double max_relative(const std::vector<double>& delta, const std::vector<double>& norm) {
  double m = 0.0;
  for (size_t k = 0; k < delta.size(); ++k) m = std::max(m, norm[k] > 0.0 ? delta[k] / norm[k] : delta[k]);
  return m;
}

}  // namespace gbo

using namespace gbo;

extern "C" {

int gbo_ppr_forward(const gbo_bsr* AT, const int32_t* d_outdeg, int32_t K, float alpha, const float* d_s, float* d_p,
                    double tol, int32_t max_iter, int32_t* h_iters, double* h_delta, gbo_stream_t stream) {
  int rc = validate_bsr(AT, 0, "gbo_ppr_forward");
  if (rc) return rc;
  if (AT->nrows != AT->ncols) return set_error(GBO_EINVAL, "gbo_ppr_forward: matrix must be square");
  if (K < 1 || K > 64) return set_error(GBO_EINVAL, "gbo_ppr_forward: 1 <= K <= 64");
  if (!(alpha > 0.0f && alpha < 1.0f) || max_iter < 1) return set_error(GBO_EINVAL, "gbo_ppr_forward: bad alpha/max_iter");
  if (!d_outdeg || !d_s || !d_p) return set_error(GBO_EINVAL, "gbo_ppr_forward: null argument");
  if (AT->bs > 32) return set_error(GBO_EUNSUPPORTED, "gbo_ppr_forward: block sizes 8, 16, 32");
  hipStream_t st = stream;
  const int32_t n = AT->nrows;
  const int64_t nk = (int64_t)n * K;
  AsyncAllocs mem(st);
  float *inv = nullptr, *q = nullptr;
  double* dpart = nullptr;
  if ((rc = mem.alloc(&inv, sizeof(float) * std::max(n, 1), "gbo_ppr_forward")) ||
      (rc = mem.alloc(&q, sizeof(float) * std::max<int64_t>(nk, 1), "gbo_ppr_forward")) ||
      (rc = mem.alloc(&dpart, sizeof(double) * (kColredBlocks * 64 + 64), "gbo_ppr_forward")))
    return rc;
  k_invdeg<<<(n + 255) / 256, 256, 0, st>>>(d_outdeg, inv, n);
  GBO_LAUNCH_CHECK();
  PprPlan P(st);
  if ((rc = make_ppr_plan(AT, &P, st))) return rc;
  // p_0 = s
  GBO_HIP_CHECK(hipMemcpyAsync(d_p, d_s, sizeof(float) * nk, hipMemcpyDeviceToDevice, st));
  std::vector<double> delta(K), snorm(K);
  // the change is measured relative to ||s_k||_1 (1 for a distribution)
  if ((rc = colreduce(d_s, nullptr, nullptr, n, K, 2, dpart, dpart + kColredBlocks * 64, snorm.data(), st))) return rc;
  int32_t it = 0;
  double dmax = 0.0;
  float* cur = d_p;
  float* nxt = q;
  while (it < max_iter) {
    if ((rc = spmm(AT, P, K, cur, inv, nullptr, alpha, 1.0f - alpha, d_s, nxt, st))) return rc;
    if ((rc = colreduce(nxt, cur, nullptr, n, K, 0, dpart, dpart + kColredBlocks * 64, delta.data(), st))) return rc;
    ++it;
    std::swap(cur, nxt);
    dmax = max_relative(delta, snorm);
    if (dmax < tol) break;
  }
  if (cur != d_p) GBO_HIP_CHECK(hipMemcpyAsync(d_p, cur, sizeof(float) * nk, hipMemcpyDeviceToDevice, st));
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if (h_iters) *h_iters = it;
  if (h_delta) *h_delta = dmax;
  if (!(dmax < tol))
    return set_error(GBO_ENOCONV, "gbo_ppr_forward: not converged after %d iterations (%.3e)", it, dmax);
  return GBO_OK;
}

int gbo_ppr_adjoint(const gbo_bsr* A, const gbo_bsr* AT, const int32_t* d_outdeg, int32_t K, float alpha,
                    const float* d_s, const float* d_p, const float* d_w, float* d_grad_s, double* h_grad_alpha,
                    double tol, int32_t max_iter, int32_t* h_iters, gbo_stream_t stream) {
  int rc = validate_bsr(A, 0, "gbo_ppr_adjoint(A)");
  if (rc) return rc;
  rc = validate_bsr(AT, 0, "gbo_ppr_adjoint(AT)");
  if (rc) return rc;
  if (A->nrows != A->ncols || AT->nrows != A->nrows) return set_error(GBO_EINVAL, "gbo_ppr_adjoint: shapes");
  if (K < 1 || K > 64) return set_error(GBO_EINVAL, "gbo_ppr_adjoint: 1 <= K <= 64");
  if (!(alpha > 0.0f && alpha < 1.0f) || max_iter < 1) return set_error(GBO_EINVAL, "gbo_ppr_adjoint: bad alpha/max_iter");
  if (!d_outdeg || !d_s || !d_p || !d_w || !d_grad_s) return set_error(GBO_EINVAL, "gbo_ppr_adjoint: null argument");
  if (A->bs > 32 || AT->bs > 32) return set_error(GBO_EUNSUPPORTED, "gbo_ppr_adjoint: block sizes 8, 16, 32");
  hipStream_t st = stream;
  const int32_t n = A->nrows;
  const int64_t nk = (int64_t)n * K;
  AsyncAllocs mem(st);
  float *inv = nullptr, *lam = nullptr;
  double* dpart = nullptr;
  if ((rc = mem.alloc(&inv, sizeof(float) * std::max(n, 1), "gbo_ppr_adjoint")) ||
      (rc = mem.alloc(&lam, sizeof(float) * std::max<int64_t>(nk, 1) * 2, "gbo_ppr_adjoint")) ||
      (rc = mem.alloc(&dpart, sizeof(double) * (kColredBlocks * 64 + 64), "gbo_ppr_adjoint")))
    return rc;
  float* q = lam + nk;
  k_invdeg<<<(n + 255) / 256, 256, 0, st>>>(d_outdeg, inv, n);
  GBO_LAUNCH_CHECK();
  PprPlan PA(st), PT(st);
  if ((rc = make_ppr_plan(A, &PA, st)) || (rc = make_ppr_plan(AT, &PT, st))) return rc;
  GBO_HIP_CHECK(hipMemcpyAsync(lam, d_w, sizeof(float) * nk, hipMemcpyDeviceToDevice, st));
  std::vector<double> delta(K), ga(K), wnorm(K);
  // the change is measured relative to ||w_k||_1, the scale of lambda_k
  if ((rc = colreduce(d_w, nullptr, nullptr, n, K, 2, dpart, dpart + kColredBlocks * 64, wnorm.data(), st))) return rc;
  int32_t it = 0;
  float* cur = lam;
  float* nxt = q;
  double dmax = 0.0;
  while (it < max_iter) {
    // lambda' = alpha * D^-1 A lambda + w
    if ((rc = spmm(A, PA, K, cur, nullptr, inv, alpha, 1.0f, d_w, nxt, st))) return rc;
    if ((rc = colreduce(nxt, cur, nullptr, n, K, 0, dpart, dpart + kColredBlocks * 64, delta.data(), st))) return rc;
    ++it;
    std::swap(cur, nxt);
    dmax = max_relative(delta, wnorm);
    if (dmax < tol) break;
  }
  // grad_s = (1 - alpha) * lambda
  k_scale_copy<<<(unsigned)((nk + 255) / 256), 256, 0, st>>>(cur, d_grad_s, nk, 1.0f - alpha);
  GBO_LAUNCH_CHECK();
  // dL/dalpha = sum_k lambda_k . (P^T p_k - s_k)
  if ((rc = spmm(AT, PT, K, d_p, inv, nullptr, 1.0f, 0.0f, nullptr, nxt, st))) return rc;
  if ((rc = colreduce(nxt, d_s, cur, n, K, 1, dpart, dpart + kColredBlocks * 64, ga.data(), st))) return rc;
  if (h_grad_alpha) {
    double s = 0.0;
    for (double v : ga) s += v;
    *h_grad_alpha = s;
  }
  GBO_HIP_CHECK(hipStreamSynchronize(st));
  if (h_iters) *h_iters = it;
  if (!(dmax < tol))
    return set_error(GBO_ENOCONV, "gbo_ppr_adjoint: not converged after %d iterations (%.3e)", it, dmax);
  return GBO_OK;
}

}  // extern "C"
