/*
 * gbo.h -- C ABI of the MI355X (gfx950) bit-packed block-sparse graph engine.
 *
 * The reference (carolc0708/graph-bit-ops) has no library ABI: its "API" is
 * source-level -- each driver #includes the packer and kernel files and one
 * binary is built per block size.  Every entry point below replaces one such
 * source-level interface; the comment on each names the reference symbol
 * (file:line under the reference tree) whose behaviour it reproduces.
 *
 * Conventions (all entry points):
 *  - return 0 (GBO_OK) on success or a negative GBO_E* code; the message of
 *    the last failure on the calling thread is available from
 *    gbo_last_error().  No C++ exception crosses this boundary.
 *  - d_* arguments are device pointers, h_* are host pointers.  The caller
 *    owns every buffer unless a create/free pair of functions says otherwise.
 *  - calls taking a gbo_stream_t are stream-ordered and asynchronous unless
 *    documented otherwise; pass NULL for the default stream.
 *  - tile offsets are computed in 64-bit (the reference's int offsets
 *    overflow once nblocks*bs > 2^31).
 *
 * Bit layout (reproduces ToBit{N}Col / ToBit{N}Row, algorithm/bsrbmv.cu:88-250):
 *  - tile t occupies words [t*bs, (t+1)*bs) of `tiles`; word type is uint8,
 *    uint16, uint32, uint64 for bs 8/16/32/64.
 *  - dir 0 ("row", CUSPARSE_DIRECTION_ROW in csr2bsr_batch_*): word r is row
 *    r of the tile and bit (bs-1-c) is column c (MSB = column 0).
 *  - dir 1 ("col", bsrbmm/csr2bsr_batch.cu *_col): word k is column k and bit
 *    (bs-1-r) is row r.
 *  - bit vectors: one bs-bit word per block, bit (bs-1-k) = element j*bs+k.
 *  - colind is ascending inside each block row; a tile exists iff it holds
 *    at least one nonzero; duplicate CSR entries OR together.
 */
#ifndef GBO_H
#define GBO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* gbo_stream_t; /* == hipStream_t */

enum {
  GBO_OK = 0,
  GBO_EINVAL = -1,       /* bad argument */
  GBO_ENOMEM = -2,       /* host or device allocation failed */
  GBO_EDEVICE = -3,      /* HIP runtime error */
  GBO_EIO = -4,          /* file not found / unreadable */
  GBO_EFORMAT = -5,      /* malformed Matrix Market input */
  GBO_ERANGE = -6,       /* size exceeds a documented limit */
  GBO_ENOCONV = -7,      /* iteration cap reached before convergence */
  GBO_EUNSUPPORTED = -8  /* block size / direction not supported */
};

/* Bit-BSR matrix (device resident).  rowptr[nblockrows+1], colind[nblocks],
 * tiles[nblocks*bs] words.  `plan`, `plan_chunk`, `plan_nchunks` are an
 * optional work decomposition made by gbo_bsr_plan(); leave them zero and
 * every operation builds a temporary one. */
typedef struct gbo_bsr {
  int32_t bs;           /* 8, 16, 32 or 64 */
  int32_t dir;          /* 0 = row words, 1 = column words */
  int32_t nrows, ncols;
  int32_t nblockrows, nblockcols;
  int64_t nblocks;
  int32_t* rowptr;
  int32_t* colind;
  void* tiles;
  int32_t* plan;        /* device, plan_nchunks entries (first block row of each chunk) */
  int32_t plan_chunk;   /* tiles per chunk */
  int32_t plan_nchunks;
  int32_t plan_variant; /* kernel variant for bs 8/16 (0 = default; see gbo_bsr_plan_ex) */
} gbo_bsr;

const char* gbo_last_error(void);
const char* gbo_version(void);

/* ---------------------------------------------------------------- memory --
 * Thin wrappers over the HIP runtime so that hosts without their own GPU
 * runtime bindings (ctypes, the CLIs) can move data. */
int gbo_device_count(int* count);
int gbo_set_device(int device);
int gbo_malloc(void** d_ptr, size_t bytes);
int gbo_free(void* d_ptr);
int gbo_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes);
int gbo_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes);
int gbo_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes);
int gbo_memset(void* d_ptr, int value, size_t bytes);
int gbo_synchronize(void);
int gbo_stream_create(gbo_stream_t* stream);
int gbo_stream_destroy(gbo_stream_t stream);
int gbo_stream_synchronize(gbo_stream_t stream);
/* Events for timing kernels on a given stream (milliseconds). */
int gbo_event_create(void** ev);
int gbo_event_destroy(void* ev);
int gbo_event_record(void* ev, gbo_stream_t stream);
int gbo_event_elapsed(float* ms, void* start, void* stop);
void gbo_free_host(void* h_ptr);
/* Plain device-to-device copy kernel (16-byte vector loads/stores; size and
 * pointers 16-byte aligned): measures the HBM copy ceiling reported next to
 * the roofline and calibrates the FETCH_SIZE/WRITE_SIZE counters.  No
 * reference counterpart. */
int gbo_stream_copy(void* d_dst, const void* d_src, size_t bytes, gbo_stream_t stream);
/* Read-only stream over d_src (16-byte aligned): every word is loaded once
 * and XOR-folded per thread; a thread stores its fold to *d_sink only when
 * it equals `never` (pick a value the fold cannot take for the data, e.g.
 * 0xFFFFFFFF over a buffer of equal bytes), so nothing is written in
 * practice.  Measures the HBM read ceiling.  No reference counterpart. */
int gbo_stream_read(const void* d_src, size_t bytes, uint32_t never, uint32_t* d_sink, gbo_stream_t stream);
/* Gather probe: reads table[idx[i]] for i in [0, n) (elements of 1 or 4
 * bytes; idx int32, 16-byte aligned, every index inside the table), four
 * gathers per lane in flight, XOR-folded as in gbo_stream_read; with
 * elem_bytes 256 a wave reads each 256-byte row idx[i] (lane l its l-th
 * word), eight rows in flight.  Measures the random-gather rate a product's
 * vector gathers meet on this GPU (fed with the product's own column
 * indices).  No reference counterpart. */
int gbo_gather_probe(const void* d_table, int elem_bytes, const int32_t* d_idx, int64_t n, uint32_t never,
                     uint32_t* d_sink, gbo_stream_t stream);

/* ------------------------------------------------- Matrix Market readers --
 * semantics 0: mmio_info + mmio_data (algorithm/mmio_highlevel.h:7,146):
 *   symmetric/hermitian banners mirror off-diagonal entries, duplicates and
 *   self-loops are kept, entries of a row stay in file order.
 * semantics 1: readMtx + coo2csr (algorithm/readMtx.hpp:352,247): symmetrize
 *   when (banner symmetric and directed != 1) or directed == 2, sort by
 *   (row, col), drop duplicates, drop self-loops unless the environment
 *   variable GRB_UTIL_REMOVE_SELFLOOP is "0" (readMtx.hpp:151-216).
 * Output arrays are allocated by the library; release with gbo_free_host.
 * h_val may be NULL when the values are not wanted.  The file is read whole
 * and parsed on all host threads (token-parallel; the result is the same
 * as the reference's sequential fscanf loops: tokens pair into entries
 * across lines, tokens past the header's entry count are ignored). */
int gbo_mtx_read(const char* path, int semantics, int directed,
                 int32_t* nrows, int32_t* ncols, int64_t* nnz,
                 int32_t** h_rowptr, int32_t** h_colind, float** h_val);
/* gbo_mtx_read with a binary CSR cache beside the file, the cache that
 * readMtx names but never writes (algorithm/readMtx.hpp:227-244,390-395):
 * semantics 1 uses readMtx's own name "<dir>/.<file>.<ud|d>.<nosl|sl>.bin",
 * semantics 0 "<dir>/.<file>.mmio.bin".  The cache records the source's size
 * and modification time, the semantics, the undirected flag and the
 * self-loop rule; a cache that does not match all of them is ignored and
 * rewritten (through a temporary file renamed into place).  *cache_state
 * (may be NULL): 1 = loaded from the cache, 0 = parsed and cache written,
 * 2 = parsed, cache not writable (not an error). */
int gbo_mtx_read_cached(const char* path, int semantics, int directed, int* cache_state,
                        int32_t* nrows, int32_t* ncols, int64_t* nnz,
                        int32_t** h_rowptr, int32_t** h_colind, float** h_val);
/* The cache file name gbo_mtx_read_cached uses for (path, semantics,
 * directed), NUL-terminated into buf[len]. */
int gbo_mtx_cache_path(const char* path, int semantics, int directed, char* buf, size_t len);
/* Writes a pattern general .mtx file of a CSR (1-based); rows are formatted
 * on all host threads. */
int gbo_mtx_write(const char* path, int32_t nrows, int32_t ncols,
                  const int32_t* h_rowptr, const int32_t* h_colind);

/* ------------------------------------------------ synthetic generators --
 * All generators are counter-based (SplitMix64 finalizer over
 * (seed, stream, index)) so results do not depend on thread count.  Outputs
 * are CSR with rows sorted ascending, no duplicates, no self-loops, and are
 * symmetric.  Arrays are allocated by the library (gbo_free_host). */
int gbo_gen_er(int32_t n, double p, uint64_t seed,
               int64_t* nnz, int32_t** h_rowptr, int32_t** h_colind);
int gbo_gen_er_m(int32_t n, int64_t m, uint64_t seed,
                 int64_t* nnz, int32_t** h_rowptr, int32_t** h_colind);
int gbo_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c,
                 uint64_t seed, int64_t* nnz, int32_t** h_rowptr, int32_t** h_colind);
int gbo_gen_king(int32_t m, int64_t* nnz, int32_t** h_rowptr, int32_t** h_colind);
/* Per-edge weight, uniform in [lo, hi], identical for (u,v) and (v,u). */
int gbo_gen_edge_weights(int32_t nrows, const int32_t* h_rowptr, const int32_t* h_colind,
                         uint32_t lo, uint32_t hi, uint64_t seed, uint32_t* h_w);
/* CSR transpose (csr2csc, algorithm/readMtx.hpp:320) and lower triangle
 * (tril_csr keeps row >= col, bsrbmm/bsrbmm.cu:1416); `strict` != 0 also
 * drops the diagonal. */
int gbo_csr_transpose(int32_t nrows, int32_t ncols, const int32_t* h_rowptr,
                      const int32_t* h_colind, int32_t** h_tptr, int32_t** h_tind);
int gbo_csr_tril(int32_t nrows, const int32_t* h_rowptr, const int32_t* h_colind,
                 int strict, int64_t* nnz, int32_t** h_lptr, int32_t** h_lind);

/* ------------------------------------------------------------- packing --
 * gbo_csr2bsr_nnzb replaces the cusparseXcsr2bsrNnz call that sizes the
 * packer's outputs (bsrbmv/benchmark-bsrbmv.cu:1039); gbo_csr2bsr_batch
 * replaces csr2bsr_batch_{8,16,32,64} (algorithm/csr2bsr_batch_bsrbmv.cu:
 * 180,406,621,786) and csr2bsr_batch_{N}_{row,col} (bsrbmm/csr2bsr_batch.cu).
 * Host CSR in (columns need not be sorted inside a row), device bit-BSR out.
 * The *_dev variants take a device CSR whose rows are sorted ascending. */
int gbo_csr2bsr_nnzb(int bs, const int32_t* h_rowptr, const int32_t* h_colind,
                     int32_t nrows, int32_t ncols, int64_t* nblocks);
int gbo_csr2bsr_batch(int bs, int dir, const int32_t* h_rowptr, const int32_t* h_colind,
                      int32_t nrows, int32_t ncols, int64_t nnz,
                      int32_t* d_bsrRowPtr, int32_t* d_bsrColInd, void* d_tiles,
                      int32_t nblockrows, int64_t nblocks, gbo_stream_t stream);
/* Device CSR: counts tiles per block row into d_bsrRowPtr (scanned) and
 * returns the total in *nblocks (synchronizes the stream). */
int gbo_csr2bsr_nnzb_dev(int bs, const int32_t* d_rowptr, const int32_t* d_colind,
                         int32_t nrows, int32_t ncols, int32_t* d_bsrRowPtr,
                         int64_t* nblocks, gbo_stream_t stream);
/* Device CSR: fills colind/tiles given d_bsrRowPtr from gbo_csr2bsr_nnzb_dev.
 * The rows are checked again (GBO_EINVAL for a column out of range or block
 * columns that go down inside a row), and a d_bsrRowPtr whose last entry is
 * not this CSR's tile count is refused before anything is written
 * (synchronizes the stream). */
int gbo_csr2bsr_fill_dev(int bs, int dir, const int32_t* d_rowptr, const int32_t* d_colind,
                         int32_t nrows, int32_t ncols, const int32_t* d_bsrRowPtr,
                         int32_t* d_bsrColInd, void* d_tiles, gbo_stream_t stream);
/* Convenience: allocate and pack in one call (device CSR in). Release with
 * gbo_bsr_free. */
int gbo_bsr_create_dev(int bs, int dir, const int32_t* d_rowptr, const int32_t* d_colind,
                       int32_t nrows, int32_t ncols, gbo_bsr* out, gbo_stream_t stream);
int gbo_bsr_create(int bs, int dir, const int32_t* h_rowptr, const int32_t* h_colind,
                   int32_t nrows, int32_t ncols, gbo_bsr* out, gbo_stream_t stream);
void gbo_bsr_free(gbo_bsr* A);
/* Work decomposition: split the tiles into chunks of `chunk_tiles` (0 = the
 * default, 256 tiles) and record each chunk's first block row.
 * Allocates A->plan (released by gbo_bsr_free / gbo_bsr_plan_free). */
int gbo_bsr_plan(gbo_bsr* A, int32_t chunk_tiles, gbo_stream_t stream);
/* As gbo_bsr_plan, also selecting the bs 8/16 kernel variant of the COUNT
 * and BFS products: 0 = one lane group per tile row word (default),
 * 1 = one lane per tile, block rows one after another, 2 = streamed: each
 * lane reads 4 (bs 8) or 2 (bs 16) consecutive tiles per step with the next
 * step's loads in flight across block-row boundaries (COUNT only; BFS runs
 * variant 0; needs 16-byte-aligned tiles/colind, else variant 0 runs),
 * 3 = bs 8 COUNT with 32-bit chunk-relative indexing and byte-transposed
 * popcounts (4 tiles per v_bcnt) and a permlane/DPP row reduction, bs 16
 * COUNT with four lanes per tile (8 bytes each); the BFS step of both is
 * dense (that count kernel gives each row's number of frontier neighbours, a
 * pack kernel keeps the nonzero unvisited rows; every tile is read, so
 * gbo_bfs takes it while fewer than half the vertices are visited and the
 * visited-row-skipping step after); at bs 8 and 16, min / min-plus / FP32 sum
 * list the tiles' set bits in LDS so the value gathers spread over all lanes
 * (the sum folds them into FP64 accumulators private to one wavefront).
 * bs 32 and 64 have one kernel set whatever the variant; their BFS step is
 * dense too (bs 32 COUNT with eight lanes per tile, bs 64 COUNT with scalar
 * per-tile reads).
 * Integer results are identical; only speed differs (the autotuner picks).
 * The FP32 sum's order of additions depends on the chunk size and variant,
 * each fixed by the plan, so its last bits may differ between plans.
 * gbo_bsr_create / gbo_bsr_create_dev set variant 3 for bs 8/16, 0 otherwise. */
int gbo_bsr_plan_ex(gbo_bsr* A, int32_t chunk_tiles, int32_t variant, gbo_stream_t stream);
void gbo_bsr_plan_free(gbo_bsr* A);
/* Per-nonzero values in tile bit order (row-major, MSB-first inside a row
 * word) plus per-tile offsets woff[nblocks+1].  Duplicate CSR entries keep
 * their minimum.  Device CSR in (rows sorted). */
int gbo_csr2bsr_values_u32(const gbo_bsr* A, const int32_t* d_rowptr, const int32_t* d_colind,
                           const uint32_t* d_val, uint32_t* d_tval, int64_t* d_woff,
                           gbo_stream_t stream);

/* Dense 0/1 float vector -> bit vector (ToBit{N}Row, algorithm/bsrbmv.cu:
 * 110,154,195,239): element > 0 sets the bit.  d_xbits holds ceil(n/bs)
 * words; the padding bits of the last word are zero. */
int gbo_pack_vec(int bs, const float* d_x, void* d_xbits, int32_t n, gbo_stream_t stream);

/* ---------------------------------------------------------- bsrbmv ops --
 * All take a row-direction (dir 0) matrix. */

/* y[i] = sum over tiles of popc(row word i & x word)  (int32 counts).
 * Reproduces bmv{8,16,32,64}_sparse(_new) (algorithm/bsrbmv.cu:2434,2776,
 * 2937,3046; bsrbmv/bsrbmv.cu:476).  accumulate != 0 adds into y like the
 * reference (`+=`); accumulate == 0 overwrites.  y has nrows entries. */
int gbo_bsrbmv_count(const gbo_bsr* A, const void* d_xbits, int32_t* d_y,
                     int accumulate, gbo_stream_t stream);

/* One pull-BFS level (bmv32_sparse_bin_masked_v4 / bmv4_sparse_bin_masked_v4,
 * algorithm/bsrbmv.cu:740,637): next word = (A*frontier != 0) & ~visited, one
 * bs-bit word per block row.  *d_count (device, int64) receives the number
 * of bits set in next.  Block rows whose visited word is all ones are
 * skipped, as are tiles whose frontier word is zero. */
int gbo_bsrbmv_bfs_step(const gbo_bsr* A, const void* d_frontier, const void* d_visited,
                        void* d_next, int64_t* d_count, gbo_stream_t stream);

/* Min-plus product (bmv32_sparse_full_minplus, algorithm/bsrbmv.cu:1083),
 * with optional weights:  out[i] = min(dist[i], min_{j: A_ij} dist[j] + w_ij)
 * saturating at INF = 0xFFFFFFFF.  d_w/d_woff come from
 * gbo_csr2bsr_values_u32; NULL means unit weights. */
int gbo_bsrbmv_minplus(const gbo_bsr* A, const uint32_t* d_w, const int64_t* d_woff,
                       const uint32_t* d_dist, uint32_t* d_out, gbo_stream_t stream);

/* y[i] = sum_{j: A_ij} x[j]  (FP32; bmv32_sparse_full, algorithm/bsrbmv.cu:
 * 1239, with the per-edge alpha/outdeg factor folded into x).  The summation
 * order is fixed by the plan, so repeated calls are bitwise identical. */
int gbo_bsrbmv_sum_f32(const gbo_bsr* A, const float* d_x, float* d_y, gbo_stream_t stream);

/* y[i] = min_{j: A_ij} x[j]  (bmv32_sparse_full_cc_new, algorithm/bsrbmv.cu:
 * 1978); rows without neighbours get INT32_MAX. */
int gbo_bsrbmv_min_i32(const gbo_bsr* A, const int32_t* d_x, int32_t* d_y, gbo_stream_t stream);

/* ---------------------------------------------------------- bsrbmm ops --
 * L_row: L packed with dir 0; LT_col: L^T (the CSC of L) packed with dir 1,
 * exactly the operands of bmm{N}_sparse_masked_v5 (bsrbmm/bsrbmm.cu:1214,
 * driver bsrbmm/tc.cu:822-853).  *d_count (device uint64) receives
 * sum_{i,k} L_ik (L L^T)_ik -- the triangle count when L is strictly lower. */
int gbo_bsrbmm_tc(const gbo_bsr* L_row, const gbo_bsr* LT_col, uint64_t* d_count,
                  gbo_stream_t stream);
/* *d_sum = sum_{i,k} (A B)_{ik} with B given as dir-1 tiles (bmm{N}_sparse,
 * bsrbmm/bsrbmm.cu:296,369,434,495; the bsrbmm<N> CLI calls it with A = L,
 * B = L^T and prints it as "ntris_bmm"). */
int gbo_bsrbmm_sum(const gbo_bsr* A_row, const gbo_bsr* B_col, uint64_t* d_sum,
                   gbo_stream_t stream);

/* ------------------------------------------- general bit-BSR SpGEMM ------
 * C = A * B with a bit-BSR output: A and B row-packed (dir 0), same block
 * size, A->ncols == B->nrows; C is row-packed, A->nrows x B->ncols.  A tile
 * of C exists exactly when the Boolean product has a set bit in it (the
 * tiles gbo_csr2bsr_batch would make of the product's pattern); C's tile
 * words are the Boolean product (bit set iff (A B)_ij > 0).  Expand-sort-
 * compress over the productive tile pairs (colmask(A tile) & rowmask(B
 * tile) != 0) on the device; each phase allocates and frees its own scratch.
 *
 * Symbolic phase (bmm32_sparse_symbolic, bsrbmm/bsrbmm.cu:1273): C's block
 * row pointer into the caller's d_C_rowptr[A->nblockrows + 1] and the tile
 * count into *C_nblocks (host). */
int gbo_bsrbmm_symbolic(const gbo_bsr* A, const gbo_bsr* B, int32_t* d_C_rowptr, int64_t* C_nblocks,
                        gbo_stream_t stream);
/* Numeric phase (bmm32_sparse_numeric, bsrbmm/bsrbmm.cu:1311): C's block
 * columns d_C_colind[C_nblocks] (ascending per block row) and tile words
 * d_C_tiles[C_nblocks * bs], caller-allocated from the symbolic phase's
 * count.  With d_C_valoff (int64[C_nblocks + 1], may be NULL) also the
 * exclusive scan of the tiles' bit counts, and *C_nnz (host, may be NULL)
 * their total: the offsets gbo_bsrbmm_values writes the counts at. */
int gbo_bsrbmm_numeric(const gbo_bsr* A, const gbo_bsr* B, const int32_t* d_C_rowptr, int64_t C_nblocks,
                       int32_t* d_C_colind, void* d_C_tiles, int64_t* d_C_valoff, int64_t* C_nnz,
                       gbo_stream_t stream);
/* The integer product (A B)_ij = sum_k A_ik B_kj at every set bit of C (the
 * values bmm32_sparse_numeric stores in C_csrval), int32, in tile bit order
 * (tiles in C's order, rows in order, columns ascending inside a row) from
 * d_C_valoff[t] (gbo_bsrbmm_numeric). */
int gbo_bsrbmm_values(const gbo_bsr* A, const gbo_bsr* B, const gbo_bsr* C, const int64_t* d_C_valoff,
                      int32_t* d_C_counts, gbo_stream_t stream);
/* Both phases into a library-owned C (release with gbo_bsr_free), with the
 * kernel variant gbo_bsr_create_dev would pick and no plan yet. */
int gbo_bsrbmm(const gbo_bsr* A, const gbo_bsr* B, gbo_bsr* C, gbo_stream_t stream);
/* Exclusive scan of the tiles' bit counts, d_off[A->nblocks + 1] (int64),
 * and their total *nnz (host, may be NULL): the offsets of per-nonzero
 * values in tile bit order. */
int gbo_bsr_bit_offsets(const gbo_bsr* A, int64_t* d_off, int64_t* nnz, gbo_stream_t stream);

/* ------------------------------------------------------------- drivers --
 * Whole algorithms with device-resident loops (host reads one scalar per
 * iteration).  `AT` is the matrix the pull step reads: for a symmetric
 * graph the adjacency matrix itself; for a directed graph its transpose
 * (the reference packs the CSC: algorithm/bfs.cu:400-423).  Synchronous. */

/* BFS (algorithm/bfs.cu:508-564): level[v] = hop distance from `source`,
 * -1 if unreachable.  h_level_sizes (optional, capacity `max_levels`)
 * receives the number of vertices discovered per level -- the reference's
 * per-level "succ" line.  *h_nlevels = number of levels incl. level 0. */
int gbo_bfs(const gbo_bsr* AT, int32_t source, int32_t* d_level,
            int64_t* h_level_sizes, int32_t max_levels, int32_t* h_nlevels,
            gbo_stream_t stream);
/* Allocation-free gbo_bfs.  No counterpart in the reference: for callers that
 * cannot allocate device memory during the call (serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as gbo_bfs (pull BFS
 * over AT); AT must already have a plan (gbo_bsr_plan), else GBO_EINVAL.
 * The frontier / next / visited words and the frontier-size counters live
 * in one static device array of 32 MiB held by each GPU the library runs on
 * until the process ends; a graph whose scratch exceeds it (about 89M
 * vertices at any block size) is refused with GBO_ERANGE before any launch.
 * At bs 8/16 (variant 3), 32 and 64 the dense step's row counts (4 bytes a vertex) go
 * there too for graphs up to about 7.6M vertices; larger ones take the
 * visited-row-skipping step at every level (same levels).
 * Calls that overlap on one GPU share that array and race: order them (one
 * stream, or synchronise between calls). */
int gbo_bfs_noalloc(const gbo_bsr* AT, int32_t source, int32_t* d_level,
                    int64_t* h_level_sizes, int32_t max_levels, int32_t* h_nlevels,
                    gbo_stream_t stream);
/* Direction-optimizing BFS, same outputs as gbo_bfs.  A is the out-edge
 * matrix (pass AT again for an undirected graph; NULL = pull only): a level
 * whose frontier holds at most n/64 vertices is pushed along the frontier
 * rows of A (only their tiles are read), the others are pulled over AT.
 * No reference counterpart (algorithm/bfs.cu always pulls). */
int gbo_bfs_do(const gbo_bsr* AT, const gbo_bsr* A, int32_t source, int32_t* d_level,
               int64_t* h_level_sizes, int32_t max_levels, int32_t* h_nlevels,
               gbo_stream_t stream);

/* SSSP (algorithm/sssp.cu:420-488) generalised to uint32 weights:
 * frontier Bellman-Ford on the min-plus semiring; dist = INF (0xFFFFFFFF)
 * when unreachable.  d_w/d_woff as for gbo_bsrbmv_minplus (NULL = unit). */
int gbo_sssp(const gbo_bsr* AT, const uint32_t* d_w, const int64_t* d_woff,
             int32_t source, uint32_t* d_dist, int32_t* h_iters, gbo_stream_t stream);

/* PageRank (algorithm/pr.cu:449-516): p0 = 1/n,
 * p' = alpha * sum_{j in in(i)} p[j]/outdeg(j) + (1-alpha)/n, no dangling
 * redistribution; stops when ||p'-p||_1 < tol or after max_iter iterations.
 * d_outdeg[n] = out-degree of each vertex (row lengths of A).  Returns
 * GBO_ENOCONV if max_iter is reached first (d_p still holds the iterate). */
int gbo_pagerank(const gbo_bsr* AT, const int32_t* d_outdeg, float alpha, double tol,
                 int32_t max_iter, float* d_p, int32_t* h_iters, double* h_last_delta,
                 gbo_stream_t stream);

/* Connected components (FastSV, algorithm/cc.cu:1018-1083): label[v] =
 * smallest vertex id of v's component.  A must be symmetric. */
int gbo_cc(const gbo_bsr* A, int32_t* d_label, int32_t* h_iters, gbo_stream_t stream);

/* Triangle count (bsrbmm/tc.cu:690-1013) of a symmetric graph given as a
 * host CSR: L = strict tril(A), pack L (dir 0) and L^T (dir 1) at block
 * size bs, run gbo_bsrbmm_tc. */
int gbo_tc(int bs, int32_t nrows, const int32_t* h_rowptr, const int32_t* h_colind,
           uint64_t* h_count, gbo_stream_t stream);

/* ------------------------------------------------------------ batching --
 * Multi-source BFS: up to 256 BFS runs from h_sources[0..nsrc) in one pass
 * per level, as a bit-BSR product with nsrc-bit frontier words per vertex
 * (no reference counterpart; it batches the BFS of algorithm/bfs.cu:508-564).
 * d_level is [nsrc][n] int32: d_level[s*n + v] = hop distance of v from
 * source s, -1 if unreachable -- identical to nsrc separate gbo_bfs runs.
 * *h_nlevels = number of levels including level 0. */
int gbo_bfs_multisource(const gbo_bsr* AT, const int32_t* h_sources, int32_t nsrc, int32_t* d_level,
                        int32_t* h_nlevels, gbo_stream_t stream);

/* Autotuner: packs the device CSR at every block size in bs_mask (bit k =
 * bs 8<<k) and times `reps` sweeps of op 0 (gbo_bsrbmv_count) or op 1
 * (gbo_bsrbmv_bfs_step) for each kernel variant and three chunk sizes;
 * returns the fastest (block size, chunk, variant, ms).  Cached per graph
 * signature for the life of the process.  No reference counterpart (the
 * reference fixes the block size at build time, bsrbmv/Makefile:57-67). */
typedef struct gbo_tuning {
  int32_t bs, chunk, variant;
  float ms;
} gbo_tuning;
int gbo_autotune(const int32_t* d_rowptr, const int32_t* d_colind, int32_t nrows, int32_t ncols, int op,
                 int bs_mask, int reps, gbo_tuning* out, gbo_stream_t stream);

/* Many-graph batch: G independent square graphs packed side by side into one
 * block-diagonal bit-BSR (each graph's first vertex aligned to bs).  One
 * persistent launch -- captured in a hipGraph when use_graph != 0 and
 * replayed by later calls with the same output buffers -- runs BFS from
 * h_sources[g] and FastSV connected components on every graph with no host
 * synchronisation between iterations: all graphs' block rows are swept
 * together by the whole GPU, level by level, separated by grid barriers
 * among the workgroups that are resident; each graph stops where a
 * one-graph run would.  The rows given are read as in gbo_bfs reads AT
 * (pull): for undirected graphs pass the symmetric CSR; connected
 * components assume symmetric graphs.  Outputs are concatenated per graph
 * (offsets from gbo_batch_offsets): d_level as gbo_bfs, d_label = smallest
 * graph-local vertex id of the component, identical to per-graph gbo_bfs /
 * gbo_cc runs.  A grid barrier wait beyond 5 s (GBO_BATCH_SPIN_MS in the
 * environment overrides the limit) ends the launch and returns GBO_EDEVICE.
 * No reference
 * counterpart (the reference runs one graph per process). */
typedef struct gbo_batch gbo_batch;
int gbo_batch_create(int bs, int32_t ngraphs, const int32_t* h_nrows, const int32_t* const* h_rowptr,
                     const int32_t* const* h_colind, gbo_batch** out, gbo_stream_t stream);
int gbo_batch_offsets(const gbo_batch* b, int64_t* h_offsets /* ngraphs+1 */);
int gbo_batch_info(const gbo_batch* b, int32_t* bs, int64_t* nblocks, int64_t* nvertices_padded);
int gbo_batch_bfs_cc(gbo_batch* b, const int32_t* h_sources, int32_t* d_level, int32_t* d_label, int use_graph,
                     gbo_stream_t stream);
void gbo_batch_free(gbo_batch* b);

/* -------------------------------------------------- differentiation (PPR) --
 * Personalized PageRank over K <= 64 vectors stored vertex-major
 * ([n][K] float):  p_k = alpha * P^T p_k + (1 - alpha) * s_k,  P = D^-1 A,
 * rows of vertices without out-edges contribute nothing (as pr.cu).  The
 * pull product reads AT (A^T; A itself for a symmetric graph).  Iterates
 * from p = s until max_k ||p_k' - p_k||_1 / ||s_k||_1 < tol (GBO_ENOCONV
 * after max_iter; the relative change is reported in *h_delta).  No
 * reference counterpart: generalises algorithm/pr.cu:449-516. */
int gbo_ppr_forward(const gbo_bsr* AT, const int32_t* d_outdeg, int32_t K, float alpha, const float* d_s,
                    float* d_p, double tol, int32_t max_iter, int32_t* h_iters, double* h_delta,
                    gbo_stream_t stream);
/* Adjoint of gbo_ppr_forward for L = sum_k w_k . p_k: solves
 * lambda_k = alpha * P lambda_k + w_k (pull over A), then
 * d_grad_s = (1 - alpha) * lambda and *h_grad_alpha =
 * sum_k lambda_k . (P^T p_k - s_k).  Uses only the converged p.  Iterates
 * from lambda = w until max_k ||lambda_k' - lambda_k||_1 / ||w_k||_1 < tol
 * (the same contraction as the forward solve, so the same tol needs about
 * the same number of iterations). */
int gbo_ppr_adjoint(const gbo_bsr* A, const gbo_bsr* AT, const int32_t* d_outdeg, int32_t K, float alpha,
                    const float* d_s, const float* d_p, const float* d_w, float* d_grad_s, double* h_grad_alpha,
                    double tol, int32_t max_iter, int32_t* h_iters, gbo_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* GBO_H */
