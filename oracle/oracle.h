/*
 * oracle.h -- CPU restatement of the reference's bit-BSR graph path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or, for the baseline, as the thing timed on the host cores).
 * The product (graph-bit-ops_amd/) never calls it.
 *
 * Every function names the reference file:line whose semantics it restates
 * (paths relative to the reference tree).  What pins each function besides
 * the GPU (dense brute force, closed forms, textbook algorithms, library
 * results, finite differences, distribution checks) is in
 * tests/test_oracle.py.
 *
 * Layout: tiles are stored as T*bs words of bs bits (uint8/16/32/64 for bs
 * 8/16/32/64), word r of tile t at index t*bs + r, bit (bs-1-c) = column c
 * (dir 0) or word k = column k with bit (bs-1-r) = row r (dir 1).
 */
#ifndef GBO_ORACLE_H
#define GBO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generators (independent restatement of the engine's generators) ---- */
uint64_t or_rand64(uint64_t seed, uint64_t stream, uint64_t idx);
/* Edge list of G(n,p) (pairs u<v), R-MAT and the king grid; returns the
 * number of undirected edges written to (u[], v[]) (capacity cap; call
 * with cap 0 to count). */
int64_t or_gen_er_edges(int32_t n, double p, uint64_t seed, int32_t* u, int32_t* v, int64_t cap);
int64_t or_gen_rmat_edges(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed,
                          int32_t* u, int32_t* v);

/* ---- packing: csr2bsr_batch_{N} (algorithm/csr2bsr_batch_bsrbmv.cu:621-780)
 * with ToBit{N}Col (algorithm/bsrbmv.cu:177-236) and the _col variant
 * (bsrbmm/csr2bsr_batch.cu, CUSPARSE_DIRECTION_COLUMN). ---- */
int64_t or_pack_count(int bs, int32_t nrows, const int32_t* rp, const int32_t* ci, int32_t* brp);
void or_pack_fill(int bs, int dir, int32_t nrows, const int32_t* rp, const int32_t* ci,
                  const int32_t* brp, int32_t* bci, void* tiles);
/* ToBit{N}Row (algorithm/bsrbmv.cu:195-207, 239-253). */
void or_pack_vec(int bs, int32_t n, const float* x, void* xbits);
/* Per-nonzero values in tile bit order + per-tile offsets (no reference
 * counterpart: the reference is unit-weight, algorithm/sssp.cu:291). */
void or_pack_values(int bs, int32_t nrows, const int32_t* rp, const int32_t* ci, const uint32_t* val,
                    const int32_t* brp, const int32_t* bci, const void* tiles, uint32_t* tval, int64_t* woff);

/* ---- bsrbmv semantics ---- */
/* bmv{N}_sparse_new: y[i] = sum_t popc(row word & x[colind]) (algorithm/
 * bsrbmv.cu:2937-2975, 3046-3093).  OpenMP over block rows [r0, r1). */
void or_bsrbmv_count(int bs, int32_t nrows, int32_t r0, int32_t r1, const int32_t* brp, const int32_t* bci,
                     const void* tiles, const void* xbits, int32_t* y);
/* bmv32_sparse_bin_masked_v4 (algorithm/bsrbmv.cu:740-783): next word =
 * brev(ballot(count > 0)) & ~visited; returns popcount of next. */
int64_t or_bsrbmv_bfs_step(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                           const void* frontier, const void* visited, void* next);
/* bmv32_sparse_full_minplus (algorithm/bsrbmv.cu:1083-1137) with weights:
 * out[i] = min(dist[i], min_j dist[j] + w_ij), saturating at 0xFFFFFFFF. */
void or_bsrbmv_minplus(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                       const uint32_t* w, const int64_t* woff, const uint32_t* dist, uint32_t* out);
/* bmv32_sparse_full (algorithm/bsrbmv.cu:1239-1283): y[i] = sum_j x[j] in
 * double (the reference multiplies alpha/deg per edge; folded into x). */
void or_bsrbmv_sum_f64(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                       const double* x, double* y);
/* bmv32_sparse_full_cc_new (algorithm/bsrbmv.cu:1978-2019): y[i] = min_j
 * x[j], INT32_MAX without neighbours. */
void or_bsrbmv_min_i32(int bs, int32_t nrows, const int32_t* brp, const int32_t* bci, const void* tiles,
                       const int32_t* x, int32_t* y);

/* ---- bsrbmm semantics (bsrbmm/bsrbmm.cu) ---- */
/* bmm{N}_sparse_masked_v5 (bsrbmm/bsrbmm.cu:1214-1265): sum_{i,k} A_ik (A B)_ik
 * with A row tiles and B column tiles; same traversal (per mask tile, per
 * tile of the block row, binary search in B's block row). */
uint64_t or_bsrbmm_masked(int bs, int32_t nbr, const int32_t* arp, const int32_t* aci, const void* atiles,
                          const int32_t* brp, const int32_t* bci, const void* btiles);
/* bmm{N}_sparse (bsrbmm/bsrbmm.cu:434-488): sum_{i,k} (A B)_ik, tile-pair loop. */
uint64_t or_bsrbmm_sum(int bs, int32_t nbr, const int32_t* arp, const int32_t* aci, const void* atiles,
                       const int32_t* brp, const int32_t* bci, const void* btiles);
/* C = A * B over the integers, element CSR with counts (bsrbmm/bsrbmm.cu:1273,
 * 1311); ci == NULL: fill rp[m+1], return nnz; else fill ci/val. */
int64_t or_spgemm(int32_t m, const int32_t* arp, const int32_t* aci, const int32_t* brp, const int32_t* bci,
                  int32_t ncols, int32_t* rp, int32_t* ci, int32_t* val);

/* ---- drivers restated on bit tiles ---- */
/* BFS loop of algorithm/bfs.cu:508-564 (visited |= f; f = A f & ~visited)
 * recording levels; returns number of levels. */
int32_t or_bfs_bsr(int bs, int32_t n, const int32_t* brp, const int32_t* bci, const void* tiles, int32_t src,
                   int32_t* level, int64_t* sizes, int32_t max_levels);

/* ---- textbook algorithms on CSR ---- */
int32_t or_bfs_queue(int32_t n, const int32_t* rp, const int32_t* ci, int32_t src, int32_t* level);
void or_dijkstra(int32_t n, const int32_t* rp, const int32_t* ci, const uint32_t* w, int32_t src, uint32_t* dist);
/* PageRank of algorithm/pr.cu:449-516 in FP64 on the pull CSR (row i lists
 * in-neighbours): p' = alpha sum p[j]/outdeg[j] + (1-alpha)/n; iterates
 * until the L1 change is below tol or max_iter; returns iterations. */
int32_t or_pagerank(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* outdeg, double alpha,
                    double tol, int32_t max_iter, double* p);
/* Personalized PageRank in FP64, K columns in vertex-major [n][K] layout.
 * mode 0, the forward map of algorithm/pr.cu:449-516 with a personalization
 * vector R (pull CSR: row i lists the in-neighbours of i):
 *   T(X)[i][k] = alpha * sum_{j in row i} X[j][k] / outdeg[j] + (1 - alpha) R[i][k]
 * mode 1, its adjoint (out-edge CSR: row i lists the out-neighbours of i):
 *   T(X)[i][k] = alpha / outdeg[i] * sum_{j in row i} X[j][k] + R[i][k]
 * (the sum term is 0 where outdeg is 0).  Both maps are alpha-contractions
 * (mode 0 in the 1-norm, mode 1 in the max-norm), so for any X the distance
 * to the fixed point is at most |X - T(X)| / (1 - alpha) in that norm.
 * or_ppr_solve iterates X <- T(X) from the X given until the largest column
 * change relative to that column's |R|_1 is below tol; returns the number
 * of iterations, or -1 if max_iter is reached first. */
int32_t or_ppr_solve(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* outdeg, double alpha, int32_t K,
                     int mode, const double* R, double* X, double tol, int32_t max_iter);
/* One application of T to FP32 X, in FP64: per column the 1-norm and the
 * max-norm of X - T(X).  With L (mode 0 only) also, per column,
 * dot[k] = sum_i L[i][k] (y[i][k] - R[i][k]) and y_l1[k] = |y_k - R_k|_1,
 * where y = sum_{j in row i} X[j][k] / outdeg[j] (the teleport-factor
 * gradient's product).  Arrays that are not wanted may be NULL. */
void or_ppr_check(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* outdeg, double alpha, int32_t K,
                  int mode, const float* X, const float* R, const float* L, double* res_l1, double* res_inf,
                  double* dot, double* y_l1);
/* Union-find connected components, label = smallest vertex id. */
void or_cc(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* label);
/* Triangles of a symmetric simple graph by sorted-list intersection on
 * L = strict lower triangle. */
uint64_t or_tc(int32_t n, const int32_t* rp, const int32_t* ci);

#ifdef __cplusplus
}
#endif
#endif
