// .mtx-in command lines of the engine, drop-ins for the reference's CLIs:
//   bsrbmv<N> file.mtx            bsrbmv/benchmark-bsrbmv.cu (main32 :971)
//   bsrbmm<N> file.mtx            bsrbmm/benchmark-bsrbmm.cu (main32 :1004)
//   tc<N>     file.mtx directed   bsrbmm/tc.cu (main32 :690)
//   bfs<N>    file.mtx directed   algorithm/bfs.cu (main32 :310)
//   sssp<N>   file.mtx            algorithm/sssp.cu (main32 :252)
//   pr<N>     file.mtx            algorithm/pr.cu (main32 :269)
//   cc<N>     file.mtx directed   algorithm/cc.cu (main32 :857)
// The binary dispatches on its own name (bsrbmv32 -> bsrbmv, bs 32) or on
// `gbo <cmd> [options] file.mtx [directed]`.  The block size is a run-time
// option (--bs) instead of a -DBLOCKSIZE build; stdout keeps the reference's
// key lines ("nrows: ...", "blocksize: ...", "BSR BMV-32 success: 1",
// "succ: N", "ntris_bmm: N", ...).  Extra options:
//   --bs N        block size (8, 16, 32, 64)
//   --seed S      seed of the random x of bsrbmv (the reference used time(0))
//   --source V    BFS/SSSP source (default 0, as the reference)
//   --max-iter K  PageRank iteration cap (default 10, as algorithm/pr.cu:4)
//   --tol T       PageRank L1 stopping tolerance (default 1e-9)
//   --out FILE    write the per-vertex result (levels, distances, scores,
//                 labels) one value per line
//   --json        print one machine-readable JSON result line at the end
//   --cache       read through the binary CSR cache beside the .mtx file
//                 (readMtx's ".<file>.<ud|d>.<nosl|sl>.bin"); also on when
//                 the environment variable GBO_MTX_CACHE is 1
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "gbo.h"

namespace {

struct Opts {
  std::string cmd, file;
  int bs = 32;
  int directed = 0;
  uint64_t seed = 0;
  int source = 0;
  int max_iter = 10;
  double tol = 1e-9;
  std::string out;
  bool json = false;
  bool cache = false;
};

[[noreturn]] void die(const char* what, int rc) {
  std::fprintf(stderr, "error: %s failed (%d): %s\n", what, rc, gbo_last_error());
  std::exit(rc < 0 ? -rc : 1);
}
#define CK(call)                 \
  do {                           \
    int _rc = (call);            \
    if (_rc != GBO_OK) die(#call, _rc); \
  } while (0)

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void usage() {
  std::fprintf(stderr,
               "usage: gbo <bsrbmv|bsrbmm|tc|bfs|sssp|pr|cc> [--bs N] [--seed S] [--source V] [--max-iter K]\n"
               "           [--tol T] [--out FILE] [--json] [--cache] file.mtx [directed]\n"
               "   or: bsrbmv32 file.mtx, bfs8 file.mtx 0, ... (name carries command and block size)\n");
  std::exit(2);
}

struct HostCsr {
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  int32_t* rp = nullptr;
  int32_t* ci = nullptr;
  ~HostCsr() {
    gbo_free_host(rp);
    gbo_free_host(ci);
  }
};

void read_mtx(const Opts& o, int semantics, HostCsr* c) {
  std::printf("input sparse matrix: %s\n", o.file.c_str());
  const char* env = std::getenv("GBO_MTX_CACHE");
  const double t0 = now_ms();
  if (o.cache || (env && std::atoi(env) == 1)) {
    int state = -1;
    CK(gbo_mtx_read_cached(o.file.c_str(), semantics, o.directed, &state, &c->nrows, &c->ncols, &c->nnz, &c->rp,
                           &c->ci, nullptr));
    char name[4096];
    CK(gbo_mtx_cache_path(o.file.c_str(), semantics, o.directed, name, sizeof(name)));
    std::printf("csr cache: %s (%s)\n", name, state == 1 ? "loaded" : state == 0 ? "written" : "not writable");
  } else {
    CK(gbo_mtx_read(o.file.c_str(), semantics, o.directed, &c->nrows, &c->ncols, &c->nnz, &c->rp, &c->ci, nullptr));
  }
  std::printf("read time: %.1f ms\n", now_ms() - t0);
  std::printf("nrows: %d, ncols: %d, nnz: %lld\n", c->nrows, c->ncols, (long long)c->nnz);
}

void print_bytes(int64_t b) {
  if (b >= (1LL << 30)) std::printf("%.3f GB", b / double(1LL << 30));
  else if (b >= (1LL << 20)) std::printf("%.3f MB", b / double(1LL << 20));
  else if (b >= (1LL << 10)) std::printf("%.3f KB", b / double(1LL << 10));
  else std::printf("%lld B", (long long)b);
}

void pack(const Opts& o, int dir, int32_t nrows, int32_t ncols, const int32_t* rp, const int32_t* ci, gbo_bsr* A) {
  CK(gbo_bsr_create(o.bs, dir, rp, ci, nrows, ncols, A, nullptr));
  CK(gbo_bsr_plan(A, 0, nullptr));
  std::printf("blocksize: %d, nblockrows: %d, nblocks: %lld\n", o.bs, A->nblockrows, (long long)A->nblocks);
  std::printf("bsr total size: ");
  print_bytes(A->nblocks * (int64_t)o.bs * o.bs / 8 + 4 * (A->nblockrows + 1 + A->nblocks));
  std::printf("\n");
}

template <class T>
void write_out(const Opts& o, const std::vector<T>& v, const char* fmt) {
  if (o.out.empty()) return;
  FILE* f = std::fopen(o.out.c_str(), "w");
  if (!f) die("open --out", GBO_EIO);
  for (const T& x : v) std::fprintf(f, fmt, x), std::fputc('\n', f);
  std::fclose(f);
}

template <class T>
T* dalloc(size_t n) {
  void* p = nullptr;
  CK(gbo_malloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return (T*)p;
}

// bsrbmv<N>: read -> pack -> random x -> 5 accumulating launches -> check
// against a host CSR matvec (the reference compares with cusparseScsrmv).
int run_bsrbmv(const Opts& o) {
  HostCsr c;
  read_mtx(o, 0, &c);
  gbo_bsr A{};
  pack(o, 0, c.nrows, c.ncols, c.rp, c.ci, &A);
  const int TEST_TIMES = 5;  // bsrbmv/benchmark-bsrbmv.cu:4
  const int32_t nbc = A.nblockcols;
  std::mt19937_64 rng(o.seed);
  std::vector<float> x((size_t)nbc * o.bs, 0.0f);
  for (int32_t i = 0; i < c.ncols; ++i) x[i] = (rng() >> 11) * (1.0 / 9007199254740992.0) > 0.5 ? 1.0f : 0.0f;
  float* dxf = dalloc<float>(x.size());
  void* dxb = dalloc<uint8_t>((size_t)nbc * o.bs / 8);
  int32_t* dy = dalloc<int32_t>(c.nrows);
  CK(gbo_memcpy_h2d(dxf, x.data(), x.size() * sizeof(float)));
  CK(gbo_pack_vec(o.bs, dxf, dxb, (int32_t)x.size(), nullptr));
  CK(gbo_memset(dy, 0, sizeof(int32_t) * std::max(c.nrows, 1)));
  void *e0, *e1;
  CK(gbo_event_create(&e0));
  CK(gbo_event_create(&e1));
  CK(gbo_event_record(e0, nullptr));
  for (int i = 0; i < TEST_TIMES; ++i) CK(gbo_bsrbmv_count(&A, dxb, dy, 1, nullptr));
  CK(gbo_event_record(e1, nullptr));
  float ms = 0;
  CK(gbo_event_elapsed(&ms, e0, e1));
  std::vector<int32_t> y(c.nrows);
  CK(gbo_memcpy_d2h(y.data(), dy, sizeof(int32_t) * c.nrows));
  // host check: TEST_TIMES * (A x), duplicates in the CSR OR together
  bool ok = true;
  std::vector<int32_t> mark(std::max(c.ncols, 1), -1);
  for (int32_t r = 0; r < c.nrows && ok; ++r) {
    int64_t s = 0;
    for (int32_t k = c.rp[r]; k < c.rp[r + 1]; ++k)
      if (mark[c.ci[k]] != r) {
        mark[c.ci[k]] = r;
        s += x[c.ci[k]] > 0;
      }
    ok = (y[r] == TEST_TIMES * s);
  }
  std::printf("BSR BMV-%d success: %d\n", o.bs, ok ? 1 : 0);
  std::printf("BSR BMV-%d: %.3lf\n", o.bs, ms / TEST_TIMES);
  write_out(o, y, "%d");
  if (o.json)
    std::printf("{\"cmd\":\"bsrbmv\",\"bs\":%d,\"nnz\":%lld,\"nblocks\":%lld,\"ms\":%.4f,\"success\":%d}\n", o.bs,
                (long long)c.nnz, (long long)A.nblocks, ms / TEST_TIMES, ok ? 1 : 0);
  gbo_free(dxf);
  gbo_free(dxb);
  gbo_free(dy);
  gbo_bsr_free(&A);
  return ok ? 0 : 1;
}

// bsrbmm<N> (the product C = L L^T by the bit-BSR SpGEMM -- symbolic,
// numeric and the integer values at C's set bits, the timed work -- and the
// sum of C's values printed as ntris_bmm like the reference) and tc<N>
// (masked: triangle count).  L = tril(A) keeps row >= col (tril_csr,
// bsrbmm/bsrbmm.cu:1416).
int run_bmm(const Opts& o, bool masked) {
  HostCsr c;
  read_mtx(o, masked ? 1 : 0, &c);
  if (c.nrows != c.ncols) die("square matrix check", GBO_EINVAL);
  int64_t lnnz = 0;
  int32_t *lp, *li, *tp, *ti;
  CK(gbo_csr_tril(c.nrows, c.rp, c.ci, 0, &lnnz, &lp, &li));
  std::printf("nnz before tril_csr(): %lld, after tril_csr(): %lld\n", (long long)c.nnz, (long long)lnnz);
  CK(gbo_csr_transpose(c.nrows, c.ncols, lp, li, &tp, &ti));
  gbo_bsr L{}, LT{}, C{};
  pack(o, 0, c.nrows, c.ncols, lp, li, &L);
  // tc: L^T column-packed (the masked kernel's operand); bsrbmm: L^T
  // row-packed, the SpGEMM's right operand
  CK(gbo_bsr_create(o.bs, masked ? 1 : 0, tp, ti, c.ncols, c.nrows, &LT, nullptr));
  uint64_t* d = dalloc<uint64_t>(1);
  int64_t* d_off = nullptr;
  int32_t* d_cnt = nullptr;
  int64_t cnnz = 0;
  void *e0, *e1;
  CK(gbo_event_create(&e0));
  CK(gbo_event_create(&e1));
  CK(gbo_synchronize());
  CK(gbo_event_record(e0, nullptr));
  if (masked) {
    CK(gbo_bsrbmm_tc(&L, &LT, d, nullptr));
  } else {
    CK(gbo_bsrbmm(&L, &LT, &C, nullptr));
    d_off = dalloc<int64_t>((size_t)C.nblocks + 1);
    CK(gbo_bsr_bit_offsets(&C, d_off, &cnnz, nullptr));
    d_cnt = dalloc<int32_t>((size_t)cnnz);
    CK(gbo_bsrbmm_values(&L, &LT, &C, d_off, d_cnt, nullptr));
  }
  CK(gbo_event_record(e1, nullptr));
  float ms = 0;
  CK(gbo_event_elapsed(&ms, e0, e1));
  uint64_t r = 0;
  if (masked) {
    CK(gbo_memcpy_d2h(&r, d, 8));
  } else {
    std::vector<int32_t> v((size_t)cnnz);
    if (cnnz) CK(gbo_memcpy_d2h(v.data(), d_cnt, sizeof(int32_t) * (size_t)cnnz));
    for (int32_t x : v) r += (uint64_t)x;
    std::printf("C = L*L^T: nblocks %lld, nnz %lld\n", (long long)C.nblocks, (long long)cnnz);
    gbo_free(d_off);
    gbo_free(d_cnt);
    gbo_bsr_free(&C);
  }
  std::printf("ntris_bmm: %llu\n", (unsigned long long)r);
  if (!masked) {  // host check: sum_c colcount_c(L)^2 (the cusparseScsrgemm sum)
    std::vector<int64_t> cc(c.ncols, 0);
    for (int64_t k = 0; k < lnnz; ++k) cc[li[k]]++;
    uint64_t s = 0;
    for (int64_t v : cc) s += (uint64_t)(v * v);
    std::printf("ntris_csr: %llu\n", (unsigned long long)s);
  }
  std::printf("BSR BMM-%d: %.3lf\n", o.bs, ms);
  if (o.json)
    std::printf("{\"cmd\":\"%s\",\"bs\":%d,\"result\":%llu,\"ms\":%.4f}\n", masked ? "tc" : "bsrbmm", o.bs,
                (unsigned long long)r, ms);
  gbo_free(d);
  gbo_free_host(lp);
  gbo_free_host(li);
  gbo_free_host(tp);
  gbo_free_host(ti);
  gbo_bsr_free(&L);
  gbo_bsr_free(&LT);
  return 0;
}

// Pull algorithms read A^T (the reference packs the CSC: algorithm/bfs.cu:400).
void pack_transpose(const Opts& o, const HostCsr& c, gbo_bsr* AT) {
  int32_t *tp, *ti;
  CK(gbo_csr_transpose(c.nrows, c.ncols, c.rp, c.ci, &tp, &ti));
  pack(o, 0, c.ncols, c.nrows, tp, ti, AT);
  gbo_free_host(tp);
  gbo_free_host(ti);
}

int run_bfs(const Opts& o) {
  HostCsr c;
  read_mtx(o, 1, &c);
  gbo_bsr AT{};
  pack_transpose(o, c, &AT);
  int32_t* dl = dalloc<int32_t>(c.nrows);
  std::vector<int64_t> sizes(1 << 16);
  int32_t nl = 0;
  std::printf("nrows: %d\n------------------------------------\n", c.nrows);
  const double t0 = now_ms();
  CK(gbo_bfs(&AT, o.source, dl, sizes.data(), (int32_t)sizes.size(), &nl, nullptr));
  const double ms = now_ms() - t0;
  for (int32_t l = 1; l < nl && l < (int32_t)sizes.size(); ++l) std::printf("succ: %lld\n", (long long)sizes[l]);
  std::printf("succ: 0\n------------------------------------\n");
  std::printf("niter: %d\n", nl);
  std::printf("BSR BMV-%d-bin: %.3lf\n", o.bs, ms);
  std::vector<int32_t> lv(c.nrows);
  CK(gbo_memcpy_d2h(lv.data(), dl, sizeof(int32_t) * c.nrows));
  write_out(o, lv, "%d");
  if (o.json) std::printf("{\"cmd\":\"bfs\",\"bs\":%d,\"levels\":%d,\"ms\":%.3f}\n", o.bs, nl, ms);
  gbo_free(dl);
  gbo_bsr_free(&AT);
  return 0;
}

int run_sssp(const Opts& o) {
  HostCsr c;
  read_mtx(o, 0, &c);
  gbo_bsr AT{};
  pack_transpose(o, c, &AT);
  uint32_t* dd = dalloc<uint32_t>(c.nrows);
  int32_t it = 0;
  std::printf("nrows: %d\n------------------------------------\n", c.nrows);
  const double t0 = now_ms();
  CK(gbo_sssp(&AT, nullptr, nullptr, o.source, dd, &it, nullptr));
  const double ms = now_ms() - t0;
  std::printf("------------------------------------\nniter: %d\n", it);
  std::printf("BSR BMV-%d-bin: %.3lf\n", o.bs, ms);
  std::vector<uint32_t> d(c.nrows);
  CK(gbo_memcpy_d2h(d.data(), dd, sizeof(uint32_t) * c.nrows));
  write_out(o, d, "%u");
  if (o.json) std::printf("{\"cmd\":\"sssp\",\"bs\":%d,\"iters\":%d,\"ms\":%.3f}\n", o.bs, it, ms);
  gbo_free(dd);
  gbo_bsr_free(&AT);
  return 0;
}

int run_pr(const Opts& o) {
  HostCsr c;
  read_mtx(o, 0, &c);
  if (c.nrows != c.ncols) die("square matrix check", GBO_EINVAL);
  gbo_bsr AT{};
  pack_transpose(o, c, &AT);
  std::vector<int32_t> deg(c.nrows);
  for (int32_t r = 0; r < c.nrows; ++r) deg[r] = c.rp[r + 1] - c.rp[r];
  int32_t* dd = dalloc<int32_t>(c.nrows);
  float* dp = dalloc<float>(c.nrows);
  CK(gbo_memcpy_h2d(dd, deg.data(), sizeof(int32_t) * c.nrows));
  int32_t it = 0;
  double delta = 0;
  std::printf("nrows: %d\n------------------------------------\n", c.nrows);
  const double t0 = now_ms();
  int rc = gbo_pagerank(&AT, dd, 0.85f, o.tol, o.max_iter, dp, &it, &delta, nullptr);
  const double ms = now_ms() - t0;
  if (rc != GBO_OK && rc != GBO_ENOCONV) die("gbo_pagerank", rc);
  std::printf("error: %e\n------------------------------------\nniter: %d\n", delta, it);
  std::printf("BSR BMV-%d-bin: %.3lf\n", o.bs, ms);
  std::vector<float> p(c.nrows);
  CK(gbo_memcpy_d2h(p.data(), dp, sizeof(float) * c.nrows));
  write_out(o, p, "%.9g");
  if (o.json) std::printf("{\"cmd\":\"pr\",\"bs\":%d,\"iters\":%d,\"l1_delta\":%.3e,\"ms\":%.3f}\n", o.bs, it, delta, ms);
  gbo_free(dd);
  gbo_free(dp);
  gbo_bsr_free(&AT);
  return 0;
}

int run_cc(const Opts& o) {
  HostCsr c;
  read_mtx(o, 1, &c);
  gbo_bsr A{};
  pack(o, 0, c.nrows, c.ncols, c.rp, c.ci, &A);
  int32_t* dl = dalloc<int32_t>(c.nrows);
  int32_t it = 0;
  const double t0 = now_ms();
  CK(gbo_cc(&A, dl, &it, nullptr));
  const double ms = now_ms() - t0;
  std::vector<int32_t> lab(c.nrows);
  CK(gbo_memcpy_d2h(lab.data(), dl, sizeof(int32_t) * c.nrows));
  // SimpleVerifyCc (algorithm/cc.cu:22-60): endpoints of every edge share a label.
  int64_t errors = 0, comps = 0;
  for (int32_t r = 0; r < c.nrows; ++r) {
    comps += lab[r] == r;
    for (int32_t k = c.rp[r]; k < c.rp[r + 1]; ++k) errors += lab[c.ci[k]] != lab[r];
  }
  std::printf("------------------------------------\nniter: %d\n", it);
  if (errors == 0) std::printf("\nCORRECT\n");
  else std::printf("%lld errors occurred.\n", (long long)errors);
  std::printf("Connected components found with %lld components.\n", (long long)comps);
  std::printf("BSR BMV-%d-bin: %.3lf\n", o.bs, ms);
  write_out(o, lab, "%d");
  if (o.json)
    std::printf("{\"cmd\":\"cc\",\"bs\":%d,\"iters\":%d,\"components\":%lld,\"ms\":%.3f}\n", o.bs, it, (long long)comps, ms);
  gbo_free(dl);
  gbo_bsr_free(&A);
  return errors ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  Opts o;
  std::string self = argv[0];
  self = self.substr(self.find_last_of('/') + 1);
  int argi = 1;
  size_t digits = self.find_first_of("0123456789");
  if (self != "gbo" && digits != std::string::npos) {
    o.cmd = self.substr(0, digits);
    o.bs = std::atoi(self.c_str() + digits);
  } else {
    if (argc < 2) usage();
    o.cmd = argv[argi++];
  }
  std::vector<std::string> pos;
  for (; argi < argc; ++argi) {
    std::string a = argv[argi];
    auto val = [&]() -> const char* {
      if (argi + 1 >= argc) usage();
      return argv[++argi];
    };
    if (a == "--bs") o.bs = std::atoi(val());
    else if (a == "--seed") o.seed = std::strtoull(val(), nullptr, 10);
    else if (a == "--source") o.source = std::atoi(val());
    else if (a == "--max-iter") o.max_iter = std::atoi(val());
    else if (a == "--tol") o.tol = std::atof(val());
    else if (a == "--out") o.out = val();
    else if (a == "--json") o.json = true;
    else if (a == "--cache") o.cache = true;
    else if (a == "-h" || a == "--help") usage();
    else pos.push_back(a);
  }
  if (pos.empty()) {
    std::printf("./exe [xxx.mtx]\n");  // the reference's message
    return 1;
  }
  o.file = pos[0];
  if (pos.size() > 1) o.directed = std::atoi(pos[1].c_str());
  if (o.cmd == "bsrbmv") return run_bsrbmv(o);
  if (o.cmd == "bsrbmm") return run_bmm(o, false);
  if (o.cmd == "tc") return run_bmm(o, true);
  if (o.cmd == "bfs") return run_bfs(o);
  if (o.cmd == "sssp") return run_sssp(o);
  if (o.cmd == "pr") return run_pr(o);
  if (o.cmd == "cc") return run_cc(o);
  usage();
}
