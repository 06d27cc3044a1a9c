// Matrix Market ingest: a parallel reader with the two reference semantics,
// a binary CSR cache beside the input file, and a parallel writer.  Host
// only (g++ + OpenMP).
//
// Reader structure (both semantics):
//   1. the file is read whole into memory with parallel preads;
//   2. the banner and size line are parsed (mm_read_banner /
//      mm_read_mtx_crd_size, algorithm/mmio.h);
//   3. the data section is split into byte ranges, one per task; a token
//      belongs to the range holding its first byte.  Pass one counts each
//      range's tokens, an exclusive scan gives every range its first token
//      index, pass two parses.  Token t is field t % F of entry t / F (F = 2
//      for pattern, 3 for real/integer, 4 for complex), which is exactly how
//      a sequential fscanf loop pairs whitespace-separated tokens, whatever
//      the line structure; tokens past the header's entry count are ignored,
//      as that loop never reads them;
//   4. rows are built by a stable parallel counting sort over the entry
//      sequence (per-range row histograms), so entries of a row keep their
//      sequence order; semantics 1 then sorts each row by column (stable)
//      and compacts it.
// Errors: the entry with the lowest index that fails (malformed token, out
// of bounds) is reported, as a sequential reader would report it.
#include <fcntl.h>
#include <omp.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "gbo_internal.hpp"

namespace gbo {
namespace {

template <class T>
T* host_alloc(size_t n) {
  T* p = static_cast<T*>(std::malloc(std::max<size_t>(n, 1) * sizeof(T)));
  if (!p) throw std::bad_alloc();
  return p;
}

// Library-allocated CSR handed to the caller (released with gbo_free_host);
// freed here on every path that does not hand it over.
struct CsrOut {
  int32_t m = 0, n = 0;
  int64_t nnz = 0;
  int32_t* rp = nullptr;
  int32_t* ci = nullptr;
  float* cv = nullptr;
  ~CsrOut() {
    std::free(rp);
    std::free(ci);
    std::free(cv);
  }
  void give(int32_t* nrows, int32_t* ncols, int64_t* nnz_out, int32_t** h_rp, int32_t** h_ci, float** h_val) {
    *nrows = m;
    *ncols = n;
    *nnz_out = nnz;
    *h_rp = rp;
    *h_ci = ci;
    rp = ci = nullptr;
    if (h_val) {
      *h_val = cv;
      cv = nullptr;
    }
  }
};

inline bool is_ws(char ch) {
  return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f';
}

// ---------------------------------------------------------------- file I/O
struct FileBuf {
  char* data = nullptr;  // size + 1 bytes, NUL-terminated
  size_t size = 0;
  ~FileBuf() { std::free(data); }
};

// Reads [0, limit) of the file (all of it when limit == 0) with parallel
// preads of 32 MiB pieces.
int read_file(const char* path, size_t limit, FileBuf* fb, struct stat* st_out) {
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return set_error(GBO_EIO, "file %s not found", path);
  struct stat st;
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    return set_error(GBO_EIO, "cannot stat %s", path);
  }
  if (st_out) *st_out = st;
  size_t size = (size_t)st.st_size;
  if (limit && size > limit) size = limit;
  fb->data = static_cast<char*>(std::malloc(size + 1));
  if (!fb->data) {
    ::close(fd);
    throw std::bad_alloc();
  }
  fb->size = size;
  const size_t piece = (size_t)32 << 20;
  const int64_t npieces = (int64_t)((size + piece - 1) / piece);
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : failed)
  for (int64_t k = 0; k < npieces; ++k) {
    size_t off = (size_t)k * piece, end = std::min(size, off + piece);
    while (off < end) {
      const ssize_t got = ::pread(fd, fb->data + off, end - off, (off_t)off);
      if (got < 0 && errno == EINTR) continue;
      if (got <= 0) {
        failed = 1;
        break;
      }
      off += (size_t)got;
    }
  }
  ::close(fd);
  if (failed) return set_error(GBO_EIO, "read error on %s", path);
  fb->data[size] = 0;
  return GBO_OK;
}

// ------------------------------------------------------------------ header
struct MtxHeader {
  char field = 'R';     // R real, C complex, I integer, P pattern
  char symmetry = 'G';  // G general, S symmetric, H hermitian, K skew
  int64_t m = 0, n = 0, nnz = 0;
  size_t data_off = 0;  // first byte after the size line
};

std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

// Next line of [p, end): returns its text (without the newline) and moves p
// past it; false at the end of the buffer.
bool next_line(const char*& p, const char* end, std::string* line) {
  if (p >= end) return false;
  const char* e = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
  if (!e) e = end;
  line->assign(p, (size_t)(e - p));
  p = (e < end) ? e + 1 : end;
  return true;
}

// Banner (mm_read_banner) and size line (mm_read_mtx_crd_size) of the
// buffer.  `truncated` is set when the buffer ends inside the header (only
// possible when a prefix of the file was read).
int parse_header(const char* buf, size_t size, MtxHeader* h, bool* truncated) {
  const char* p = buf;
  const char* end = buf + size;
  std::string line;
  if (truncated) *truncated = false;
  if (!next_line(p, end, &line)) return set_error(GBO_EFORMAT, "empty file");
  char b0[64] = {0}, b1[64] = {0}, b2[64] = {0}, b3[64] = {0}, b4[64] = {0};
  if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", b0, b1, b2, b3, b4) != 5)
    return set_error(GBO_EFORMAT, "could not process Matrix Market banner");
  if (std::string(b0) != "%%MatrixMarket" || lower(b1) != "matrix")
    return set_error(GBO_EFORMAT, "could not process Matrix Market banner");
  const std::string fmt = lower(b2), fld = lower(b3), sym = lower(b4);
  if (fmt != "coordinate") return set_error(GBO_EFORMAT, "only coordinate format is supported");
  if (fld == "real") h->field = 'R';
  else if (fld == "complex") h->field = 'C';
  else if (fld == "integer") h->field = 'I';
  else if (fld == "pattern") h->field = 'P';
  else return set_error(GBO_EFORMAT, "unknown field type '%s'", b3);
  if (sym == "general") h->symmetry = 'G';
  else if (sym == "symmetric") h->symmetry = 'S';
  else if (sym == "hermitian") h->symmetry = 'H';
  else if (sym == "skew-symmetric") h->symmetry = 'K';
  else return set_error(GBO_EFORMAT, "unknown symmetry '%s'", b4);
  while (true) {  // comments and blank lines, then the size line
    if (!next_line(p, end, &line)) {
      if (truncated) *truncated = true;
      return set_error(GBO_EFORMAT, "missing size line");
    }
    if (truncated && p == end) {  // size line may continue past a prefix
      *truncated = true;
      return set_error(GBO_EFORMAT, "missing size line");
    }
    const char* q = line.c_str();
    while (*q == ' ' || *q == '\t') ++q;
    if (*q == '%' || *q == '\r' || *q == 0) continue;
    long long m, n, nz;
    if (std::sscanf(q, "%lld %lld %lld", &m, &n, &nz) != 3)
      return set_error(GBO_EFORMAT, "malformed size line");
    if (m < 0 || n < 0 || nz < 0 || m >= (1LL << 31) || n >= (1LL << 31) || nz >= (1LL << 31))
      return set_error(GBO_ERANGE, "matrix size out of range");
    h->m = m;
    h->n = n;
    h->nnz = nz;
    h->data_off = (size_t)(p - buf);
    return GBO_OK;
  }
}

// ------------------------------------------------------------- data parse
struct Entries {
  std::vector<int32_t> r, c;  // 0-based
  std::vector<float> v;       // empty unless values were asked for
};

// A failure at token `tok` (kinds below); the lowest token wins.
struct ParseErr {
  int64_t tok = INT64_MAX;
  int kind = 0;  // 1 malformed token, 2 row out of range, 3 column out of range
  long long val = 0;
};

inline bool parse_int(const char* p, const char** q, long long* out) {
  bool neg = false;
  if (*p == '+' || *p == '-') neg = (*p++ == '-');
  const char* s = p;
  long long x = 0;
  while (*p >= '0' && *p <= '9') {
    if (p - s >= 18) return false;  // far outside any valid index
    x = x * 10 + (*p - '0');
    ++p;
  }
  if (p == s || !(is_ws(*p) || *p == 0)) return false;
  *q = p;
  *out = neg ? -x : x;
  return true;
}

// A value token, converted as the reader converts it: readMtx's "%f" goes
// straight to float (single), mmio_data's "%lg" through double.
inline bool parse_real(const char* p, const char** q, bool single, float* out) {
  char* e = nullptr;
  const float x = single ? std::strtof(p, &e) : (float)std::strtod(p, &e);
  if (e == p || !(is_ws(*e) || *e == 0)) return false;
  *q = e;
  *out = x;
  return true;
}

int parse_entries(const char* d, const char* end, const MtxHeader& h, bool want_vals, bool single, Entries* e) {
  const int F = (h.field == 'P') ? 2 : (h.field == 'C') ? 4 : 3;
  const int64_t need = h.nnz * F;
  const size_t len = (size_t)(end - d);
  const int64_t nrange =
      std::max<int64_t>(1, std::min<int64_t>((int64_t)omp_get_max_threads() * 8, (int64_t)(len >> 16) + 1));
  std::vector<const char*> b((size_t)nrange + 1);
  for (int64_t i = 0; i <= nrange; ++i) b[(size_t)i] = d + (size_t)((__int128)len * i / nrange);
  auto starts = [&](const char* p) { return !is_ws(*p) && (p == d || is_ws(p[-1])); };
  std::vector<int64_t> ntok((size_t)nrange + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < nrange; ++i) {
    int64_t k = 0;
    for (const char* p = b[(size_t)i]; p < b[(size_t)i + 1]; ++p) k += starts(p);
    ntok[(size_t)i + 1] = k;
  }
  for (int64_t i = 0; i < nrange; ++i) ntok[(size_t)i + 1] += ntok[(size_t)i];
  const int64_t total = ntok[(size_t)nrange];
  // a short file is refused before the entry arrays are sized by its header
  if (total < need) {
    // the lowest malformed token still wins over the missing ones
    for (int64_t i = 0; i < nrange && total < need; ++i) {
      int64_t t = ntok[(size_t)i];
      for (const char* p = b[(size_t)i]; p < b[(size_t)i + 1]; ++p) {
        if (!starts(p)) continue;
        const char* q = p;
        long long xi;
        float xf;
        const int f = (int)(t % F);
        const bool ok = f < 2 ? parse_int(p, &q, &xi) : parse_real(p, &q, single, &xf);
        if (!ok) return set_error(GBO_EFORMAT, "malformed token in entry %lld of the mtx file", (long long)(t / F));
        if (f < 2 && (xi < 1 || xi > (f == 0 ? h.m : h.n)))
          return set_error(GBO_EFORMAT, "entry %lld out of bounds (%s index %lld, size %lld)", (long long)(t / F),
                           f == 0 ? "row" : "column", xi, (long long)(f == 0 ? h.m : h.n));
        ++t;
      }
    }
    return set_error(GBO_EFORMAT, "not enough entries in mtx file (%lld of %lld)", (long long)(total / F),
                     (long long)h.nnz);
  }
  e->r.resize((size_t)h.nnz);
  e->c.resize((size_t)h.nnz);
  if (want_vals) e->v.assign((size_t)h.nnz, 1.0f);
  std::vector<ParseErr> errs((size_t)nrange);
  const int64_t m = h.m, n = h.n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < nrange; ++i) {
    int64_t t = ntok[(size_t)i];
    if (t >= need) continue;
    const char* p = b[(size_t)i];
    const char* stop = b[(size_t)i + 1];
    ParseErr& er = errs[(size_t)i];
    while (t < need) {
      while (p < stop && !starts(p)) ++p;
      if (p >= stop) break;
      const int64_t ent = t / F;
      const int f = (int)(t % F);
      const char* q = p;
      if (f < 2) {
        long long x;
        if (!parse_int(p, &q, &x)) {
          er = ParseErr{t, 1, 0};
          break;
        }
        if (x < 1 || x > (f == 0 ? m : n)) {
          er = ParseErr{t, f == 0 ? 2 : 3, x};
          break;
        }
        (f == 0 ? e->r : e->c)[(size_t)ent] = (int32_t)(x - 1);
      } else {
        float x;
        if (!parse_real(p, &q, single, &x)) {
          er = ParseErr{t, 1, 0};
          break;
        }
        if (f == 2 && want_vals) e->v[(size_t)ent] = x;
      }
      p = q;
      ++t;
    }
  }
  ParseErr first;
  for (const auto& er : errs)
    if (er.tok < first.tok) first = er;
  if (first.kind == 1)
    return set_error(GBO_EFORMAT, "malformed token in entry %lld of the mtx file", (long long)(first.tok / F));
  if (first.kind == 2 || first.kind == 3)
    return set_error(GBO_EFORMAT, "entry %lld out of bounds (%s index %lld, size %lld)",
                     (long long)(first.tok / F), first.kind == 2 ? "row" : "column", first.val,
                     (long long)(first.kind == 2 ? m : n));
  return GBO_OK;
}

// ------------------------------------------------------- row bucketing
// Stable parallel counting sort of a sequence of items by row.  The
// sequence has L slots; emit(k, f) calls f(row, col, val) for the items of
// slot k (zero, one or two), in order.  Items of a row end up in sequence
// order.  Returns the row counts' exclusive scan in `off` (m + 1 entries).
template <class Emit>
int bucket_rows(int32_t m, int64_t L, Emit emit, bool with_vals, std::vector<int64_t>* off,
                std::vector<int32_t>* ci, std::vector<float>* cv) {
  const int64_t P = omp_get_max_threads();
  // per-range row histograms: at most 2^27 counters in all
  int64_t nr = std::max<int64_t>(1, std::min<int64_t>(P, ((int64_t)1 << 27) / std::max<int32_t>(m, 1)));
  nr = std::min<int64_t>(nr, std::max<int64_t>(1, L / 4096));
  std::vector<uint32_t> H((size_t)nr * (size_t)m, 0u);
  auto lo = [&](int64_t i) { return (int64_t)((__int128)L * i / nr); };
#pragma omp parallel for schedule(static, 1)
  for (int64_t i = 0; i < nr; ++i) {
    uint32_t* hi = H.data() + (size_t)i * m;
    for (int64_t k = lo(i); k < lo(i + 1); ++k) emit(k, [&](int32_t r, int32_t, float) { ++hi[r]; });
  }
  off->assign((size_t)m + 1, 0);
  std::vector<int64_t>& o = *off;
#pragma omp parallel for schedule(static)
  for (int32_t r = 0; r < m; ++r) {
    int64_t s = 0;
    for (int64_t i = 0; i < nr; ++i) s += H[(size_t)i * m + r];
    o[(size_t)r + 1] = s;
  }
  for (int32_t r = 0; r < m; ++r) o[(size_t)r + 1] += o[(size_t)r];
  if (o[(size_t)m] >= ((int64_t)1 << 31)) return set_error(GBO_ERANGE, "nnz exceeds int32 CSR");
#pragma omp parallel for schedule(static)
  for (int32_t r = 0; r < m; ++r) {
    uint32_t at = (uint32_t)o[(size_t)r];
    for (int64_t i = 0; i < nr; ++i) {
      const uint32_t c = H[(size_t)i * m + r];
      H[(size_t)i * m + r] = at;
      at += c;
    }
  }
  ci->resize((size_t)o[(size_t)m]);
  if (with_vals) cv->resize((size_t)o[(size_t)m]);
#pragma omp parallel for schedule(static, 1)
  for (int64_t i = 0; i < nr; ++i) {
    uint32_t* hi = H.data() + (size_t)i * m;
    for (int64_t k = lo(i); k < lo(i + 1); ++k)
      emit(k, [&](int32_t r, int32_t c, float v) {
        const uint32_t at = hi[r]++;
        (*ci)[at] = c;
        if (with_vals) (*cv)[at] = v;
      });
  }
  return GBO_OK;
}

// Index of the lowest entry for which bad(k) holds, or -1.
template <class Bad>
int64_t first_bad(int64_t ne, Bad bad) {
  int64_t first = INT64_MAX;
#pragma omp parallel for schedule(static) reduction(min : first)
  for (int64_t k = 0; k < ne; ++k)
    if (bad(k) && k < first) first = k;
  return first == INT64_MAX ? -1 : first;
}

bool drop_selfloops() {
  const char* env = std::getenv("GRB_UTIL_REMOVE_SELFLOOP");
  return env ? (std::atoi(env) != 0) : true;
}

bool is_undirected(const MtxHeader& h, int directed) {
  bool und = (h.symmetry == 'S') || directed == 2;
  return directed == 1 ? false : und;
}

// Entries (file order) -> CSR under semantics 0 or 1.
int build_csr(const MtxHeader& h, const Entries& e, int semantics, int directed, bool want_vals, CsrOut* out) {
  const int32_t m = (int32_t)h.m, n = (int32_t)h.n;
  const int64_t ne = h.nnz;
  const int32_t* R = e.r.data();
  const int32_t* C = e.c.data();
  const float* V = want_vals ? e.v.data() : nullptr;
  std::vector<int64_t> off;
  std::vector<int32_t> ci;
  std::vector<float> cv;
  out->m = m;
  out->n = n;
  if (semantics == 0) {
    // mmio_data: a mirrored entry of a symmetric/hermitian banner lands
    // right after its source, duplicates and self-loops stay.
    const bool sym = (h.symmetry == 'S' || h.symmetry == 'H');
    if (sym && first_bad(ne, [&](int64_t k) { return R[k] != C[k] && C[k] >= m; }) >= 0)
      return set_error(GBO_EFORMAT, "symmetric matrix is not square");
    auto emit = [&](int64_t k, auto&& f) {
      const float v = V ? V[k] : 1.0f;
      f(R[k], C[k], v);
      if (sym && R[k] != C[k]) f(C[k], R[k], v);
    };
    int rc = bucket_rows(m, ne, emit, want_vals, &off, &ci, &cv);
    if (rc != GBO_OK) return rc;
    out->nnz = off[(size_t)m];
    out->rp = host_alloc<int32_t>((size_t)m + 1);
    out->ci = host_alloc<int32_t>((size_t)out->nnz);
    if (want_vals) out->cv = host_alloc<float>((size_t)out->nnz);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r <= m; ++r) out->rp[r] = (int32_t)off[(size_t)r];
    std::memcpy(out->ci, ci.data(), sizeof(int32_t) * (size_t)out->nnz);
    if (want_vals) std::memcpy(out->cv, cv.data(), sizeof(float) * (size_t)out->nnz);
    return GBO_OK;
  }
  // readMtx: the sequence is the file's entries, then (undirected) their
  // mirrors r != c in file order; a stable sort by (row, col) keeps the
  // first of each run of equal keys, self-loops dropped unless
  // GRB_UTIL_REMOVE_SELFLOOP is 0.
  const bool und = is_undirected(h, directed);
  const bool drop = drop_selfloops();
  if (und && first_bad(ne, [&](int64_t k) { return R[k] != C[k] && (C[k] >= m || R[k] >= n); }) >= 0)
    return set_error(GBO_EFORMAT, "undirected matrix is not square");
  auto emit = [&](int64_t k, auto&& f) {
    if (k < ne) {
      f(R[k], C[k], V ? V[k] : 1.0f);
    } else {
      k -= ne;
      if (R[k] != C[k]) f(C[k], R[k], V ? V[k] : 1.0f);
    }
  };
  int rc = bucket_rows(m, und ? 2 * ne : ne, emit, want_vals, &off, &ci, &cv);
  if (rc != GBO_OK) return rc;
  std::vector<int64_t> keep((size_t)m + 1, 0);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, float>> tmp;
#pragma omp for schedule(dynamic, 256)
    for (int32_t r = 0; r < m; ++r) {
      int32_t* c0 = ci.data() + off[(size_t)r];
      const int64_t len = off[(size_t)r + 1] - off[(size_t)r];
      int64_t w = 0;
      if (want_vals) {
        float* v0 = cv.data() + off[(size_t)r];
        tmp.resize((size_t)len);
        for (int64_t k = 0; k < len; ++k) tmp[(size_t)k] = {c0[k], v0[k]};
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const std::pair<int32_t, float>& a, const std::pair<int32_t, float>& b) {
                           return a.first < b.first;
                         });
        for (int64_t k = 0; k < len; ++k) {
          const int32_t c = tmp[(size_t)k].first;
          if (k > 0 && c == tmp[(size_t)k - 1].first) continue;
          if (drop && c == r) continue;
          c0[w] = c;
          v0[w] = tmp[(size_t)k].second;
          ++w;
        }
      } else {
        std::sort(c0, c0 + len);  // equal columns are indistinguishable without values
        for (int64_t k = 0; k < len; ++k) {
          if (k > 0 && c0[k] == c0[k - 1]) continue;
          if (drop && c0[k] == r) continue;
          c0[w++] = c0[k];
        }
      }
      keep[(size_t)r + 1] = w;
    }
  }
  for (int32_t r = 0; r < m; ++r) keep[(size_t)r + 1] += keep[(size_t)r];
  out->nnz = keep[(size_t)m];
  out->rp = host_alloc<int32_t>((size_t)m + 1);
  out->ci = host_alloc<int32_t>((size_t)out->nnz);
  if (want_vals) out->cv = host_alloc<float>((size_t)out->nnz);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t r = 0; r < m; ++r) {
    const int64_t len = keep[(size_t)r + 1] - keep[(size_t)r];
    std::memcpy(out->ci + keep[(size_t)r], ci.data() + off[(size_t)r], sizeof(int32_t) * (size_t)len);
    if (want_vals) std::memcpy(out->cv + keep[(size_t)r], cv.data() + off[(size_t)r], sizeof(float) * (size_t)len);
  }
  for (int32_t r = 0; r <= m; ++r) out->rp[r] = (int32_t)keep[(size_t)r];
  return GBO_OK;
}

int check_args(const char* path, int semantics, int directed, const void* a, const void* b, const void* c,
               const void* d, const void* e) {
  if (!path || !a || !b || !c || !d || !e) return set_error(GBO_EINVAL, "gbo_mtx_read: null argument");
  if (semantics != 0 && semantics != 1) return set_error(GBO_EINVAL, "gbo_mtx_read: semantics must be 0 or 1");
  if (directed < 0 || directed > 2) return set_error(GBO_EINVAL, "gbo_mtx_read: directed must be 0, 1 or 2");
  return GBO_OK;
}

int read_mtx(const char* path, int semantics, int directed, bool want_vals, CsrOut* out, struct stat* st) {
  FileBuf fb;
  int rc = read_file(path, 0, &fb, st);
  if (rc != GBO_OK) return rc;
  MtxHeader h;
  rc = parse_header(fb.data, fb.size, &h, nullptr);
  if (rc != GBO_OK) return rc;
  Entries e;
  rc = parse_entries(fb.data + h.data_off, fb.data + fb.size, h, want_vals, semantics == 1 && h.field == 'R', &e);
  if (rc != GBO_OK) return rc;
  std::free(fb.data);  // the text is no longer needed
  fb.data = nullptr;
  return build_csr(h, e, semantics, directed, want_vals, out);
}

// ------------------------------------------------------------ binary cache
// Layout: CacheHeader, rowptr[m + 1], colind[nnz], then val[nnz] when
// has_val.  A cache is valid for one (source size, source mtime, semantics,
// undirected, self-loop rule) tuple; any difference makes the reader parse
// the text again and rewrite the cache.
struct CacheHeader {
  char magic[8];
  uint32_t version;
  uint32_t semantics;
  uint32_t undirected;
  uint32_t drop_loops;
  uint64_t src_size;
  int64_t src_mtime_s;
  int64_t src_mtime_ns;
  int32_t m, n;
  int64_t nnz;
  uint32_t has_val;
  uint32_t reserved;
};
constexpr char kMagic[8] = {'G', 'B', 'O', 'C', 'S', 'R', 0, 1};
constexpr uint32_t kVersion = 1;

// The cache key of semantics 1 is readMtx's own file name for its binary
// cache (convert(), algorithm/readMtx.hpp:227-244): the undirected flag and
// the self-loop rule are part of the name.
int cache_name(const char* path, int semantics, int directed, std::string* name, MtxHeader* h_out) {
  MtxHeader h;
  bool truncated = false;
  int rc;
  {
    FileBuf fb;  // the first MiB holds the header of any ordinary file
    rc = read_file(path, (size_t)1 << 20, &fb, nullptr);
    if (rc != GBO_OK) return rc;
    rc = parse_header(fb.data, fb.size, &h, &truncated);
  }
  if (rc != GBO_OK && truncated) {
    FileBuf fb;
    rc = read_file(path, 0, &fb, nullptr);
    if (rc != GBO_OK) return rc;
    rc = parse_header(fb.data, fb.size, &h, nullptr);
  }
  if (rc != GBO_OK) return rc;
  if (h_out) *h_out = h;
  const std::string p(path);
  const size_t slash = p.find_last_of('/');
  const std::string dir = slash == std::string::npos ? "." : (slash == 0 ? "/" : p.substr(0, slash));
  const std::string base = slash == std::string::npos ? p : p.substr(slash + 1);
  const std::string sep = dir == "/" ? "" : "/";
  if (semantics == 0)
    *name = dir + sep + "." + base + ".mmio.bin";
  else
    *name = dir + sep + "." + base + "." + (is_undirected(h, directed) ? "ud" : "d") + "." +
            (drop_selfloops() ? "nosl" : "sl") + ".bin";
  return GBO_OK;
}

bool read_exact(int fd, void* dst, size_t bytes, off_t at) {
  char* p = static_cast<char*>(dst);
  while (bytes) {
    const ssize_t got = ::pread(fd, p, bytes, at);
    if (got < 0 && errno == EINTR) continue;
    if (got <= 0) return false;
    p += got;
    at += got;
    bytes -= (size_t)got;
  }
  return true;
}

bool write_all(int fd, const void* src, size_t bytes) {
  const char* p = static_cast<const char*>(src);
  while (bytes) {
    const ssize_t put = ::write(fd, p, bytes);
    if (put < 0 && errno == EINTR) continue;
    if (put <= 0) return false;
    p += put;
    bytes -= (size_t)put;
  }
  return true;
}

CacheHeader make_key(int semantics, bool und, const struct stat& st) {
  CacheHeader k;
  std::memset(&k, 0, sizeof(k));
  std::memcpy(k.magic, kMagic, sizeof(kMagic));
  k.version = kVersion;
  k.semantics = (uint32_t)semantics;
  k.undirected = und ? 1u : 0u;
  k.drop_loops = (semantics == 1 && drop_selfloops()) ? 1u : 0u;
  k.src_size = (uint64_t)st.st_size;
  k.src_mtime_s = (int64_t)st.st_mtim.tv_sec;
  k.src_mtime_ns = (int64_t)st.st_mtim.tv_nsec;
  return k;
}

// Loads the cache when its key matches; false (nothing allocated) otherwise.
bool load_cache(const std::string& name, const CacheHeader& key, bool want_vals, CsrOut* out) {
  const int fd = ::open(name.c_str(), O_RDONLY);
  if (fd < 0) return false;
  CacheHeader h;
  struct stat st;
  bool ok = read_exact(fd, &h, sizeof(h), 0) && ::fstat(fd, &st) == 0;
  ok = ok && std::memcmp(h.magic, key.magic, 8) == 0 && h.version == key.version &&
       h.semantics == key.semantics && h.undirected == key.undirected && h.drop_loops == key.drop_loops &&
       h.src_size == key.src_size && h.src_mtime_s == key.src_mtime_s && h.src_mtime_ns == key.src_mtime_ns &&
       h.m >= 0 && h.n >= 0 && h.nnz >= 0 && h.nnz < ((int64_t)1 << 31);
  if (ok) {
    const uint64_t want = sizeof(h) + 4ull * ((uint64_t)h.m + 1) + 4ull * (uint64_t)h.nnz * (h.has_val ? 2 : 1);
    ok = (uint64_t)st.st_size == want;
  }
  if (ok) {
    out->m = h.m;
    out->n = h.n;
    out->nnz = h.nnz;
    out->rp = host_alloc<int32_t>((size_t)h.m + 1);
    out->ci = host_alloc<int32_t>((size_t)h.nnz);
    off_t at = sizeof(h);
    ok = read_exact(fd, out->rp, 4 * ((size_t)h.m + 1), at);
    at += 4 * ((off_t)h.m + 1);
    ok = ok && read_exact(fd, out->ci, 4 * (size_t)h.nnz, at);
    at += 4 * (off_t)h.nnz;
    if (ok && want_vals) {
      out->cv = host_alloc<float>((size_t)h.nnz);
      if (h.has_val)
        ok = read_exact(fd, out->cv, 4 * (size_t)h.nnz, at);
      else
        std::fill(out->cv, out->cv + h.nnz, 1.0f);
    }
    ok = ok && out->rp[0] == 0 && out->rp[h.m] == (int32_t)h.nnz;
  }
  if (ok) {
    // a damaged cache (rows out of order, columns out of range) is parsed
    // again instead of reaching the packer
    int bad = 0;
    const int32_t* rp = out->rp;
    const int32_t* ci = out->ci;
    const int32_t m = h.m, n = h.n;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int32_t r = 0; r < m; ++r) {
      const int32_t s = rp[r], e = rp[r + 1];
      if (s < 0 || s > e || e > (int32_t)h.nnz) {
        bad = 1;
        continue;
      }
      for (int32_t k = s; k < e; ++k) bad |= (ci[k] < 0 || ci[k] >= n) ? 1 : 0;
    }
    ok = !bad;
  }
  ::close(fd);
  if (!ok) {
    std::free(out->rp);
    std::free(out->ci);
    std::free(out->cv);
    out->rp = out->ci = nullptr;
    out->cv = nullptr;
  }
  return ok;
}

// Writes the cache through a temporary file renamed into place, so a reader
// never sees a partial cache.  False when the directory is not writable.
bool store_cache(const std::string& name, CacheHeader key, const CsrOut& c, bool has_val) {
  char tmp_suffix[64];
  std::snprintf(tmp_suffix, sizeof(tmp_suffix), ".tmp%ld", (long)::getpid());
  const std::string tmp = name + tmp_suffix;
  const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) return false;
  key.m = c.m;
  key.n = c.n;
  key.nnz = c.nnz;
  key.has_val = has_val ? 1u : 0u;
  bool ok = write_all(fd, &key, sizeof(key)) && write_all(fd, c.rp, 4 * ((size_t)c.m + 1)) &&
            write_all(fd, c.ci, 4 * (size_t)c.nnz) && (!has_val || write_all(fd, c.cv, 4 * (size_t)c.nnz));
  ok = (::close(fd) == 0) && ok;
  if (ok) ok = ::rename(tmp.c_str(), name.c_str()) == 0;
  if (!ok) ::unlink(tmp.c_str());
  return ok;
}

// Decimal formatting of a non-negative int into p; returns the end.
inline char* put_uint(char* p, uint32_t x) {
  char t[12];
  int k = 0;
  do {
    t[k++] = (char)('0' + x % 10);
    x /= 10;
  } while (x);
  while (k) *p++ = t[--k];
  return p;
}

}  // namespace
}  // namespace gbo

using namespace gbo;

extern "C" {

int gbo_mtx_read(const char* path, int semantics, int directed, int32_t* nrows, int32_t* ncols, int64_t* nnz,
                 int32_t** h_rowptr, int32_t** h_colind, float** h_val) {
  GBO_TRY_BEGIN
  int rc = check_args(path, semantics, directed, nrows, ncols, nnz, h_rowptr, h_colind);
  if (rc != GBO_OK) return rc;
  CsrOut out;
  rc = read_mtx(path, semantics, directed, h_val != nullptr, &out, nullptr);
  if (rc != GBO_OK) return rc;
  out.give(nrows, ncols, nnz, h_rowptr, h_colind, h_val);
  return GBO_OK;
  GBO_TRY_END
}

int gbo_mtx_cache_path(const char* path, int semantics, int directed, char* buf, size_t len) {
  GBO_TRY_BEGIN
  if (!path || !buf || len == 0) return set_error(GBO_EINVAL, "gbo_mtx_cache_path: null argument");
  if ((semantics != 0 && semantics != 1) || directed < 0 || directed > 2)
    return set_error(GBO_EINVAL, "gbo_mtx_cache_path: bad semantics or directed");
  std::string name;
  int rc = cache_name(path, semantics, directed, &name, nullptr);
  if (rc != GBO_OK) return rc;
  if (name.size() + 1 > len) return set_error(GBO_ERANGE, "gbo_mtx_cache_path: buffer too small");
  std::memcpy(buf, name.c_str(), name.size() + 1);
  return GBO_OK;
  GBO_TRY_END
}

int gbo_mtx_read_cached(const char* path, int semantics, int directed, int* cache_state, int32_t* nrows,
                        int32_t* ncols, int64_t* nnz, int32_t** h_rowptr, int32_t** h_colind, float** h_val) {
  GBO_TRY_BEGIN
  int rc = check_args(path, semantics, directed, nrows, ncols, nnz, h_rowptr, h_colind);
  if (rc != GBO_OK) return rc;
  std::string name;
  MtxHeader h;
  rc = cache_name(path, semantics, directed, &name, &h);
  if (rc != GBO_OK) return rc;
  struct stat st;
  if (::stat(path, &st) != 0) return set_error(GBO_EIO, "file %s not found", path);
  const CacheHeader key = make_key(semantics, semantics == 1 && is_undirected(h, directed), st);
  CsrOut out;
  if (load_cache(name, key, h_val != nullptr, &out)) {
    if (cache_state) *cache_state = 1;
    out.give(nrows, ncols, nnz, h_rowptr, h_colind, h_val);
    return GBO_OK;
  }
  // values are kept in the cache when the file has them, whoever asks
  const bool vals = h_val != nullptr || h.field != 'P';
  struct stat st2;
  rc = read_mtx(path, semantics, directed, vals, &out, &st2);
  if (rc != GBO_OK) return rc;
  const bool same = st2.st_size == st.st_size && st2.st_mtim.tv_sec == st.st_mtim.tv_sec &&
                    st2.st_mtim.tv_nsec == st.st_mtim.tv_nsec;
  const bool stored = same && store_cache(name, key, out, h.field != 'P');
  if (cache_state) *cache_state = stored ? 0 : 2;
  if (!h_val) {
    std::free(out.cv);
    out.cv = nullptr;
  }
  out.give(nrows, ncols, nnz, h_rowptr, h_colind, h_val);
  return GBO_OK;
  GBO_TRY_END
}

int gbo_mtx_write(const char* path, int32_t nrows, int32_t ncols, const int32_t* rp, const int32_t* ci) {
  GBO_TRY_BEGIN
  if (!path || !rp || nrows < 0 || (!ci && rp[nrows] > 0)) return set_error(GBO_EINVAL, "gbo_mtx_write: null");
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(GBO_EIO, "cannot open %s for writing", path);
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate pattern general\n%d %d %d\n", nrows, ncols, rp[nrows]);
  // Row ranges of about 1M entries are formatted in parallel, then written
  // in order.
  const int64_t nnz = rp[nrows];
  const int64_t nrange = std::max<int64_t>(1, nnz >> 20);
  std::vector<int32_t> r0((size_t)nrange + 1);
  for (int64_t i = 0; i <= nrange; ++i) {
    const int64_t want = nnz * i / nrange;
    r0[(size_t)i] = (int32_t)(std::lower_bound(rp, rp + nrows, (int32_t)want) - rp);
  }
  r0[0] = 0;
  r0[(size_t)nrange] = nrows;
  bool ok = true;
  const int64_t batch = 64;
  for (int64_t b0 = 0; b0 < nrange && ok; b0 += batch) {
    const int64_t b1 = std::min(nrange, b0 + batch);
    std::vector<std::string> text((size_t)(b1 - b0));
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = b0; i < b1; ++i) {
      const int32_t ra = r0[(size_t)i], rb = std::max(ra, r0[(size_t)i + 1]);
      std::string& s = text[(size_t)(i - b0)];
      s.resize((size_t)(rp[rb] - rp[ra]) * 24 + 1);
      char* p = &s[0];
      for (int32_t r = ra; r < rb; ++r)
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
          p = put_uint(p, (uint32_t)r + 1);
          *p++ = ' ';
          p = put_uint(p, (uint32_t)ci[k] + 1);
          *p++ = '\n';
        }
      s.resize((size_t)(p - &s[0]));
    }
    for (const auto& s : text)
      if (!s.empty() && std::fwrite(s.data(), 1, s.size(), f) != s.size()) ok = false;
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) return set_error(GBO_EIO, "write error on %s", path);
  return GBO_OK;
  GBO_TRY_END
}

}  // extern "C"
