// Numeric text tables of the mesh and stimulus files (host code).
//
// The reference reads openCARP .pts / .elem / .lon files and .vtx node
// lists with np.loadtxt (torchcor/core/mesh.py:26-57,
// torchcor/core/stimulation.py:27-36), minutes for a mesh of millions of
// elements.  Here a file is memory-mapped, its data lines are split into
// one contiguous range per host thread at line boundaries, each thread
// counts its rows, and after a prefix sum each parses its rows straight
// into the output (std::from_chars: correctly rounded, like Python's float()).
//
// Table semantics follow np.loadtxt: `skip` header lines are dropped, blank
// lines and text after '#' are ignored, tokens are separated by blanks or
// tabs, and every data row must have as many tokens as the first one.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "tc_internal.h"

namespace tc {
namespace {

struct Mapped {
  const char* p = nullptr;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p && n) munmap((void*)p, n);
    if (fd >= 0) close(fd);
  }
};

int map_file(const char* path, Mapped& m) {
  TC_CHECK(path, TC_ERR_ARG, "table reader: null path");
  m.fd = open(path, O_RDONLY);
  TC_CHECK(m.fd >= 0, TC_ERR_ARG, "table reader: cannot open %s", path);
  struct stat st;
  TC_CHECK(fstat(m.fd, &st) == 0, TC_ERR_ARG, "table reader: cannot stat %s", path);
  m.n = (size_t)st.st_size;
  if (m.n == 0) return TC_OK;
  void* p = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
  TC_CHECK(p != MAP_FAILED, TC_ERR_ARG, "table reader: cannot map %s", path);
  m.p = (const char*)p;
  return TC_OK;
}

inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Tokens of the line [b, e) (comments cut), as (begin, end) pairs.
inline int tokens(const char* b, const char* e, const char** tb, const char** te, int cap) {
  int k = 0;
  const char* q = b;
  while (q < e) {
    while (q < e && blank(*q)) ++q;
    if (q >= e || *q == '#') break;
    const char* s = q;
    while (q < e && !blank(*q) && *q != '#') ++q;
    if (k < cap) {
      tb[k] = s;
      te[k] = q;
    }
    ++k;
  }
  return k;
}

// Start of the data: after `skip` lines.
const char* data_start(const Mapped& m, int skip) {
  const char* q = m.p;
  const char* end = m.p + m.n;
  for (int i = 0; i < skip && q < end; ++i) {
    const char* nl = (const char*)memchr(q, '\n', end - q);
    q = nl ? nl + 1 : end;
  }
  return q;
}

// Thread ranges of [d, end) cut at line starts.
std::vector<const char*> split(const char* d, const char* end, int parts) {
  std::vector<const char*> cut(parts + 1, end);
  cut[0] = d;
  const size_t len = end - d;
  for (int t = 1; t < parts; ++t) {
    const char* q = d + len * t / parts;
    if (q < cut[t - 1]) q = cut[t - 1];
    const char* nl = q < end ? (const char*)memchr(q, '\n', end - q) : nullptr;
    cut[t] = nl ? nl + 1 : end;
  }
  return cut;
}

int threads_for(size_t bytes) {
  unsigned hw = std::thread::hardware_concurrency();
  int t = hw ? (int)std::min(hw, 16u) : 4;
  const int by_size = (int)(bytes / (1 << 20)) + 1;  // at least 1 MiB per thread
  return std::max(1, std::min(t, by_size));
}

// Rows of [b, e) and their token counts (all must equal `cols`, or set
// *cols when it is -1 from the first row met).  Returns the row count or
// -1 (ragged; *bad_row = index of the offending row in this range).
int64_t count_rows(const char* b, const char* e, int* cols, int64_t* bad_row) {
  const char* tb[1];
  const char* te[1];
  int64_t rows = 0;
  const char* q = b;
  while (q < e) {
    const char* nl = (const char*)memchr(q, '\n', e - q);
    const char* le = nl ? nl : e;
    const int k = tokens(q, le, tb, te, 0);
    if (k > 0) {
      if (*cols < 0) *cols = k;
      else if (k != *cols) {
        *bad_row = rows;
        return -1;
      }
      ++rows;
    }
    q = nl ? nl + 1 : e;
  }
  return rows;
}

template <class V>
bool parse_token(const char* b, const char* e, V& v) {
  if (b < e && *b == '+') ++b;  // from_chars takes no leading '+'
  auto r = std::from_chars(b, e, v);
  return r.ec == std::errc() && r.ptr == e;
}

template <class V>
int read_table(const char* path, int skip, int col0, int ncols, int64_t rows, V* out) {
  TC_CHECK(skip >= 0 && col0 >= 0 && ncols > 0 && rows >= 0 && (rows == 0 || out), TC_ERR_ARG,
           "table reader: bad arguments");
  Mapped m;
  int rc = map_file(path, m);
  if (rc) return rc;
  if (rows == 0) return TC_OK;
  const char* d = data_start(m, skip);
  const char* end = m.p + m.n;
  const int T = threads_for(end - d);
  const std::vector<const char*> cut = split(d, end, T);
  std::vector<int64_t> cnt(T, 0), bad(T, -1);
  std::vector<int> cols(T, -1);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
      th.emplace_back([&, t] { cnt[t] = count_rows(cut[t], cut[t + 1], &cols[t], &bad[t]); });
    for (auto& x : th) x.join();
  }
  int width = -1;
  int64_t total = 0;
  for (int t = 0; t < T; ++t) {
    TC_CHECK(cnt[t] >= 0, TC_ERR_ARG, "table reader: %s: row %lld has a different number of "
             "columns", path, (long long)(total + bad[t] + 1));
    if (cnt[t] > 0) {
      if (width < 0) width = cols[t];
      TC_CHECK(cols[t] == width, TC_ERR_ARG, "table reader: %s: rows with %d and %d columns",
               path, width, cols[t]);
    }
    total += cnt[t];
  }
  TC_CHECK(total == rows, TC_ERR_ARG, "table reader: %s holds %lld rows, %lld expected", path,
           (long long)total, (long long)rows);
  TC_CHECK(col0 + ncols <= width, TC_ERR_ARG, "table reader: %s has %d columns, %d needed",
           path, width, col0 + ncols);
  std::vector<int64_t> first(T, 0);
  for (int t = 1; t < T; ++t) first[t] = first[t - 1] + cnt[t - 1];
  std::vector<int64_t> err(T, -1);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
      th.emplace_back([&, t] {
        std::vector<const char*> tb(width), te(width);
        int64_t r = first[t];
        const char* q = cut[t];
        const char* e = cut[t + 1];
        while (q < e) {
          const char* nl = (const char*)memchr(q, '\n', e - q);
          const char* le = nl ? nl : e;
          const int k = tokens(q, le, tb.data(), te.data(), width);
          if (k > 0) {
            V* o = out + r * ncols;
            for (int c = 0; c < ncols; ++c)
              if (!parse_token(tb[col0 + c], te[col0 + c], o[c])) {
                err[t] = r;
                return;
              }
            ++r;
          }
          q = nl ? nl + 1 : e;
        }
      });
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < T; ++t)
    TC_CHECK(err[t] < 0, TC_ERR_ARG, "table reader: %s: row %lld is not numeric", path,
             (long long)(err[t] + 1));
  return TC_OK;
}

int table_shape(const char* path, int skip, int64_t* rows, int32_t* cols) {
  TC_CHECK(rows && cols && skip >= 0, TC_ERR_ARG, "tc_table_shape: bad arguments");
  Mapped m;
  int rc = map_file(path, m);
  if (rc) return rc;
  *rows = 0;
  *cols = 0;
  if (m.n == 0) return TC_OK;
  const char* d = data_start(m, skip);
  const char* end = m.p + m.n;
  const int T = threads_for(end - d);
  const std::vector<const char*> cut = split(d, end, T);
  std::vector<int64_t> cnt(T, 0), bad(T, -1);
  std::vector<int> c(T, -1);
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] { cnt[t] = count_rows(cut[t], cut[t + 1], &c[t], &bad[t]); });
  for (auto& x : th) x.join();
  int width = -1;
  int64_t total = 0;
  for (int t = 0; t < T; ++t) {
    TC_CHECK(cnt[t] >= 0, TC_ERR_ARG, "tc_table_shape: %s: row %lld has a different number of "
             "columns", path, (long long)(total + bad[t] + 1));
    if (cnt[t] > 0) {
      if (width < 0) width = c[t];
      TC_CHECK(c[t] == width, TC_ERR_ARG, "tc_table_shape: %s: rows with %d and %d columns",
               path, width, c[t]);
    }
    total += cnt[t];
  }
  *rows = total;
  *cols = width < 0 ? 0 : width;
  return TC_OK;
}

}  // namespace
}  // namespace tc

template <class F>
static int guarded(F f) {
  try {
    return f();
  } catch (const std::exception& e) {
    tc::set_error("table reader: %s", e.what());
    return TC_ERR_SIZE;
  }
}

extern "C" int tc_table_shape(const char* path, int skip_lines, int64_t* rows_host,
                              int32_t* cols_host) {
  return guarded([&] { return tc::table_shape(path, skip_lines, rows_host, cols_host); });
}

extern "C" int tc_table_read_f64(const char* path, int skip_lines, int col0, int ncols,
                                 int64_t rows, double* out_host) {
  return guarded([&] { return tc::read_table(path, skip_lines, col0, ncols, rows, out_host); });
}

extern "C" int tc_table_read_i64(const char* path, int skip_lines, int col0, int ncols,
                                 int64_t rows, int64_t* out_host) {
  return guarded([&] { return tc::read_table(path, skip_lines, col0, ncols, rows, out_host); });
}
