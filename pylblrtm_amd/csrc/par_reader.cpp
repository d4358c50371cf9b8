// HITRAN-2004 .par line-list reader (host code) with the semantics of
// pylblrtm's par_to_dict (pylblrtm/hitran_manager.py:816-872):
//
// * the file is split into text lines the way Python's universal-newline
//   text mode splits it (\n, \r\n and a lone \r all end a line);
// * a line shorter than 155 characters is skipped;
// * otherwise the fields mol_id [0:2], iso_id [2:3], nu0 [3:15], Sref [15:25],
//   g_air [35:40], g_self [40:45], Elow [45:55], n_air [55:59], shift [59:67]
//   (column layout of hitran_manager.py:663-686) are converted in that order
//   with Python's int() / float() rules, stopping at the first field that
//   does not convert: the reference appends field by field, so a malformed
//   line still contributes the fields before the bad one.
//
// Python accepts, around and inside a number: leading and trailing space,
// \t, \n, \v, \f, \r, one sign, single underscores between digits,
// inf/infinity/nan in any case.  Lines holding non-ASCII bytes are reported, not parsed: the
// caller decodes those as UTF-8 with replacement the way Python does.
#include <fcntl.h>
#include <locale.h>
#include <stdlib.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "internal.h"

namespace {

// The whitespace int() and float() strip: space and \t \n \v \f \r (not
// the \x1c-\x1f that str.isspace() also accepts).
bool py_space(unsigned char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
bool digit(char c) { return c >= '0' && c <= '9'; }

// Strip Python whitespace from [s, s+n); returns false if nothing is left.
bool strip(const char*& s, size_t& n) {
  while (n > 0 && py_space((unsigned char)s[0])) { ++s; --n; }
  while (n > 0 && py_space((unsigned char)s[n - 1])) --n;
  return n > 0;
}

// digit ('_'? digit)* starting at s[i]; appends the digits to out.
// Returns the index after the run, or i if no digit starts there; sets bad
// on a misplaced underscore.
size_t digit_run(const char* s, size_t n, size_t i, char* out, size_t& m, size_t cap, bool& bad) {
  const size_t start = i;
  while (i < n) {
    if (digit(s[i])) {
      if (m + 1 >= cap) { bad = true; return i; }
      out[m++] = s[i++];
    } else if (s[i] == '_' && i > start && digit(s[i - 1]) && i + 1 < n && digit(s[i + 1])) {
      ++i;
    } else {
      break;
    }
  }
  return i;
}

// Python int(field) (base 10) for an ASCII field.
bool py_int(const char* s, size_t n, int32_t* out) {
  if (!strip(s, n)) return false;
  size_t i = 0;
  bool neg = false;
  if (s[0] == '+' || s[0] == '-') { neg = s[0] == '-'; i = 1; }
  char buf[40];
  size_t m = 0;
  bool bad = false;
  const size_t j = digit_run(s, n, i, buf, m, sizeof buf, bad);
  if (bad || j == i || j != n || m > 9) return false;   // fields are 1-2 characters wide
  int64_t v = 0;
  for (size_t k = 0; k < m; ++k) v = v * 10 + (buf[k] - '0');
  *out = (int32_t)(neg ? -v : v);
  return true;
}

// Exact decimal-to-double conversion for a validated "digits[.digits][e[+-]digits]"
// string when the significand has at most 15 significant digits (so it is an
// exact double) and the power of ten is at most 22 (so is 10^|e|): one IEEE
// multiply or divide of two exact values is then correctly rounded, the same
// result strtod gives (Clinger's fast path).  Returns false otherwise.
bool fast_decimal(const char* b, size_t m, double* out) {
  static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                    1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                    1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  uint64_t sig = 0;
  int sig_digits = 0, frac = 0;
  bool in_frac = false;
  size_t i = 0;
  for (; i < m && b[i] != 'e'; ++i) {
    if (b[i] == '.') { in_frac = true; continue; }
    const int d = b[i] - '0';
    if (sig_digits > 0 || d != 0) {
      if (++sig_digits > 15) return false;
    }
    sig = sig * 10 + (uint64_t)d;
    if (in_frac) ++frac;
  }
  int e10 = 0;
  if (i < m) {   // exponent
    ++i;
    bool eneg = false;
    if (i < m && (b[i] == '+' || b[i] == '-')) { eneg = b[i] == '-'; ++i; }
    for (; i < m; ++i) {
      e10 = e10 * 10 + (b[i] - '0');
      if (e10 > 400) return false;
    }
    if (eneg) e10 = -e10;
  }
  const int e = e10 - frac;
  if (e > 22 || e < -22) return false;
  const double x = (double)sig;
  *out = e >= 0 ? x * kPow10[e] : x / kPow10[-e];
  return true;
}

bool iequal(const char* s, size_t n, const char* lit) {
  if (std::strlen(lit) != n) return false;
  for (size_t k = 0; k < n; ++k) {
    char c = s[k];
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    if (c != lit[k]) return false;
  }
  return true;
}

// Python float(field) for an ASCII field: the decimal grammar
// [sign] (digits [. [digits]] | . digits) [(e|E) [sign] digits], digits
// possibly split by single underscores, or [sign] inf|infinity|nan.
bool py_float(const char* s, size_t n, double* out) {
  if (!strip(s, n)) return false;
  size_t i = 0;
  bool neg = false;
  if (s[0] == '+' || s[0] == '-') { neg = s[0] == '-'; i = 1; }
  if (iequal(s + i, n - i, "inf") || iequal(s + i, n - i, "infinity")) {
    *out = neg ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::infinity();
    return true;
  }
  if (iequal(s + i, n - i, "nan")) {
    *out = neg ? -std::numeric_limits<double>::quiet_NaN() : std::numeric_limits<double>::quiet_NaN();
    return true;
  }
  char buf[96];
  size_t m = 0;
  bool bad = false;
  size_t j = digit_run(s, n, i, buf, m, sizeof buf, bad);
  bool mant = j > i;
  if (j < n && s[j] == '.') {
    buf[m++] = '.';
    const size_t k = digit_run(s, n, j + 1, buf, m, sizeof buf, bad);
    mant = mant || k > j + 1;
    j = k;
  }
  if (bad || !mant) return false;
  if (j < n && (s[j] == 'e' || s[j] == 'E')) {
    buf[m++] = 'e';
    size_t k = j + 1;
    if (k < n && (s[k] == '+' || s[k] == '-')) buf[m++] = s[k++];
    const size_t e = digit_run(s, n, k, buf, m, sizeof buf, bad);
    if (bad || e == k) return false;
    j = e;
  }
  if (j != n || m >= sizeof buf) return false;
  buf[m] = '\0';
  double v;
  if (!fast_decimal(buf, m, &v)) {
    // glibc's strtod is correctly rounded (as Python's float() is) and gives
    // inf / 0 on overflow / underflow; a fixed "C" locale keeps '.' the point
    static const locale_t c_locale = newlocale(LC_ALL_MASK, "C", (locale_t)0);
    char* end = nullptr;
    v = strtod_l(buf, &end, c_locale);
    if (end != buf + m) return false;
  }
  *out = neg ? -v : v;
  return true;
}

struct Field { int a, b; };
constexpr Field kInts[2] = {{0, 2}, {2, 3}};
constexpr Field kFloats[7] = {{3, 15}, {15, 25}, {35, 40}, {40, 45}, {45, 55}, {55, 59}, {59, 67}};
constexpr size_t kMinLen = 155;

// One text line [s, s + len): its status and fields.
void parse_line(const char* s, size_t len, int8_t* status, int32_t* id, double* v) {
  bool ascii = true;
  for (size_t k = 0; k < len; ++k) ascii &= (unsigned char)s[k] < 0x80;
  if (!ascii) { *status = -2; return; }
  if (len < kMinLen) { *status = -1; return; }
  int8_t st = 0;
  bool ok = true;
  for (int k = 0; k < 2 && ok; ++k)
    if ((ok = py_int(s + kInts[k].a, kInts[k].b - kInts[k].a, id + k))) ++st;
  for (int k = 0; k < 7 && ok; ++k)
    if ((ok = py_float(s + kFloats[k].a, kFloats[k].b - kFloats[k].a, v + k))) ++st;
  *status = st;
}

// End of the line that starts at pos (index of its terminator) and the start
// of the next line; \n, \r\n and \r end a line.
inline size_t line_end(const char* p, size_t n, size_t pos, size_t* next) {
  size_t end = pos;
  while (end < n && p[end] != '\n' && p[end] != '\r') ++end;
  size_t nx = end;
  if (nx < n) nx += (p[nx] == '\r' && nx + 1 < n && p[nx + 1] == '\n') ? 2 : 1;
  *next = nx;
  return end;
}

// Split [0, n) into up to `parts` ranges that begin at line starts.
std::vector<size_t> split_at_lines(const char* p, size_t n, int parts) {
  std::vector<size_t> cut{0};
  for (int k = 1; k < parts; ++k) {
    size_t c = std::max(cut.back(), n * (size_t)k / (size_t)parts);
    if (c > 0 && c < n) {   // advance to the start of the next line
      if (p[c - 1] == '\n' || (p[c - 1] == '\r' && p[c] != '\n')) {
        // c already starts a line
      } else {
        size_t nx;
        line_end(p, n, c, &nx);
        c = nx;
      }
    }
    if (c > cut.back() && c < n) cut.push_back(c);
  }
  cut.push_back(n);
  return cut;
}

int64_t count_lines(const char* p, size_t a, size_t b) {
  int64_t lines = 0;
  size_t pos = a, nx;
  while (pos < b) {
    line_end(p, b, pos, &nx);
    ++lines;
    pos = nx;
  }
  return lines;
}

}  // namespace

extern "C" int lbl_par_read(const char* path, int64_t capacity, int64_t* n_text_lines,
                            int8_t* status, int32_t* ids, double* values) {
  lbl::set_error(LBL_OK, "%s", "");
  if (!path || !n_text_lines || (capacity > 0 && (!status || !ids || !values)))
    return lbl::set_error(LBL_EINVAL, "lbl_par_read: null argument");
  // A regular file is mapped (no copy); anything else is read into memory.
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return lbl::set_error(LBL_EINVAL, "cannot open %s", path);
  struct stat sb;
  const char* p = nullptr;
  size_t n = 0;
  void* map = MAP_FAILED;
  std::vector<char> data;
  if (fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size > 0) {
    map = mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (map != MAP_FAILED) { p = static_cast<const char*>(map); n = (size_t)sb.st_size; }
  }
  if (map == MAP_FAILED) {
    char chunk[1 << 16];
    ssize_t got;
    while ((got = read(fd, chunk, sizeof chunk)) > 0) data.insert(data.end(), chunk, chunk + got);
    if (got < 0) { close(fd); return lbl::set_error(LBL_EINVAL, "read error on %s", path); }
    p = data.data();
    n = data.size();
  }
  close(fd);
  struct Unmap {
    void* m; size_t n;
    ~Unmap() { if (m != MAP_FAILED) munmap(m, n); }
  } unmap{map, n};

  // Lines are independent: chunks that start at line starts go to threads,
  // first to count their lines, then (with the running offsets) to parse them.
  const unsigned hw = std::thread::hardware_concurrency();
  const int threads = (int)std::max(1u, std::min(hw == 0 ? 1u : hw, 16u));
  const std::vector<size_t> cut = split_at_lines(p, n, n < (1u << 20) ? 1 : threads);
  const int parts = (int)cut.size() - 1;
  std::vector<int64_t> counts(parts, 0);
  auto run = [&](auto&& fn) {
    std::vector<std::thread> pool;
    for (int k = 1; k < parts; ++k) pool.emplace_back(fn, k);
    if (parts > 0) fn(0);
    for (auto& t : pool) t.join();
  };
  run([&](int k) { counts[k] = count_lines(p, cut[k], cut[k + 1]); });
  int64_t total = 0;
  std::vector<int64_t> first(parts, 0);
  for (int k = 0; k < parts; ++k) { first[k] = total; total += counts[k]; }
  *n_text_lines = total;
  if (capacity <= 0) return LBL_OK;
  if (capacity < total) return lbl::set_error(LBL_EINVAL, "lbl_par_read: capacity too small");
  run([&](int k) {
    int64_t line = first[k];
    size_t pos = cut[k], nx;
    while (pos < cut[k + 1]) {
      const size_t end = line_end(p, cut[k + 1], pos, &nx);
      parse_line(p + pos, end - pos, status + line, ids + 2 * line, values + 7 * line);
      ++line;
      pos = nx;
    }
  });
  return LBL_OK;
}
