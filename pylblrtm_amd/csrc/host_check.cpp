// Host-side checks of libpylblrtm_amd, for a build under AddressSanitizer
// and UndefinedBehaviorSanitizer (the Makefile's `sanitize` target).  Runs
// on the CPU without a GPU and calls no entry point that would reach the
// device: the .par reader (lbl_par_read) on files that cover its paths
// (mmap and read(), one and several threads, every line ending, short,
// non-ASCII and malformed lines, missing files), and the argument
// validation of the C ABI (grid, workspace sizing, the optical-depth, adjoint,
// surrogate and Voigt entry points), which returns before any HIP call.
//
// Exit status 0 when every check holds; a sanitizer finding aborts.
#include <fcntl.h>
#include <signal.h>
#include <stdlib.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "pylblrtm_amd.h"

namespace {

int g_failures = 0;
int g_checks = 0;

#define CHECK(cond)                                                              \
  do {                                                                           \
    ++g_checks;                                                                  \
    if (!(cond)) {                                                               \
      ++g_failures;                                                              \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
    }                                                                            \
  } while (0)

std::string g_dir;

std::string write_file(const char* name, const std::string& text) {
  const std::string path = g_dir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  return path;
}

// One 160-column HITRAN-2004 record: mol_id [0:2], iso_id [2:3], nu0 [3:15],
// Sref [15:25], A [25:35], g_air [35:40], g_self [40:45], Elow [45:55],
// n_air [55:59], shift [59:67], then quantum fields to column 160.
std::string par_line(int mol, int iso, double nu0, double sref, double gair, double gself,
                     double elow, double nair, double shift) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%2d%1d%12.6f%10.3e%10.3e%5.3f%5.3f%10.4f%4.2f%8.5f", mol, iso,
                nu0, sref, 1.0e-2, gair, gself, elow, nair, shift);
  std::string s(buf);
  s.resize(160, ' ');
  return s;
}

struct Parsed {
  int rc = -1;
  int64_t n = -1;
  std::vector<int8_t> status;
  std::vector<int32_t> ids;
  std::vector<double> values;
};

Parsed read_par(const std::string& path) {
  Parsed p;
  p.rc = lbl_par_read(path.c_str(), 0, &p.n, nullptr, nullptr, nullptr);
  if (p.rc != LBL_OK) return p;
  const int64_t cap = p.n > 0 ? p.n : 1;
  p.status.assign(cap, 99);
  p.ids.assign(2 * cap, -7);
  p.values.assign(7 * cap, -7.0);
  p.rc = lbl_par_read(path.c_str(), p.n, &p.n, p.status.data(), p.ids.data(), p.values.data());
  return p;
}

double field(const std::string& line, size_t a, size_t b) {
  return std::strtod(line.substr(a, b - a).c_str(), nullptr);
}

bool values_match(const Parsed& p, int64_t i, const std::string& line) {
  static const size_t cols[7][2] = {{3, 15}, {15, 25}, {35, 40}, {40, 45},
                                    {45, 55}, {55, 59}, {59, 67}};
  if (p.ids[2 * i] != std::atoi(line.substr(0, 2).c_str())) return false;
  if (p.ids[2 * i + 1] != std::atoi(line.substr(2, 1).c_str())) return false;
  for (int f = 0; f < 7; ++f)
    if (p.values[7 * i + f] != field(line, cols[f][0], cols[f][1])) return false;
  return true;
}

void check_par_reader() {
  std::vector<std::string> good;
  for (int i = 0; i < 40; ++i)
    good.push_back(par_line(1 + i % 7, 1 + i % 3, 500.0 + 0.37 * i, 1.5e-21 * (i + 1),
                            0.05 + 0.001 * i, 0.07 + 0.001 * i, 10.0 * i, 0.6 + 0.005 * i,
                            -0.002 + 1e-5 * i));

  // every line ending, and a last line without one
  {
    std::string text;
    const char* ends[3] = {"\n", "\r\n", "\r"};
    for (size_t i = 0; i < good.size(); ++i)
      text += good[i] + (i + 1 < good.size() ? ends[i % 3] : "");
    const Parsed p = read_par(write_file("endings.par", text));
    CHECK(p.rc == LBL_OK && p.n == (int64_t)good.size());
    for (int64_t i = 0; p.rc == LBL_OK && i < p.n; ++i) {
      CHECK(p.status[i] == 9);
      CHECK(values_match(p, i, good[i]));
    }
  }
  // short (154 and 0 columns), exactly 155 columns, non-ASCII, malformed
  {
    std::string short154 = good[0].substr(0, 154);
    std::string exact155 = good[1].substr(0, 155);
    std::string utf8 = good[2];
    utf8[100] = (char)0xC3;
    utf8[101] = (char)0xA9;
    std::string bad_nu = good[3];
    bad_nu.replace(3, 12, "   12x.5    ");
    std::string bad_gself = good[4];
    bad_gself.replace(40, 5, "1__2 ");
    std::string under = good[5];
    under.replace(45, 10, "  1_000.5 ");
    std::string specials = good[6];
    specials.replace(15, 10, "   -INF   ");
    specials.replace(59, 8, "  nan   ");
    std::string text = short154 + "\n\n" + exact155 + "\n" + utf8 + "\n" + bad_nu + "\n" +
                       bad_gself + "\n" + under + "\n" + specials + "\n";
    const Parsed p = read_par(write_file("mixed.par", text));
    CHECK(p.rc == LBL_OK && p.n == 8);
    if (p.rc == LBL_OK && p.n == 8) {
      CHECK(p.status[0] == -1 && p.status[1] == -1);
      CHECK(p.status[2] == 9 && values_match(p, 2, good[1]));
      CHECK(p.status[3] == -2);
      CHECK(p.status[4] == 2 && p.ids[8] == std::atoi(good[3].substr(0, 2).c_str()));
      CHECK(p.status[5] == 5);          // mol, iso, nu0, Sref, g_air convert
      CHECK(p.status[6] == 9 && p.values[7 * 6 + 4] == 1000.5);
      CHECK(p.status[7] == 9 && std::isinf(p.values[7 * 7 + 1]) && p.values[7 * 7 + 1] < 0 &&
            std::isnan(p.values[7 * 7 + 6]));
    }
  }
  // empty file, a file of newlines only
  {
    Parsed p = read_par(write_file("empty.par", ""));
    CHECK(p.rc == LBL_OK && p.n == 0);
    p = read_par(write_file("blank.par", "\n\r\n\r"));
    CHECK(p.rc == LBL_OK && p.n == 3);
    for (int64_t i = 0; p.rc == LBL_OK && i < p.n; ++i) CHECK(p.status[i] == -1);
  }
  // several MiB: the reader splits the file over threads at line starts
  {
    std::string text;
    const int n = 30000;
    text.reserve((size_t)n * 162);
    for (int i = 0; i < n; ++i) text += good[i % good.size()] + ((i % 5 == 0) ? "\r\n" : "\n");
    const Parsed p = read_par(write_file("large.par", text));
    CHECK(p.rc == LBL_OK && p.n == n);
    bool all = p.rc == LBL_OK;
    for (int64_t i = 0; all && i < p.n; ++i)
      all = p.status[i] == 9 && values_match(p, i, good[i % good.size()]);
    CHECK(all);
  }
  // a FIFO is not a regular file: read() instead of mmap.  The reader opens
  // it once per call (count, then parse); the writer serves one call at a
  // time, so the second call never sees the first call's data.
  {
    const std::string path = g_dir + "/pipe.par";
    CHECK(mkfifo(path.c_str(), 0600) == 0);
    const std::string text = good[7] + "\n" + good[8] + "\n";
    std::atomic<int> calls_done{0};
    std::thread writer([&] {
      for (int pass = 0; pass < 2; ++pass) {
        while (calls_done.load() < pass) std::this_thread::yield();
        const int fd = open(path.c_str(), O_WRONLY);
        if (fd < 0) return;
        size_t off = 0;
        while (off < text.size()) {
          const ssize_t w = write(fd, text.data() + off, text.size() - off);
          if (w <= 0) break;
          off += (size_t)w;
        }
        close(fd);
      }
    });
    int64_t n = -1;
    const int rc0 = lbl_par_read(path.c_str(), 0, &n, nullptr, nullptr, nullptr);
    calls_done.store(1);
    int8_t st[2] = {0, 0};
    int32_t ids[4];
    double vals[14];
    int rc1 = LBL_EINVAL;
    if (rc0 == LBL_OK && n == 2) {
      rc1 = lbl_par_read(path.c_str(), 2, &n, st, ids, vals);
      writer.join();
    } else {
      // the writer is waiting to open for the second call: open and close
      // the read end once so that it finishes
      calls_done.store(2);
      const int fd = open(path.c_str(), O_RDONLY);
      if (fd >= 0) close(fd);
      writer.join();
    }
    CHECK(rc0 == LBL_OK && rc1 == LBL_OK && n == 2);
    if (rc1 == LBL_OK) {
      Parsed p;
      p.status.assign(st, st + 2);
      p.ids.assign(ids, ids + 4);
      p.values.assign(vals, vals + 14);
      CHECK(st[0] == 9 && st[1] == 9 && values_match(p, 0, good[7]) && values_match(p, 1, good[8]));
    }
  }
  // failures: missing file, a directory, capacity too small, null arguments
  {
    int64_t n = 0;
    const std::string missing = g_dir + "/" + std::string(600, 'x') + ".par";
    CHECK(lbl_par_read(missing.c_str(), 0, &n, nullptr, nullptr, nullptr) == LBL_EINVAL);
    CHECK(std::strlen(lbl_last_error()) > 0 && std::strlen(lbl_last_error()) < 512);
    CHECK(lbl_par_read(g_dir.c_str(), 0, &n, nullptr, nullptr, nullptr) == LBL_EINVAL);
    const std::string path = write_file("two.par", good[0] + "\n" + good[1] + "\n");
    int8_t st[1];
    int32_t ids[2];
    double vals[7];
    CHECK(lbl_par_read(path.c_str(), 1, &n, st, ids, vals) == LBL_EINVAL);
    CHECK(lbl_par_read(nullptr, 0, &n, nullptr, nullptr, nullptr) == LBL_EINVAL);
    CHECK(lbl_par_read(path.c_str(), 0, nullptr, nullptr, nullptr, nullptr) == LBL_EINVAL);
    CHECK(lbl_par_read(path.c_str(), 2, &n, nullptr, ids, vals) == LBL_EINVAL);
  }
}

// nu_min + i*step, multiply then add (no contraction), as numpy builds it
std::vector<double> affine(double lo, double step, int64_t n) {
#pragma clang fp contract(off)
  std::vector<double> nu((size_t)n);
  for (int64_t i = 0; i < n; ++i) nu[(size_t)i] = lo + (double)i * step;
  return nu;
}

void check_grid() {
  lbl_grid_t g{};
  std::vector<double> nu = affine(500.0, 0.001, 2000001);
  CHECK(lbl_grid_from_nu(nu.data(), (int64_t)nu.size(), 2.0, 25.0, &g) == LBL_OK);
  CHECK(g.n_nu == 2000001 && g.nu_min == 500.0 && g.step == 0.001 && g.tile_samples == 2000);
  nu = affine(2400.0, 0.001, 1001);
  CHECK(lbl_grid_from_nu(nu.data(), (int64_t)nu.size(), 2.0, 25.0, &g) == LBL_OK);
  CHECK(g.tile_samples == 1999);
  const double one[1] = {700.0};
  CHECK(lbl_grid_from_nu(one, 1, 2.0, 25.0, &g) == LBL_OK && g.tile_samples == 2000);
  const double tiny[2] = {0.0, 1e-300};   // a tile of 2e300 points: capped, no overflow
  CHECK(lbl_grid_from_nu(tiny, 2, 2.0, 25.0, &g) == LBL_OK && g.tile_samples == ((int64_t)1 << 62));
  const double decreasing[3] = {3.0, 2.0, 1.0};
  CHECK(lbl_grid_from_nu(decreasing, 3, 2.0, 25.0, &g) == LBL_EINVAL);
  const double uneven[3] = {1.0, 2.0, 4.0};
  CHECK(lbl_grid_from_nu(uneven, 3, 2.0, 25.0, &g) == LBL_EINVAL);
  const double flat[3] = {1.0, 1.0, 1.0};
  CHECK(lbl_grid_from_nu(flat, 3, 2.0, 25.0, &g) == LBL_EINVAL);
  const double with_nan[2] = {1.0, NAN};
  CHECK(lbl_grid_from_nu(with_nan, 2, 2.0, 25.0, &g) == LBL_EINVAL);
  const double with_inf[2] = {1.0, INFINITY};
  CHECK(lbl_grid_from_nu(with_inf, 2, 2.0, 25.0, &g) == LBL_EINVAL);
  CHECK(lbl_grid_from_nu(one, 0, 2.0, 25.0, &g) == LBL_EINVAL);
  CHECK(lbl_grid_from_nu(one, 1, 0.0, 25.0, &g) == LBL_EINVAL);
  CHECK(lbl_grid_from_nu(one, 1, NAN, 25.0, &g) == LBL_EINVAL);
  CHECK(lbl_grid_from_nu(one, 1, 2.0, INFINITY, &g) == LBL_EINVAL);
  CHECK(lbl_grid_from_nu(nullptr, 1, 2.0, 25.0, &g) == LBL_EINVAL);
  CHECK(lbl_grid_from_nu(one, 1, 2.0, 25.0, nullptr) == LBL_EINVAL);
  // a length beyond the size arithmetic is refused before any point is read
  CHECK(lbl_grid_from_nu(one, ((int64_t)1 << 40) + 1, 2.0, 25.0, &g) == LBL_EINVAL);

  // the error text is per thread
  std::vector<std::thread> ts;
  std::vector<int> ok(8, 0);
  for (int t = 0; t < 8; ++t)
    ts.emplace_back([&, t] {
      lbl_grid_t gg{};
      for (int r = 0; r < 200; ++r) {
        if (t % 2) {
          if (lbl_grid_from_nu(uneven, 3, 2.0, 25.0, &gg) != LBL_EINVAL || !lbl_last_error()[0]) return;
        } else {
          if (lbl_grid_from_nu(one, 1, 2.0, 25.0, &gg) != LBL_OK || lbl_last_error()[0]) return;
        }
      }
      ok[t] = 1;
    });
  for (auto& t : ts) t.join();
  for (int t = 0; t < 8; ++t) CHECK(ok[t] == 1);
}

// Host stand-ins for the device arrays: only their being non-null is
// looked at on the paths below, which all return before any HIP call.
struct Problem {
  double d[64] = {};
  int32_t i32[64] = {};
  double mu[2] = {1.0, 0.5};
  float f[64] = {};
  lbl_lines_t lines{};
  lbl_groups_t groups{};
  lbl_layers_t layers{};
  lbl_grid_t grid{};
  lbl_outputs_t out{};
  Problem() {
    lines = {10, d, d, d, d, d, d, d, i32};
    groups = {1, i32, d, i32, d, d};
    layers = {4, 1, d, d, d, d};
    grid = {500.0, 0.001, 2000001, 2000, 25.0, 0.0};
    out = {2, mu, f, nullptr, nullptr};
  }
};

void check_sizes() {
  Problem p;
  size_t b = 0;
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &p.grid, LBL_FP64, &b) == LBL_OK && b > 0);
  size_t bm = 0;
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &p.grid, LBL_MIXED, &bm) == LBL_OK && bm > 0);
  CHECK(lbl_vjp_workspace_bytes(&p.lines, &p.layers, &p.grid, &b) == LBL_OK && b > 0);
  // a tile longer than the grid sizes like the whole grid as one tile
  lbl_grid_t one_tile = p.grid, long_tile = p.grid;
  one_tile.n_nu = long_tile.n_nu = 10001;
  one_tile.tile_samples = 10001;
  long_tile.tile_samples = (int64_t)1 << 62;
  size_t b1 = 0, b2 = 0;
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &one_tile, LBL_MIXED, &b1) == LBL_OK);
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &long_tile, LBL_MIXED, &b2) == LBL_OK);
  CHECK(b1 == b2);
  // the largest sizes the arithmetic accepts, and the first it refuses
  lbl_lines_t big_lines = p.lines;
  lbl_layers_t big_layers = p.layers;
  lbl_grid_t big_grid = p.grid;
  big_lines.n_lines = INT32_MAX;
  big_layers.n_layers = 1 << 19;
  big_grid.n_nu = (int64_t)1 << 40;
  big_grid.tile_samples = 2000;
  CHECK(lbl_workspace_bytes(&big_lines, &big_layers, &big_grid, LBL_FP64, &b) == LBL_OK);
  CHECK(b > ((size_t)1 << 60));
  big_grid.n_nu += 1;
  CHECK(lbl_workspace_bytes(&big_lines, &big_layers, &big_grid, LBL_FP64, &b) == LBL_EINVAL);
  big_grid.n_nu = INT64_MAX;
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &big_grid, LBL_MIXED, &b) == LBL_EINVAL);
  big_lines.n_lines = (int64_t)INT32_MAX + 1;
  CHECK(lbl_workspace_bytes(&big_lines, &p.layers, &p.grid, LBL_MIXED, &b) == LBL_EINVAL);
  big_layers.n_layers = (1 << 19) + 1;
  CHECK(lbl_workspace_bytes(&p.lines, &big_layers, &p.grid, LBL_MIXED, &b) == LBL_EINVAL);
  big_lines.n_lines = -1;
  CHECK(lbl_workspace_bytes(&big_lines, &p.layers, &p.grid, LBL_MIXED, &b) == LBL_EINVAL);
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &p.grid, 7, &b) == LBL_EINVAL);
  CHECK(lbl_workspace_bytes(nullptr, &p.layers, &p.grid, LBL_MIXED, &b) == LBL_EINVAL);
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, nullptr, LBL_MIXED, &b) == LBL_EINVAL);
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &p.grid, LBL_MIXED, nullptr) == LBL_EINVAL);
  CHECK(lbl_vjp_workspace_bytes(&p.lines, &p.layers, nullptr, &b) == LBL_EINVAL);
  lbl_grid_t zero = p.grid;
  zero.tile_samples = 0;
  CHECK(lbl_workspace_bytes(&p.lines, &p.layers, &zero, LBL_MIXED, &b) == LBL_EINVAL);
  size_t tb = 0;
  CHECK(lbl_mlp_train_workspace_bytes(4096, &tb) == LBL_OK && tb > 0);
  CHECK(lbl_mlp_train_workspace_bytes(100, &tb) == LBL_EINVAL);
  CHECK(lbl_mlp_train_workspace_bytes(-64, &tb) == LBL_EINVAL);
  CHECK(lbl_mlp_train_workspace_bytes(64, nullptr) == LBL_EINVAL);
  CHECK(lbl_mlp_num_params() > 0);
}

// Each call must fail its argument check (before any HIP call) with
// LBL_EINVAL and a message.
void expect_einval(int rc, const char* what) {
  ++g_checks;
  if (rc != LBL_EINVAL || !lbl_last_error()[0]) {
    ++g_failures;
    std::fprintf(stderr, "%s: rc %d, message '%s'\n", what, rc, lbl_last_error());
  }
}

void check_entry_points() {
  int32_t el = -1, eg = -1;
  {
    Problem p;
    char ws[256];
    expect_einval(lbl_optical_depth(nullptr, &p.groups, &p.layers, &p.grid, LBL_MIXED, &p.out, ws,
                                    sizeof ws, &el, &eg, nullptr), "null lines");
    expect_einval(lbl_optical_depth(&p.lines, &p.groups, &p.layers, &p.grid, 5, &p.out, ws,
                                    sizeof ws, &el, &eg, nullptr), "unknown precision");
    expect_einval(lbl_optical_depth(&p.lines, &p.groups, &p.layers, &p.grid, LBL_SURROGATE, &p.out,
                                    ws, sizeof ws, &el, &eg, nullptr), "surrogate without network");
    // the workspace check comes before any device work
    expect_einval(lbl_optical_depth(&p.lines, &p.groups, &p.layers, &p.grid, LBL_MIXED, &p.out, ws,
                                    sizeof ws, &el, &eg, nullptr), "workspace too small");
    expect_einval(lbl_optical_depth(&p.lines, &p.groups, &p.layers, &p.grid, LBL_FP64, &p.out,
                                    nullptr, 1u << 30, &el, &eg, nullptr), "no workspace");
    // the fixed scratch of the no-allocation entry point is too small for
    // this grid (checked before the device is looked at)
    expect_einval(lbl_optical_depth_noalloc(&p.lines, &p.groups, &p.layers, &p.grid, LBL_MIXED,
                                            &p.out, &el, &eg, nullptr), "noalloc too large");
    expect_einval(lbl_optical_depth_noalloc(&p.lines, &p.groups, &p.layers, &p.grid, LBL_SURROGATE,
                                            &p.out, &el, &eg, nullptr), "noalloc surrogate");
  }
  auto mutate = [&](const char* what, auto&& change) {
    Problem p;
    change(p);
    char ws[256];
    expect_einval(lbl_optical_depth(&p.lines, &p.groups, &p.layers, &p.grid, LBL_MIXED, &p.out, ws,
                                    sizeof ws, &el, &eg, nullptr), what);
  };
  mutate("n_lines < 0", [](Problem& p) { p.lines.n_lines = -1; });
  mutate("n_lines > INT32_MAX", [](Problem& p) { p.lines.n_lines = (int64_t)INT32_MAX + 1; });
  mutate("n_layers > 2^19", [](Problem& p) { p.layers.n_layers = (1 << 19) + 1; });
  mutate("n_groups > max", [](Problem& p) { p.groups.n_groups = 1 << 20; });
  mutate("n_groups 0 with lines", [](Problem& p) { p.groups.n_groups = 0; });
  mutate("n_gases 0", [](Problem& p) { p.layers.n_gases = 0; });
  mutate("null nu0", [](Problem& p) { p.lines.nu0 = nullptr; });
  mutate("null group index", [](Problem& p) { p.lines.group = nullptr; });
  mutate("null tips", [](Problem& p) { p.groups.tips_q = nullptr; });
  mutate("null p_atm", [](Problem& p) { p.layers.p_atm = nullptr; });
  mutate("null col", [](Problem& p) { p.layers.col = nullptr; });
  mutate("n_nu 0", [](Problem& p) { p.grid.n_nu = 0; });
  mutate("n_nu > 2^40", [](Problem& p) { p.grid.n_nu = ((int64_t)1 << 40) + 1; });
  mutate("tile 0", [](Problem& p) { p.grid.tile_samples = 0; });
  mutate("step 0", [](Problem& p) { p.grid.step = 0.0; });
  mutate("step NaN", [](Problem& p) { p.grid.step = NAN; });
  mutate("nu_min inf", [](Problem& p) { p.grid.nu_min = INFINITY; });
  mutate("cutoff NaN", [](Problem& p) { p.grid.cutoff = NAN; });
  mutate("n_mu 0", [](Problem& p) { p.out.n_mu = 0; });
  mutate("n_mu 9", [](Problem& p) { p.out.n_mu = 9; });
  mutate("mu 0", [](Problem& p) { p.mu[1] = 0.0; });
  mutate("mu > 1", [](Problem& p) { p.mu[0] = 1.5; });
  mutate("mu NaN", [](Problem& p) { p.mu[0] = NAN; });
  mutate("null mu", [](Problem& p) { p.out.mu = nullptr; });
  mutate("null transmittance", [](Problem& p) { p.out.transmittance = nullptr; });
  {
    Problem p;
    lbl_surrogate_t sur{};
    char ws[256];
    expect_einval(lbl_optical_depth_surrogate(&p.lines, &p.groups, &p.layers, &p.grid, &p.out,
                                              &sur, ws, sizeof ws, &el, &eg, nullptr),
                  "surrogate without parameters");
    expect_einval(lbl_transmittance_vjp(&p.lines, &p.groups, &p.layers, &p.grid, 0, p.mu, p.f, p.f,
                                        p.f, p.d, p.d, ws, sizeof ws, nullptr), "vjp n_mu 0");
    expect_einval(lbl_transmittance_vjp(&p.lines, &p.groups, &p.layers, &p.grid, 2, p.mu, p.f, p.f,
                                        p.f, nullptr, p.d, ws, sizeof ws, nullptr), "vjp null grad");
    expect_einval(lbl_transmittance_vjp(&p.lines, &p.groups, &p.layers, &p.grid, 2, p.mu, p.f,
                                        nullptr, p.f, p.d, p.d, ws, sizeof ws, nullptr),
                  "vjp null layer optical depth");
    expect_einval(lbl_transmittance_vjp(nullptr, &p.groups, &p.layers, &p.grid, 2, p.mu, p.f, p.f,
                                        p.f, p.d, p.d, ws, sizeof ws, nullptr), "vjp null lines");
    expect_einval(lbl_voigt_eval(-1, p.d, p.d, p.d, LBL_FP64, nullptr), "voigt n < 0");
    expect_einval(lbl_voigt_eval(4, nullptr, p.d, p.d, LBL_FP64, nullptr), "voigt null x");
    expect_einval(lbl_voigt_eval(4, p.d, p.d, p.d, 3, nullptr), "voigt precision");
    expect_einval(lbl_mlp_eval(-1, p.f, p.f, p.f, nullptr, p.f, nullptr), "mlp eval n < 0");
    expect_einval(lbl_mlp_make_data(8, 1, nullptr, p.f, p.f, nullptr), "mlp data null");
    expect_einval(lbl_mlp_pack(nullptr, nullptr, nullptr), "mlp pack null");
    unsigned short w[4];
    expect_einval(lbl_mlp_train_step(p.f, p.f, p.f, w, p.f, p.f, p.f, 1000, 3, 0, 0, 64, 1e-3f, 1,
                                     p.f, p.f, ws, sizeof ws, nullptr), "train n_data");
    expect_einval(lbl_mlp_train_step(p.f, p.f, p.f, w, p.f, p.f, p.f, 1024, 2, 0, 0, 64, 1e-3f, 1,
                                     p.f, p.f, ws, sizeof ws, nullptr), "train perm_a even");
    expect_einval(lbl_mlp_train_step(p.f, nullptr, p.f, w, p.f, p.f, p.f, 1024, 3, 0, 0, 64, 1e-3f,
                                     1, p.f, p.f, ws, sizeof ws, nullptr), "train null adam");
    expect_einval(lbl_mlp_train_step(p.f, p.f, p.f, w, p.f, p.f, p.f, 1024, 3, 0, 0, 64, 1e-3f, 1,
                                     p.f, p.f, ws, 4, nullptr), "train workspace");
  }
  uint64_t counts[16];
  CHECK(lbl_last_work_count(counts, 16) == LBL_OK && counts[15] == 0);
  CHECK(lbl_last_work_count(nullptr, 4) == LBL_EINVAL);
  CHECK(lbl_last_work_count(counts, -1) == LBL_EINVAL);
  double a = -1, b = -1;
  CHECK(lbl_last_timings(&a, &b) == LBL_OK && lbl_last_timings(nullptr, nullptr) == LBL_OK);
  CHECK(lbl_last_timings_split(&a, nullptr) == LBL_OK && lbl_last_timing_epilogue(&a) == LBL_OK);
}

}  // namespace

int main() {
  signal(SIGPIPE, SIG_IGN);   // a FIFO writer whose reader has gone gets EPIPE
  char tmpl[] = "/tmp/lbl_host_check_XXXXXX";
  if (!mkdtemp(tmpl)) { std::perror("mkdtemp"); return 2; }
  g_dir = tmpl;
  check_par_reader();
  check_grid();
  check_sizes();
  check_entry_points();
  for (const char* f : {"endings.par", "mixed.par", "empty.par", "blank.par", "large.par",
                        "pipe.par", "two.par"})
    unlink((g_dir + "/" + f).c_str());
  rmdir(g_dir.c_str());
  std::printf("host_check: %d checks, %d failed\n", g_checks, g_failures);
  return g_failures == 0 ? 0 : 1;
}
