// host_checks.cpp -- the library's host-side code (decomposition metadata,
// process-grid choice, status handling, profiler regions, synthetic-input
// generators: csrc/api.cpp) and the oracle's C restatement (oracle/*.c),
// built with AddressSanitizer and UndefinedBehaviorSanitizer and driven over
// many grids.  Kernel launchers are stubbed: nothing here touches a GPU.
// Built and run by tests/test_host_sanitizers.py; prints "host checks passed"
// and exits 0, or reports the first mismatch and exits 1 (a sanitizer report
// aborts the run on its own).
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "d2d_oracle.h"
#include "d2dhip.h"
#include "internal.h"

// kernel launchers (copy_kernels.hip, halo_kernels.hip) are not part of this
// build
namespace d2dhip {
int launch_pack(int, int, int, const void *, void *, const d2dhip_decomp *, int, hipStream_t) {
  return D2DHIP_ERR_UNSUPPORTED;
}
int launch_halo(int, const void *, void *, int, int, int, int, int, const int[3], hipStream_t) {
  return D2DHIP_ERR_UNSUPPORTED;
}
int launch_fill_uniform(void *, long long, int, unsigned long long, hipStream_t) {
  return D2DHIP_ERR_UNSUPPORTED;
}
int launch_column_copy(int, const void *, void *, long long, long long, long long, long long, int,
                       hipStream_t) {
  return D2DHIP_ERR_UNSUPPORTED;
}
int launch_fill_modal(void *, int, int, int, int, const d2dhip_modes *, hipStream_t) {
  return D2DHIP_ERR_UNSUPPORTED;
}
}  // namespace d2dhip

static int g_fail = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      std::fprintf(stderr, "mismatch: " __VA_ARGS__);     \
      std::fprintf(stderr, "  (%s:%d)\n", __FILE__, __LINE__); \
      if (++g_fail > 10) std::exit(1);                    \
    }                                                     \
  } while (0)

static bool same3(const int *a, const int *b) { return !std::memcmp(a, b, 3 * sizeof(int)); }

// d2dhip_decomp_info_init against the oracle's restatement, every rank
static void check_decomp(int nx, int ny, int nz, int pr, int pc) {
  for (int r = 0; r < pr; ++r)
    for (int c = 0; c < pc; ++c) {
      d2dhip_decomp d;
      oracle_decomp o;
      const int s1 = d2dhip_decomp_info_init(nx, ny, nz, pr, pc, r, c, &d);
      const int s2 = oracle_decomp_info_init(nx, ny, nz, pr, pc, r, c, &o);
      CHECK((s1 == 0) == (s2 == 0), "status %d vs %d at %d %d %d / %dx%d", s1, s2, nx, ny, nz, pr, pc);
      if (s1 || s2) continue;
      CHECK(same3(d.xst, o.xst) && same3(d.xen, o.xen) && same3(d.xsz, o.xsz), "x pencil");
      CHECK(same3(d.yst, o.yst) && same3(d.yen, o.yen) && same3(d.ysz, o.ysz), "y pencil");
      CHECK(same3(d.zst, o.zst) && same3(d.zen, o.zen) && same3(d.zsz, o.zsz), "z pencil");
      for (int m = 0; m < pr; ++m)
        CHECK(d.x1dist[m] == o.x1dist[m] && d.y1dist[m] == o.y1dist[m] &&
                  d.x1cnts[m] == o.x1cnts[m] && d.x1disp[m] == o.x1disp[m] &&
                  d.y1cnts[m] == o.y1cnts[m] && d.y1disp[m] == o.y1disp[m],
              "row blocks");
      for (int m = 0; m < pc; ++m)
        CHECK(d.y2dist[m] == o.y2dist[m] && d.z2dist[m] == o.z2dist[m] &&
                  d.y2cnts[m] == o.y2cnts[m] && d.y2disp[m] == o.y2disp[m] &&
                  d.z2cnts[m] == o.z2cnts[m] && d.z2disp[m] == o.z2disp[m],
              "column blocks");
      CHECK(d.x1count == o.x1count && d.y1count == o.y1count && d.y2count == o.y2count &&
                d.z2count == o.z2count && d.even == o.even,
            "counts");
      for (int dir = 0; dir < 4; ++dir)
        for (int even = 0; even < 2; ++even) CHECK(d2dhip_pack_buffer_elems(dir, &d, even) >= 0, "buffer");
      // a refused multi-rank transpose and the stubbed packs: argument checks only
      if (pr * pc > 1) {
        const int st = d2dhip_transpose(D2DHIP_X_TO_Y, D2DHIP_C16, &d, &d, &d, nullptr);
        CHECK(st == D2DHIP_ERR_UNSUPPORTED || pr == 1, "transpose status %d", st);
      }
    }
}

// one rank's split + merge through the oracle's maps: a pencil survives
static void check_pack_maps(int nx, int ny, int nz, int pr, int pc) {
  oracle_decomp o;
  if (oracle_decomp_info_init(nx, ny, nz, pr, pc, pr - 1, pc - 1, &o)) return;
  for (int even = 0; even < 2; ++even) {
    const int *src[4] = {o.xsz, o.ysz, o.zsz, o.ysz};
    for (int dir = 0; dir < 4; ++dir) {
      const int *s = src[dir];
      const size_t n = (size_t)s[0] * s[1] * s[2];
      const bool xy = dir == 0 || dir == 3;
      const size_t nb = even ? (size_t)(xy ? o.x1count * pr : o.y2count * pc) : n;
      std::vector<double> in(n), buf(std::max(nb, n) + 1), back(n, -1.0);
      for (size_t i = 0; i < n; ++i) in[i] = (double)i;
      oracle_mem_split(dir, in.data(), s[0], s[1], s[2], buf.data(), &o, 8, even);
      // at one rank per direction the split is its own inverse's input
      if ((xy ? pr : pc) == 1) {
        oracle_mem_merge(dir, buf.data(), s[0], s[1], s[2], back.data(), &o, 8, even);
        CHECK(back == in, "split/merge dir %d even %d", dir, even);
      }
    }
  }
}

// Glassman against a long-double DFT, every length 1..48 and both signs
static void check_glassman() {
  for (int n = 1; n <= 48; ++n)
    for (int isign = -1; isign <= 1; isign += 2) {
      std::vector<double> u(2 * n), w(2 * n), ref(2 * n);
      for (int i = 0; i < 2 * n; ++i) u[i] = std::sin(0.37 * i + n) + 0.1 * i;
      for (int k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        for (int j = 0; j < n; ++j) {
          const long double a = isign * 2.0L * 3.14159265358979323846264338327950288L * j * k / n;
          re += u[2 * j] * cosl(a) - u[2 * j + 1] * sinl(a);
          im += u[2 * j] * sinl(a) + u[2 * j + 1] * cosl(a);
        }
        ref[2 * k] = (double)re;
        ref[2 * k + 1] = (double)im;
      }
      oracle_spcfft(u.data(), n, isign, w.data());
      double err = 0, nrm = 0;
      for (int i = 0; i < 2 * n; ++i) {
        err += (u[i] - ref[i]) * (u[i] - ref[i]);
        nrm += ref[i] * ref[i];
      }
      CHECK(std::sqrt(err / nrm) < 1e-12, "glassman n=%d", n);
    }
}

// 3D drivers of the oracle on small grids: r2c then c2r returns N times the input
static void check_oracle_3d(int fmt, int nx, int ny, int nz) {
  const size_t n = (size_t)nx * ny * nz;
  const size_t ns = fmt == 1 ? (size_t)(nx / 2 + 1) * ny * nz : (size_t)nx * ny * (nz / 2 + 1);
  std::vector<double> x(n), back(n), spec(2 * ns), c(2 * n), c2(2 * n);
  for (size_t i = 0; i < n; ++i) x[i] = std::cos(0.11 * i) + 0.5;
  CHECK(oracle_fft3d_r2c(fmt, nx, ny, nz, x.data(), spec.data(), nullptr) == 0, "r2c status");
  CHECK(oracle_fft3d_c2r(fmt, nx, ny, nz, spec.data(), back.data(), nullptr) == 0, "c2r status");
  double e = 0;
  for (size_t i = 0; i < n; ++i) e = std::max(e, std::fabs(back[i] / (double)n - x[i]));
  CHECK(e < 1e-12, "round trip %d %d %d fmt %d: %g", nx, ny, nz, fmt, e);
  for (size_t i = 0; i < 2 * n; ++i) c[i] = std::sin(0.07 * i);
  CHECK(oracle_fft3d_c2c(fmt, nx, ny, nz, c.data(), c2.data(), -1, nullptr) == 0, "c2c status");
}

int main() {
  const int dims[] = {1, 2, 3, 5, 8, 13, 17, 32, 33};
  const int procs[] = {1, 2, 3, 4, 5, 8};
  for (int nx : dims)
    for (int ny : dims)
      for (int nz : dims)
        for (int pr : procs)
          for (int pc : procs)
            if (pr <= ny && pc <= nz) check_decomp(nx, ny, nz, pr, pc);
  check_decomp(2048, 2048, 2048, 4, 4);
  check_decomp(1024, 1024, 1024, 4, 2);
  for (int pr : procs)
    for (int pc : procs) check_pack_maps(17, 13, 11, pr, pc);
  for (int np = 1; np <= 512; ++np) {
    int pr = 0, pc = 0;
    CHECK(d2dhip_best_2d_grid(np, &pr, &pc) == 0 && pr * pc == np, "best_2d_grid %d", np);
  }
  for (int s = -12; s <= 128; ++s) CHECK(d2dhip_status_string(s) != nullptr, "status %d", s);
  d2dhip_profiler_prep(1, 1, 1, 1);
  CHECK(d2dhip_profiler_start("outer") == 0 && d2dhip_profiler_start("inner") == 0, "start");
  CHECK(d2dhip_profiler_end("inner") == 0 && d2dhip_profiler_end("outer") == 0, "end");
  CHECK(d2dhip_profiler_end("none") == D2DHIP_ERR_INVALID_VALUE, "unbalanced end");
  d2dhip_modes m;
  CHECK(d2dhip_modes_make(&m, 32, 128, 1, -6.0, 2.0, 1e-6, 3) == 0, "modes");
  CHECK(d2dhip_modes_make(&m, 65, 8, 0, 0, 0, 0, 1) != 0, "too many modes");
  std::vector<double> f(9 * 8 * 7);
  d2dhip_host_fill_modal(f.data(), 9, 8, 7, D2DHIP_R8, &m);
  std::vector<float> g(2 * 100);
  d2dhip_host_fill_uniform(g.data(), 100, D2DHIP_C8, 5);
  for (float v : g) CHECK(v >= -1.f && v < 1.f, "uniform range");
  check_glassman();
  for (int fmt = 1; fmt <= 3; fmt += 2) {
    check_oracle_3d(fmt, 17, 13, 11);
    check_oracle_3d(fmt, 8, 6, 5);
    check_oracle_3d(fmt, 1, 7, 2);
  }
  if (g_fail) return 1;
  std::printf("host checks passed\n");
  return 0;
}
