// api.cpp -- status handling, decomposition metadata, transposes/packs and
// synthetic-input generators of libd2dhip's C ABI.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <string>

#include "gen.h"
#include "internal.h"

namespace d2dhip {

static thread_local std::string g_last_message;

// profiler categories that open regions; all on by default, as the
// reference's default_profiler = .true. (src/profiler_caliper.f90:23-27)
static std::atomic<unsigned> g_regions{PROF_TRANSPOSE | PROF_IO | PROF_FFT | PROF_D2D};

Region::Region(int category, const char *name)
    : on((g_regions.load(std::memory_order_relaxed) & (unsigned)category) != 0) {
  if (on) roctxRangePushA(name);
}
Region::~Region() {
  if (on) roctxRangePop();
}

int set_error(int status, const std::string &msg) {
  g_last_message = msg;
  return status;
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return D2DHIP_SUCCESS;
  (void)hipGetLastError();  // do not let this failure resurface at the next launch
  return set_error(e == hipErrorOutOfMemory ? D2DHIP_ERR_ALLOC : D2DHIP_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace d2dhip

using namespace d2dhip;

namespace {

// distribute: src/decomp_2d.f90:1070-1101.  Floor-sized blocks first, the
// data1 mod proc one-larger blocks last.  st/en are 1-based.
int distribute(int data1, int proc, int *st, int *en, int *sz) {
  int size1 = data1 / proc;
  const int nu = data1 - size1 * proc;
  const int nl = proc - nu;
  st[0] = 1;
  sz[0] = size1;
  en[0] = size1;
  for (int i = 1; i < nl; ++i) {
    st[i] = st[i - 1] + size1;
    sz[i] = size1;
    en[i] = en[i - 1] + size1;
  }
  size1 += 1;
  for (int i = nl; i < proc; ++i) {
    st[i] = en[i - 1] + 1;
    sz[i] = size1;
    en[i] = en[i - 1] + size1;
  }
  if (en[proc - 1] != data1 || sz[proc - 1] != data1 - st[proc - 1] + 1)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid distribution.");
  return D2DHIP_SUCCESS;
}

// partition: src/decomp_2d.f90:1016-1064.  pdim 1 = local, 2 = over p_row
// at coord_row, 3 = over p_col at coord_col.
int partition(const int g[3], const int pdim[3], const d2dhip_decomp *d, int st[3], int en[3],
              int sz[3]) {
  int s[D2DHIP_MAX_PROCS], e[D2DHIP_MAX_PROCS], z[D2DHIP_MAX_PROCS];
  for (int i = 0; i < 3; ++i) {
    if (pdim[i] == 1) {
      st[i] = 1;
      en[i] = g[i];
      sz[i] = g[i];
      continue;
    }
    const int np = pdim[i] == 2 ? d->p_row : d->p_col;
    const int c = pdim[i] == 2 ? d->coord_row : d->coord_col;
    int rc = distribute(g[i], np, s, e, z);
    if (rc) return rc;
    st[i] = s[c];
    en[i] = e[c];
    sz[i] = z[c];
  }
  return D2DHIP_SUCCESS;
}

int elem_size(int elem) {
  switch (elem) {
    case D2DHIP_R8: return 8;
    case D2DHIP_C16: return 16;
    case D2DHIP_R4: return 4;
    case D2DHIP_C8: return 8;
    default: return 0;
  }
}

}  // namespace

extern "C" {

const char *d2dhip_status_string(int status) {
  switch (status) {
    case D2DHIP_SUCCESS: return "success";
    case D2DHIP_ERR_INVALID_VALUE: return "invalid value";
    case D2DHIP_ERR_UNSUPPORTED: return "unsupported";
    case D2DHIP_ERR_ALLOC: return "device allocation failed";
    case D2DHIP_ERR_HIP: return "HIP runtime error";
    case D2DHIP_ERR_NULL: return "NULL plan or pointer";
    case D2DHIP_ERR_GRID: return "Invalid 2D processor grid";
    case D2DHIP_ERR_PLAN_MISMATCH: return "call does not match the plan";
    case D2DHIP_WARN_SKIP_IGNORED: return "r2c / c2r transform can not be skipped";
    case D2DHIP_WARN_TUNE_CACHE: return "tuning choices could not be written to the cache";
    default: return "unknown status";
  }
}

const char *d2dhip_last_message(void) { return g_last_message.c_str(); }

int d2dhip_version(void) { return 100; }

// decomp_profiler_prep (src/profiler_caliper.f90:138-150): which categories
// open regions, in the reference's order (transpose, io, fft, d2d).
int d2dhip_profiler_prep(int transpose, int io, int fft, int d2d) {
  g_regions.store((transpose ? PROF_TRANSPOSE : 0u) | (io ? PROF_IO : 0u) | (fft ? PROF_FFT : 0u) |
                      (d2d ? PROF_D2D : 0u),
                  std::memory_order_relaxed);
  return D2DHIP_SUCCESS;
}

// decomp_profiler_start / _end (src/profiler_caliper.f90:156-177): a caller's
// own region, always opened; regions nest, and _end closes the innermost one
// that is open on the calling thread.
static thread_local int g_open_regions = 0;  // opened by d2dhip_profiler_start
int d2dhip_profiler_start(const char *name) {
  if (!name) return set_error(D2DHIP_ERR_NULL, "region name is NULL");
  roctxRangePushA(name);
  ++g_open_regions;
  return D2DHIP_SUCCESS;
}
int d2dhip_profiler_end(const char *name) {
  if (!name) return set_error(D2DHIP_ERR_NULL, "region name is NULL");
  if (g_open_regions == 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "no region is open");
  --g_open_regions;
  roctxRangePop();
  return D2DHIP_SUCCESS;
}

// decomp_info_init: src/decomp_2d.f90:382-490 with get_dist (:1108-1130),
// partition x (1,2,3), y (2,1,3), z (2,3,1) (:413-418), prepare_buffer
// (:1135-1206) and the even test (:443-454).
int d2dhip_decomp_info_init(int nx, int ny, int nz, int p_row, int p_col, int coord_row,
                            int coord_col, d2dhip_decomp *d) {
  Region region(PROF_D2D, "decomp_2d_init");
  if (!d) return set_error(D2DHIP_ERR_NULL, "decomp pointer is NULL");
  std::memset(d, 0, sizeof(*d));
  if (nx <= 0 || ny <= 0 || nz <= 0)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid grid size");
  if (p_row < 1 || p_col < 1 || p_row > D2DHIP_MAX_PROCS || p_col > D2DHIP_MAX_PROCS)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "p_row and p_col must be in [1, 64]");
  if (coord_row < 0 || coord_row >= p_row || coord_col < 0 || coord_col >= p_col)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "rank coordinates out of range");
  if (nx < p_row || ny < p_row || ny < p_col || nz < p_col)
    return set_error(D2DHIP_ERR_GRID,
                     "Invalid 2D processor grid. Make sure that min(nx,ny) >= p_row and "
                     "min(ny,nz) >= p_col");
  d->p_row = p_row;
  d->p_col = p_col;
  d->coord_row = coord_row;
  d->coord_col = coord_col;
  int s[D2DHIP_MAX_PROCS], e[D2DHIP_MAX_PROCS];
  int rc;
  if ((rc = distribute(nx, p_row, s, e, d->x1dist))) return rc;
  if ((rc = distribute(ny, p_row, s, e, d->y1dist))) return rc;
  if ((rc = distribute(ny, p_col, s, e, d->y2dist))) return rc;
  if ((rc = distribute(nz, p_col, s, e, d->z2dist))) return rc;
  const int g[3] = {nx, ny, nz};
  const int px[3] = {1, 2, 3}, py[3] = {2, 1, 3}, pz[3] = {2, 3, 1};
  if ((rc = partition(g, px, d, d->xst, d->xen, d->xsz))) return rc;
  if ((rc = partition(g, py, d, d->yst, d->yen, d->ysz))) return rc;
  if ((rc = partition(g, pz, d, d->zst, d->zen, d->zsz))) return rc;
  for (int i = 0; i < p_row; ++i) {
    d->x1cnts[i] = (long long)d->x1dist[i] * d->xsz[1] * d->xsz[2];
    d->y1cnts[i] = (long long)d->ysz[0] * d->y1dist[i] * d->ysz[2];
    d->x1disp[i] = i ? d->x1disp[i - 1] + d->x1cnts[i - 1] : 0;
    d->y1disp[i] = i ? d->y1disp[i - 1] + d->y1cnts[i - 1] : 0;
  }
  for (int i = 0; i < p_col; ++i) {
    d->y2cnts[i] = (long long)d->ysz[0] * d->y2dist[i] * d->ysz[2];
    d->z2cnts[i] = (long long)d->zsz[0] * d->zsz[1] * d->z2dist[i];
    d->y2disp[i] = i ? d->y2disp[i - 1] + d->y2cnts[i - 1] : 0;
    d->z2disp[i] = i ? d->z2disp[i - 1] + d->z2cnts[i - 1] : 0;
  }
  d->x1count = (long long)d->x1dist[p_row - 1] * d->y1dist[p_row - 1] * d->xsz[2];
  d->y1count = d->x1count;
  d->y2count = (long long)d->y2dist[p_col - 1] * d->z2dist[p_col - 1] * d->zsz[0];
  d->z2count = d->y2count;
  d->even = (nx % p_row == 0 && ny % p_row == 0 && ny % p_col == 0 && nz % p_col == 0);
  return D2DHIP_SUCCESS;
}

// best_2d_grid: src/decomp_2d_init_fin.f90:270-300 with findfactor
// (src/factor.f90:15-54): all factors of nproc in increasing order, p_col
// the factor at position nfact/2+1 (1-based), p_row = nproc / p_col.
int d2dhip_best_2d_grid(int nproc, int *p_row, int *p_col) {
  if (nproc < 1 || !p_row || !p_col) return set_error(D2DHIP_ERR_INVALID_VALUE, "nproc < 1");
  int f[512];
  int nf = 0;
  const int m = (int)std::sqrt((double)nproc);
  for (int i = 1; i <= m; ++i)
    if (nproc % i == 0) f[nf++] = i;
  if (f[nf - 1] * f[nf - 1] != nproc) {
    for (int i = nf; i < 2 * nf; ++i) f[i] = nproc / f[2 * nf - 1 - i];
    nf *= 2;
  } else {
    for (int i = nf; i < 2 * nf - 1; ++i) f[i] = nproc / f[2 * nf - 2 - i];
    nf = 2 * nf - 1;
  }
  *p_col = f[nf / 2];
  *p_row = nproc / *p_col;
  return D2DHIP_SUCCESS;
}

int d2dhip_transpose(int dir, int elem, const void *src, void *dst, const d2dhip_decomp *d,
                     void *stream) {
  if (!d || !src || !dst) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  const int es = elem_size(elem);
  if (!es || dir < 0 || dir > 3) return set_error(D2DHIP_ERR_INVALID_VALUE, "bad dir/elem");
  // region names of src/transpose_*.f90 (transp_x_y_r at transpose_x_to_y.f90:39, ...):
  // _r for real data, _c for complex
  static const char *const kTransp[4][2] = {{"transp_x_y_r", "transp_x_y_c"},
                                            {"transp_y_z_r", "transp_y_z_c"},
                                            {"transp_z_y_r", "transp_z_y_c"},
                                            {"transp_y_x_r", "transp_y_x_c"}};
  const bool cplx = elem == D2DHIP_C16 || elem == D2DHIP_C8;
  Region region(PROF_TRANSPOSE, kTransp[dir][cplx ? 1 : 0]);
  const bool xy = (dir == D2DHIP_X_TO_Y || dir == D2DHIP_Y_TO_X);
  const int np = xy ? d->p_row : d->p_col;
  if (np != 1)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "multi-rank transposes need the alltoallv exchange; use pack/unpack");
  // dims == 1: dst = src over the source pencil (src/transpose_x_to_y.f90:41-50).
  const int *sz = (dir == D2DHIP_X_TO_Y) ? d->xsz : (dir == D2DHIP_Z_TO_Y) ? d->zsz : d->ysz;
  const size_t bytes = (size_t)sz[0] * sz[1] * sz[2] * es;
  if (src == dst || bytes == 0) return D2DHIP_SUCCESS;
  return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream),
                   "hipMemcpyAsync(transpose)");
}

long long d2dhip_pack_buffer_elems(int dir, const d2dhip_decomp *d, int even) {
  if (!d || dir < 0 || dir > 3) return -1;
  const bool xy = (dir == D2DHIP_X_TO_Y || dir == D2DHIP_Y_TO_X);
  if (even) return xy ? d->x1count * d->p_row : d->y2count * d->p_col;
  const int *sz = (dir == D2DHIP_X_TO_Y) ? d->xsz : (dir == D2DHIP_Z_TO_Y) ? d->zsz : d->ysz;
  return (long long)sz[0] * sz[1] * sz[2];
}

int d2dhip_pack(int dir, int elem, const void *src, void *send, const d2dhip_decomp *d, int even,
                void *stream) {
  if (!d || !src || !send) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  const int es = elem_size(elem);
  if (!es || dir < 0 || dir > 3) return set_error(D2DHIP_ERR_INVALID_VALUE, "bad dir/elem");
  return launch_pack(dir, 1, es, src, send, d, even, (hipStream_t)stream);
}

int d2dhip_unpack(int dir, int elem, const void *recv, void *dst, const d2dhip_decomp *d,
                  int even, void *stream) {
  if (!d || !recv || !dst) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  const int es = elem_size(elem);
  if (!es || dir < 0 || dir > 3) return set_error(D2DHIP_ERR_INVALID_VALUE, "bad dir/elem");
  return launch_pack(dir, 0, es, recv, dst, d, even, (hipStream_t)stream);
}

// ---- halo cells ------------------------------------------------------------

static int halo_check(int elem, int s1, int s2, int s3, int level, int ipencil,
                      const int *periodic) {
  if (!periodic) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (!elem_size(elem)) return set_error(D2DHIP_ERR_INVALID_VALUE, "bad elem");
  if (s1 < 0 || s2 < 0 || s3 < 0 || level < 0)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "negative size or level");
  if (ipencil < 1 || ipencil > 3)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid decomposition size");
  const int s[3] = {s1, s2, s3};
  for (int a = 0; a < 3; ++a)
    if (a != ipencil - 1 && periodic[a] && level > s[a])
      return set_error(D2DHIP_ERR_INVALID_VALUE,
                       "halo level exceeds the pencil along a periodic axis");
  return D2DHIP_SUCCESS;
}

int d2dhip_update_halo(int elem, const void *in, void *out, int s1, int s2, int s3, int level,
                       int ipencil, const int periodic[3], void *stream) {
  if (!in || !out) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (int st = halo_check(elem, s1, s2, s3, level, ipencil, periodic)) return st;
  return launch_halo(elem_size(elem), in, out, s1, s2, s3, level, ipencil, periodic,
                     (hipStream_t)stream);
}

int d2dhip_halo_exchange(int elem, void *arr, int s1, int s2, int s3, int level, int ipencil,
                         const int periodic[3], void *stream) {
  if (!arr) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (int st = halo_check(elem, s1, s2, s3, level, ipencil, periodic)) return st;
  return launch_halo(elem_size(elem), nullptr, arr, s1, s2, s3, level, ipencil, periodic,
                     (hipStream_t)stream);
}

// ---- synthetic inputs ------------------------------------------------------

int d2dhip_fill_uniform(void *dst, long long n, int elem, unsigned long long seed, void *stream) {
  if (!dst) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (!elem_size(elem) || n < 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "bad elem/n");
  return launch_fill_uniform(dst, n, elem, seed, (hipStream_t)stream);
}

int d2dhip_column_copy(const void *src, void *dst, int elem, long long C0, long long C1,
                       long long s1, long long sj, int n, void *stream) {
  if (!src || !dst) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if ((elem != D2DHIP_C16 && elem != D2DHIP_C8) || C0 < 0 || C1 < 0 || n < 0 || s1 < 0 || sj < 0)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "bad column-copy arguments");
  return launch_column_copy(elem, src, dst, C0, C1, s1, sj, n, (hipStream_t)stream);
}

void d2dhip_host_fill_uniform(void *dst, long long n, int elem, unsigned long long seed) {
  const bool cplx = (elem == D2DHIP_C16 || elem == D2DHIP_C8);
  const long long nv = cplx ? 2 * n : n;
  if (elem == D2DHIP_R8 || elem == D2DHIP_C16) {
    double *p = (double *)dst;
    for (long long i = 0; i < nv; ++i) p[i] = gen_uniform_pm1((unsigned long long)seed, (unsigned long long)i);
  } else {
    float *p = (float *)dst;
    for (long long i = 0; i < nv; ++i) p[i] = (float)gen_uniform_pm1((unsigned long long)seed, (unsigned long long)i);
  }
}

int d2dhip_modes_make(d2dhip_modes *m, int nmodes, int kmax, int amp_kind, double amp_lo,
                      double amp_hi, double sigma, unsigned long long seed) {
  if (!m) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (nmodes < 0 || nmodes > 64 || kmax < 0)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "nmodes must be in [0,64], kmax >= 0");
  std::memset(m, 0, sizeof(*m));
  m->nmodes = nmodes;
  m->sigma = sigma;
  const unsigned long long ks = gen_stream_key(seed, 1);
  m->noise_seed = gen_stream_key(seed, 2);
  for (int i = 0; i < nmodes; ++i) {
    const unsigned long long b = 5ull * (unsigned long long)i;
    const double u0 = gen_u01(ks, b), u1 = gen_u01(ks, b + 1), u2 = gen_u01(ks, b + 2);
    const double u3 = gen_u01(ks, b + 3), u4 = gen_u01(ks, b + 4);
    const int span = 2 * kmax + 1;
    m->kx[i] = (int)(u0 * span) - kmax;
    m->ky[i] = (int)(u1 * span) - kmax;
    m->kz[i] = (int)(u2 * span) - kmax;
    if (amp_kind == 0) {
      m->amp[i] = gen_box_muller(u3, u4);
    } else {
      const double mag = std::pow(10.0, amp_lo + (amp_hi - amp_lo) * u3);
      m->amp[i] = (u4 < 0.5) ? -mag : mag;
    }
  }
  return D2DHIP_SUCCESS;
}

int d2dhip_fill_modal(void *dst, int nx, int ny, int nz, int elem, const d2dhip_modes *m,
                      void *stream) {
  if (!dst || !m) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (elem != D2DHIP_R8 && elem != D2DHIP_R4)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "modal fields are real (R8 or R4)");
  if (nx <= 0 || ny <= 0 || nz <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "bad grid");
  return launch_fill_modal(dst, nx, ny, nz, elem, m, (hipStream_t)stream);
}

void d2dhip_host_fill_modal(void *dst, int nx, int ny, int nz, int elem, const d2dhip_modes *m) {
  const long long N = (long long)nx * ny * nz;
  for (long long k = 0; k < nz; ++k)
    for (long long j = 0; j < ny; ++j)
      for (long long i = 0; i < nx; ++i) {
        const long long idx = i + nx * (j + ny * k);
        const double v = gen_modal_value(m, nx, ny, nz, i, j, k, idx);
        if (elem == D2DHIP_R8)
          ((double *)dst)[idx] = v;
        else
          ((float *)dst)[idx] = (float)v;
      }
  (void)N;
}

}  // extern "C"
