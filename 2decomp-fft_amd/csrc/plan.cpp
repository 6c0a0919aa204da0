// plan.cpp -- FFT plans and the one-rank 3D drivers of libd2dhip.
//
// A plan holds, per axis length, a Stockham radix schedule and its FP64- or
// FP32-rounded twiddles (generated in long double with exact argument
// reduction), plus the workspace a non-inplace c2r needs to preserve its
// input.  The 3D drivers follow the stage order of the reference's
// fft_3d_c2c / fft_3d_r2c / fft_3d_c2r (src/fft_common_3d.f90:9-296) at
// dims = (1,1), where every transpose is the identity: each stage reads its
// input pencil once and writes its output once, the first stage reading the
// caller's input and the later ones working in place on the output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace d2dhip;

struct d2dhip_plan_s {
  int format = 1;
  int nx = 0, ny = 0, nz = 0;
  int prec = 64;
  int rb = 8;  // bytes per real of the plan's storage
  unsigned flags = 0;
  hipStream_t stream = nullptr;
  AxisPlan ax[3];
  // Stage descriptors: 0-2 c2c along x, y, z of the (nx,ny,nz) pencil;
  // 3 the real-to-complex stage; 4, 5 the two c2c stages of the spectral
  // pencil (y then the remaining axis); 6 the complex-to-real stage.
  static constexpr int kStages = 7;
  AxisDesc h_desc[kStages];
  AxisShape shape[kStages];
  int stage_io[kStages];
  AxisDesc *d_desc = nullptr;
  void *wk = nullptr;  // c2r workspace (spectral pencil)
  size_t wk_bytes = 0;
  // stage timing: an event pair around every launch while enabled
  int prof = 0;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
};

namespace {

// (cos, sin) of 2*pi*num/den in long double, reduced exactly to [0, pi/4].
void unit_root(long long num, long long den, long double &c, long double &s) {
  static const long double kTwoPi = 6.283185307179586476925286766559005768L;
  num %= den;
  if (num < 0) num += den;
  int sgn_c = 1, sgn_s = 1;
  bool swap = false;
  // f = num/den in [0,1)
  if (2 * num >= den && num != 0) {  // f -> 1 - f
    num = den - num;
    sgn_s = -sgn_s;
  }
  if (4 * num > den) {  // f -> 1/2 - f  (as (den - 2num)/(2den))
    num = den - 2 * num;
    den = 2 * den;
    sgn_c = -sgn_c;
  }
  if (8 * num > den) {  // f -> 1/4 - f  (as (den - 4num)/(4den))
    num = den - 4 * num;
    den = 4 * den;
    swap = true;
  }
  const long double x = kTwoPi * (long double)num / (long double)den;
  long double cc = cosl(x), ss = sinl(x);
  if (swap) std::swap(cc, ss);
  c = sgn_c * cc;
  s = sgn_s * ss;
}

// Radix schedule of length n from {16, 8, 4, 2, 9, 3, 5, 7, 11, 13, 17};
// empty with ok = false if n has a prime factor above 17.
std::vector<int> factor_radices(int n, bool &ok) {
  std::vector<int> r;
  ok = true;
  if (n == 1) {
    r.push_back(1);
    return r;
  }
  int e = 0, m = n;
  while (m % 2 == 0) {
    m /= 2;
    ++e;
  }
  int q = e / 4, rem = e % 4;
  std::vector<int> p2(q, 16);
  if (rem == 1) {
    if (q >= 1) {
      p2.pop_back();
      p2.push_back(8);
      p2.push_back(4);
    } else {
      p2.push_back(2);
    }
  } else if (rem == 2) {
    if (q >= 1) {
      p2.pop_back();
      p2.push_back(8);
      p2.push_back(8);
    } else {
      p2.push_back(4);
    }
  } else if (rem == 3) {
    p2.push_back(8);
  }
  for (int v : p2) r.push_back(v);
  while (m % 9 == 0) {
    r.push_back(9);
    m /= 9;
  }
  const int odd[] = {3, 5, 7, 11, 13, 17};
  for (int p : odd) {
    while (m % p == 0) {
      r.push_back(p);
      m /= p;
    }
  }
  if (m != 1) ok = false;
  return r;
}

int build_axis(AxisPlan &ap, int n, int rb) {
  bool ok = true;
  ap.n = n;
  ap.radices = factor_radices(n, ok);
  if (!ok)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "FFT length " + std::to_string(n) +
                         " has a prime factor above 17 (Bluestein path not available)");
  ap.passes.clear();
  int L = 1, off = 0;
  std::vector<long double> twc, tws;
  for (int r : ap.radices) {
    PassInfo pi;
    pi.r = r;
    pi.L = L;
    pi.tw_off = off;
    pi.invL = 1.0f / (float)L;
    ap.passes.push_back(pi);
    for (int k = 0; k < L; ++k)
      for (int s = 1; s < r; ++s) {
        long double c, sn;
        unit_root((long long)s * k, (long long)L * r, c, sn);
        twc.push_back(c);
        tws.push_back(-sn);  // forward sign: exp(-2 pi i s k / (L r))
      }
    off += L * (r - 1);
    L *= r;
  }
  if ((int)ap.passes.size() > kMaxPass)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "too many FFT passes");
  const size_t cnt = std::max<size_t>(twc.size(), 1);
  ap.tw_bytes = cnt * 2 * rb;
  std::vector<unsigned char> host(ap.tw_bytes, 0);
  for (size_t i = 0; i < twc.size(); ++i) {
    if (rb == 8) {
      double *h = reinterpret_cast<double *>(host.data());
      h[2 * i] = (double)twc[i];
      h[2 * i + 1] = (double)tws[i];
    } else {
      float *h = reinterpret_cast<float *>(host.data());
      h[2 * i] = (float)twc[i];
      h[2 * i + 1] = (float)tws[i];
    }
  }
  int st = hip_check(hipMalloc(&ap.d_tw, ap.tw_bytes), "hipMalloc(twiddles)");
  if (st) {
    ap.d_tw = nullptr;
    return D2DHIP_ERR_ALLOC;
  }
  return hip_check(hipMemcpy(ap.d_tw, host.data(), ap.tw_bytes, hipMemcpyHostToDevice),
                   "hipMemcpy(twiddles)");
}

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Tile width, block size and LDS footprint of one stage.  Strided axes
// (lines = false) take TX columns of one 128-byte run per point; contiguous
// lines (lines = true) take TX whole lines.  LDS is capped at 64 KiB so that
// two workgroups fit a CU, unless a single column needs more.
int choose_shape(const AxisPlan &ap, bool lines, int cb, long long C0, AxisShape &s) {
  const int n = ap.n;
  int txmax = lines ? 64 : std::max(1, 128 / cb);
  long long c0p2 = 1;
  while (c0p2 < C0 && c0p2 < txmax) c0p2 <<= 1;
  txmax = (int)std::min<long long>(txmax, c0p2);
  const int ldt_lines = (n % 2 == 0) ? n + 1 : n;
  for (int pass = 0; pass < 2; ++pass) {
    const size_t cap = pass == 0 ? 64 * 1024 : 160 * 1024;
    for (int tx = txmax; tx >= 1; tx >>= 1) {
      int nt = 64;
      for (const PassInfo &p : ap.passes) {
        const int per = kRegCap / p.r;
        const long long need = ((long long)tx * (n / p.r) + per - 1) / per;
        nt = (int)std::max<long long>(nt, need);
      }
      nt = (nt + 63) / 64 * 64;
      const size_t lds = lines ? (size_t)tx * ldt_lines * cb : (size_t)tx * n * cb;
      if (nt <= 512 && lds <= cap) {
        s.TX = tx;
        s.nthreads = nt;
        s.lds_bytes = lds;
        s.ldj = lines ? 1 : tx;
        s.ldt = lines ? ldt_lines : 1;
        return D2DHIP_SUCCESS;
      }
    }
  }
  return set_error(D2DHIP_ERR_UNSUPPORTED,
                   "FFT length " + std::to_string(n) + " does not fit one workgroup's LDS");
}

struct StageIo {
  long long C0r;  // columns (real columns for paired IO)
  long long C1;
  long long in_s0, in_s1, in_sj;
  long long out_s0, out_s1, out_sj;
};

// Column views of a pencil (n1, n2, n3), i fastest, for a stage along
// `axis`: x runs along contiguous lines, y and z along strided columns.
StageIo cols_view(int axis, long long n1, long long n2, long long n3) {
  StageIo v{};
  if (axis == 0) {  // lines (j,k): start l*n1, unit point stride
    v.C0r = n2 * n3;
    v.C1 = 1;
    v.in_s0 = v.out_s0 = n1;
    v.in_s1 = v.out_s1 = 0;
    v.in_sj = v.out_sj = 1;
  } else if (axis == 1) {  // columns i within plane k, points n1 apart
    v.C0r = n1;
    v.C1 = n3;
    v.in_s0 = v.out_s0 = 1;
    v.in_s1 = v.out_s1 = n1 * n2;
    v.in_sj = v.out_sj = n1;
  } else {  // columns p = i + n1*j, points n1*n2 apart
    v.C0r = n1 * n2;
    v.C1 = 1;
    v.in_s0 = v.out_s0 = 1;
    v.in_s1 = v.out_s1 = 0;
    v.in_sj = v.out_sj = n1 * n2;
  }
  return v;
}

int make_stage(d2dhip_plan_s *p, int idx, int axis, int io, const StageIo &sio) {
  const AxisPlan &ap = p->ax[axis];
  const int cb = 2 * p->rb;
  AxisDesc &a = p->h_desc[idx];
  std::memset(&a, 0, sizeof(a));
  a.tw = ap.d_tw;
  a.n = ap.n;
  a.npass = (int)ap.passes.size();
  for (int i = 0; i < a.npass; ++i) a.pass[i] = ap.passes[i];
  a.C0r = sio.C0r;
  a.C0 = (io == IO_C2C) ? sio.C0r : (sio.C0r + 1) / 2;
  a.C1 = sio.C1;
  AxisShape &s = p->shape[idx];
  int st = choose_shape(ap, axis == 0, cb, a.C0, s);
  if (st) return st;
  a.TX = s.TX;
  a.log2TX = ilog2(s.TX);
  const long long tiles0 = (a.C0 + s.TX - 1) / s.TX;
  if (tiles0 > (1LL << 30)) return set_error(D2DHIP_ERR_UNSUPPORTED, "pencil too large");
  a.tiles0 = (int)tiles0;
  a.in_s0 = sio.in_s0;
  a.in_s1 = sio.in_s1;
  a.in_sj = sio.in_sj;
  a.out_s0 = sio.out_s0;
  a.out_s1 = sio.out_s1;
  a.out_sj = sio.out_sj;
  a.ldj = s.ldj;
  a.ldt = s.ldt;
  a.nh = ap.n / 2 + 1;
  p->stage_io[idx] = io;
  return D2DHIP_SUCCESS;
}

// Builds the seven stage descriptors of a plan and uploads them.
int make_stages(d2dhip_plan_s *p) {
  const long long nx = p->nx, ny = p->ny, nz = p->nz;
  int st;
  for (int a = 0; a < 3; ++a)
    if ((st = make_stage(p, a, a, IO_C2C, cols_view(a, nx, ny, nz)))) return st;
  if (p->format == D2DHIP_PHYSICAL_IN_X) {
    const long long nxc = nx / 2 + 1;
    StageIo r = cols_view(0, nx, ny, nz);  // real lines of nx -> complex lines of nxc
    r.out_s0 = nxc;
    if ((st = make_stage(p, 3, 0, IO_R2C, r))) return st;
    if ((st = make_stage(p, 4, 1, IO_C2C, cols_view(1, nxc, ny, nz)))) return st;
    if ((st = make_stage(p, 5, 2, IO_C2C, cols_view(2, nxc, ny, nz)))) return st;
    StageIo c = cols_view(0, nx, ny, nz);
    c.in_s0 = nxc;
    if ((st = make_stage(p, 6, 0, IO_C2R, c))) return st;
  } else {
    const long long nzc = nz / 2 + 1;
    // real columns i + nx*j of the z axis; the complex output has the same
    // column and point strides
    if ((st = make_stage(p, 3, 2, IO_R2C, cols_view(2, nx, ny, nz)))) return st;
    if ((st = make_stage(p, 4, 1, IO_C2C, cols_view(1, nx, ny, nzc)))) return st;
    if ((st = make_stage(p, 5, 0, IO_C2C, cols_view(0, nx, ny, nzc)))) return st;
    if ((st = make_stage(p, 6, 2, IO_C2R, cols_view(2, nx, ny, nz)))) return st;
  }
  const size_t bytes = sizeof(AxisDesc) * d2dhip_plan_s::kStages;
  if (hipMalloc(&p->d_desc, bytes) != hipSuccess) {
    p->d_desc = nullptr;
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(stage descriptors) failed");
  }
  return hip_check(hipMemcpy(p->d_desc, p->h_desc, bytes, hipMemcpyHostToDevice),
                   "hipMemcpy(stage descriptors)");
}

hipEvent_t take_event(d2dhip_plan_s *p) {
  if (!p->pool.empty()) {
    hipEvent_t e = p->pool.back();
    p->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int run(d2dhip_plan_s *p, int idx, const void *in, void *out, int conj) {
  const AxisIo io = {in, out, conj};
  hipEvent_t a = nullptr, b = nullptr;
  if (p->prof) {
    a = take_event(p);
    b = take_event(p);
    if (!a || !b) return set_error(D2DHIP_ERR_HIP, "hipEventCreate failed");
    int st = hip_check(hipEventRecord(a, p->stream), "hipEventRecord");
    if (st) return st;
  }
  int st = launch_axis(p->rb, p->stage_io[idx], p->d_desc + idx, p->h_desc[idx], io,
                       p->shape[idx], p->stream);
  if (p->prof) {
    if (!st) st = hip_check(hipEventRecord(b, p->stream), "hipEventRecord");
    p->recs.push_back({idx, a, b});
  }
  return st;
}

}  // namespace

extern "C" {

int d2dhip_plan_create(d2dhip_plan *plan, int format, int nx, int ny, int nz, int prec,
                       unsigned flags, void *stream) {
  if (!plan) return set_error(D2DHIP_ERR_NULL, "plan pointer is NULL");
  *plan = nullptr;
  if (nx <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for nx");
  if (ny <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for ny");
  if (nz <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for nz");
  if (format != D2DHIP_PHYSICAL_IN_X && format != D2DHIP_PHYSICAL_IN_Z)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for pencil");
  if (prec != D2DHIP_PREC_F64 && prec != D2DHIP_PREC_F32 && prec != D2DHIP_PREC_SPLIT_BF16)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid precision");
  if (prec == D2DHIP_PREC_SPLIT_BF16)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "split-BF16 plans are not available yet");
  d2dhip_plan_s *p = new (std::nothrow) d2dhip_plan_s();
  if (!p) return set_error(D2DHIP_ERR_ALLOC, "host allocation failed");
  p->format = format;
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->prec = prec;
  p->rb = (prec == D2DHIP_PREC_F64) ? 8 : 4;
  p->flags = flags;
  p->stream = (hipStream_t)stream;
  int st = axis_kernel_prepare(p->rb);
  const int ns[3] = {nx, ny, nz};
  for (int a = 0; a < 3 && !st; ++a) st = build_axis(p->ax[a], ns[a], p->rb);
  if (!st) st = make_stages(p);
  if (!st && !(flags & D2DHIP_FLAG_INPLACE)) {
    const long long nsp = (format == D2DHIP_PHYSICAL_IN_X)
                              ? (long long)(nx / 2 + 1) * ny * nz
                              : (long long)nx * ny * (nz / 2 + 1);
    p->wk_bytes = (size_t)nsp * 2 * p->rb;
    if (hipMalloc(&p->wk, p->wk_bytes) != hipSuccess) {
      p->wk = nullptr;
      st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(c2r workspace) failed");
    }
  }
  if (st) {
    d2dhip_plan_destroy(p);
    return st;
  }
  *plan = p;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_destroy(d2dhip_plan p) {
  if (!p) return D2DHIP_SUCCESS;
  for (auto &r : p->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : p->pool) (void)hipEventDestroy(e);
  for (auto &a : p->ax)
    if (a.d_tw) (void)hipFree(a.d_tw);
  if (p->wk) (void)hipFree(p->wk);
  if (p->d_desc) (void)hipFree(p->d_desc);
  delete p;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_set_stream(d2dhip_plan p, void *stream) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  p->stream = (hipStream_t)stream;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_get_sizes(d2dhip_plan p, int ph_sz[3], int sp_sz[3]) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (ph_sz) {
    ph_sz[0] = p->nx;
    ph_sz[1] = p->ny;
    ph_sz[2] = p->nz;
  }
  if (sp_sz) {
    sp_sz[0] = p->format == D2DHIP_PHYSICAL_IN_X ? p->nx / 2 + 1 : p->nx;
    sp_sz[1] = p->ny;
    sp_sz[2] = p->format == D2DHIP_PHYSICAL_IN_Z ? p->nz / 2 + 1 : p->nz;
  }
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_profile(d2dhip_plan p, int enable) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  p->prof = enable ? 1 : 0;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_profile_report(d2dhip_plan p, double ms[7], int launches[7]) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  for (int i = 0; i < d2dhip_plan_s::kStages; ++i) {
    if (ms) ms[i] = 0.0;
    if (launches) launches[i] = 0;
  }
  int st = D2DHIP_SUCCESS;
  for (auto &r : p->recs) {
    if (!st) st = hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
    float t = 0.f;
    if (!st) st = hip_check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
    if (!st) {
      if (ms) ms[r.stage] += t;
      if (launches) launches[r.stage] += 1;
    }
    p->pool.push_back(r.a);
    p->pool.push_back(r.b);
  }
  p->recs.clear();
  return st;
}

long long d2dhip_plan_device_bytes(d2dhip_plan p) {
  if (!p) return 0;
  long long b = (long long)p->wk_bytes;
  for (auto &a : p->ax) b += (long long)a.tw_bytes;
  return b;
}

// fft_3d_c2c: src/fft_common_3d.f90:9-116.  X-forward and Z-backward run
// x, y, z; X-backward and Z-forward run z, y, x.  Skipped axes
// (opt_skip_XYZ_c2c) are left out; the first stage that runs reads `in`.
int d2dhip_exec_c2c(d2dhip_plan p, void *in, void *out, int isign) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (!in || !out) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  if (isign != D2DHIP_FORWARD && isign != D2DHIP_BACKWARD)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "isign must be -1 or 1");
  const bool x_first = (p->format == D2DHIP_PHYSICAL_IN_X && isign == D2DHIP_FORWARD) ||
                       (p->format == D2DHIP_PHYSICAL_IN_Z && isign == D2DHIP_BACKWARD);
  const int order[3] = {x_first ? 0 : 2, 1, x_first ? 2 : 0};
  const unsigned skipbit[3] = {D2DHIP_FLAG_SKIP_X, D2DHIP_FLAG_SKIP_Y, D2DHIP_FLAG_SKIP_Z};
  const int conj = (isign == D2DHIP_BACKWARD) ? 1 : 0;
  const void *src = in;
  int ran = 0;
  for (int s = 0; s < 3; ++s) {
    const int ax = order[s];
    if (p->flags & skipbit[ax]) continue;
    int st = run(p, ax, src, out, conj);
    if (st) return st;
    src = out;
    ++ran;
  }
  if (!ran && in != out) {
    const size_t bytes = (size_t)p->nx * p->ny * p->nz * 2 * p->rb;
    return hip_check(hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, p->stream),
                     "hipMemcpyAsync");
  }
  return D2DHIP_SUCCESS;
}

// fft_3d_r2c: src/fft_common_3d.f90:121-189.  PHYSICAL_IN_X: r2c along x,
// then y, then z; PHYSICAL_IN_Z: r2c along z, then y, then x.
int d2dhip_exec_r2c(d2dhip_plan p, const void *in_r, void *out_c) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (!in_r || !out_c) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  const bool fx = p->format == D2DHIP_PHYSICAL_IN_X;
  // the r2c axis cannot be skipped (src/fft_generic.f90:222-223, :265-266)
  const int warn = (p->flags & (fx ? D2DHIP_FLAG_SKIP_X : D2DHIP_FLAG_SKIP_Z)) ? 1 : 0;
  const unsigned skip_last = fx ? D2DHIP_FLAG_SKIP_Z : D2DHIP_FLAG_SKIP_X;
  int st;
  if ((st = run(p, 3, in_r, out_c, 0))) return st;
  if (!(p->flags & D2DHIP_FLAG_SKIP_Y) && (st = run(p, 4, out_c, out_c, 0))) return st;
  if (!(p->flags & skip_last) && (st = run(p, 5, out_c, out_c, 0))) return st;
  if (warn)
    return set_error(D2DHIP_WARN_SKIP_IGNORED, "r2c / c2r transform can not be skipped");
  return D2DHIP_SUCCESS;
}

// fft_3d_c2r: src/fft_common_3d.f90:194-296.  The c2c stages write the
// plan's workspace (or in_c itself with D2DHIP_FLAG_INPLACE), so that in_c
// is preserved as the reference's wk1 copy preserves it, without the copy.
int d2dhip_exec_c2r(d2dhip_plan p, void *in_c, void *out_r) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (!in_c || !out_r) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  const bool fx = p->format == D2DHIP_PHYSICAL_IN_X;
  const int warn = (p->flags & (fx ? D2DHIP_FLAG_SKIP_X : D2DHIP_FLAG_SKIP_Z)) ? 1 : 0;
  const unsigned skip_first = fx ? D2DHIP_FLAG_SKIP_Z : D2DHIP_FLAG_SKIP_X;
  void *wk = (p->flags & D2DHIP_FLAG_INPLACE) ? in_c : p->wk;
  const void *src = in_c;
  int st;
  if (!(p->flags & skip_first)) {  // z (X format) or x (Z format), inverse
    if ((st = run(p, 5, src, wk, 1))) return st;
    src = wk;
  }
  if (!(p->flags & D2DHIP_FLAG_SKIP_Y)) {
    if ((st = run(p, 4, src, wk, 1))) return st;
    src = wk;
  }
  if ((st = run(p, 6, src, out_r, 1))) return st;
  if (warn)
    return set_error(D2DHIP_WARN_SKIP_IGNORED, "r2c / c2r transform can not be skipped");
  return D2DHIP_SUCCESS;
}

}  // extern "C"
