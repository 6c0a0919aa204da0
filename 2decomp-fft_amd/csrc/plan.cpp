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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "fast_kernels.h"
#include "internal.h"

using namespace d2dhip;

// Twiddles of one radix schedule (per-pass tables, forward sign).
struct TwTable {
  void *d = nullptr;
  size_t bytes = 0;
  std::vector<PassInfo> passes;
};

// Geometry of one stage and the kernels that can run it.
struct StageGeom {
  int axis = 0, io = 0, n = 0;
  long long C0r = 0, C1 = 0;
  long long in_s0 = 0, in_s1 = 0, in_sj = 0, out_s0 = 0, out_s1 = 0, out_sj = 0;
  long long in_elems = 0, out_elems = 0;  // elements (real or complex) of each side
  std::vector<const FastVariant *> cands;  // specialised kernels for this stage
  int choice = -1;                         // index into cands, -1 = general kernel
  // real-to-real stages (io IO_R2R): type 1-8, first transformed point,
  // transformed points, points per line; n is the extension's length
  int dtt = 0, dtt_first = 0, dtt_len = 0, dtt_n = 0;
  const void *dtw = nullptr;  // half-length types: (cos, sin)(pi k / (2 dtt_len))
  const void *dtw4 = nullptr;  // DCT-IV / DST-IV on L/2-point tiles: pre- and post-twiddles
};

// Chirp tables of a Bluestein length.
struct BlueTables {
  int M = 0;
  void *chirp = nullptr;
  void *bhat = nullptr;
};

struct d2dhip_plan_s {
  int format = 1;
  int nx = 0, ny = 0, nz = 0;
  int prec = 64;
  int rb = 8;  // bytes per real of the plan's storage
  unsigned flags = 0;
  hipStream_t stream = nullptr;
  std::map<std::vector<int>, TwTable> tw;
  std::map<int, BlueTables> blue;
  std::map<int, void *> dtw;  // DCT / DST twiddle tables by length
  // Stage descriptors: 0-2 c2c along x, y, z of the (nx,ny,nz) pencil;
  // 3 the real-to-complex stage; 4 the spectral y stage; 5 the other
  // spectral c2c stage (z for PHYSICAL_IN_X, x for _Z); 6 the c2r stage;
  // 7-9 the forward real-to-real (DCT / DST) stages along x, y, z and 10-12
  // the backward ones, once d2dhip_plan_set_dtt has set them.
  static constexpr int kFftStages = 7;
  static constexpr int kStages = 13;
  bool has_dtt = false;
  int dtt[15] = {};  // types 1-3, ifirst 4-6, ndismiss 7-9, ofirst 10-12, backward types 13-15
  StageGeom geom[kStages];
  AxisDesc h_desc[kStages];
  AxisShape shape[kStages];
  const FastVariant *fast[kStages] = {};
  AxisDesc *d_desc = nullptr;
  // Plane chunking (PHYSICAL_IN_X): the forward r2c x stage and y stage, and
  // the inverse y stage and c2r x stage, alternate over chunks of `chunk`
  // z-planes, so that the second stage of each pair reads a chunk the first
  // has just written while it still sits in the memory-side cache (256 MiB
  // of Infinity Cache on MI355X).  0 = whole stages.  Chunk descriptors:
  // slot 2*q + t, q = 0 stage 3 forward, 1 stage 4 forward, 2 stage 4
  // inverse, 3 stage 6 inverse; t = 0 full chunk, 1 last (shorter) chunk.
  int chunk_fwd = 0, chunk_bwd = 0;
  AxisDesc h_cdesc[8];
  AxisDesc *d_cdesc = nullptr;
  // Blocked spectral layout (D2DHIP_LAYOUT_BLOCKED / _BLOCKED32; PHYSICAL_IN_X
  // whole-grid plans with a workspace): between the x stage and the last
  // stage of exec_r2c / exec_c2r the spectral pencil sits in the workspace in
  // blocks of W = 16 or 32 i-columns, element (i, j, k) at
  //   (i / W) * bs + k * W * nyb + j * W + (i % W),
  //   nyb = ny + 1, bs = W * nyb * nz,
  // so that a z-stage tile reads and writes rows lying W * nyb elements
  // apart, several to an address-translation page (the plain z stage's rows
  // lie a whole i-j plane apart, one page each, and keep the translation
  // cache busy), and a y-stage tile is one contiguous run.  An x-stage tile
  // then touches one page per block, so the wider blocks halve its pages.
  // The spare j row keeps the row distance off a power of two, which HBM
  // serves more slowly.  Forward: x (into the blocks), z (in place), y
  // (blocks -> the plain output); inverse: y, z, x.  Descriptors: 0 x r2c,
  // 1 z over the whole blocks, 2 z over the last partial block, 3 y forward,
  // 4 y inverse, 5 x c2r.
  int layout = D2DHIP_LAYOUT_PLAIN;
  AxisDesc h_bdesc[6];
  AxisDesc *d_bdesc = nullptr;
  void *wk = nullptr;  // c2r workspace (spectral pencil)
  size_t wk_bytes = 0;
  // stage timing: an event pair around every launch while enabled
  int prof = 0;
  struct Rec {
    int stage;
    hipEvent_t a, b;
    double w;  // fraction of the stage's work (chunks of a chunked stage)
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  // fused pack / unpack: the x stage's descriptor with a block table of the
  // process grid (d2dhip_exec_r2c_pack, d2dhip_exec_c2r_unpack)
  long long *d_pack = nullptr;  // per-point table, pack_cap entries
  size_t pack_cap = 0;
  AxisDesc *d_pdesc = nullptr;
  std::vector<long long> h_pack;
  AxisDesc h_pdesc{};
  // hipGraph of the exec calls made between capture_begin and capture_end
  hipStream_t cap_stream = nullptr, cap_keep = nullptr;
  bool capturing = false;
  bool tune_write_failed = false;  // a tuning choice could not be written to the cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // Pencils the stages run on: x, y, z of the physical decomposition, then
  // x, y, z of the spectral one (sizes, i fastest).  At one rank every one
  // is the whole grid (or its spectral shape); a plan of one rank of a
  // p_row x p_col grid (d2dhip_plan_create_rank) has that rank's pencils,
  // and its 3D calls are refused (they need the exchange between stages).
  long long pen[6][3] = {};
  bool multi = false;
  d2dhip_decomp ph{}, sp{};
  // y and z stages fused with their exchanges' pack / unpack
  // (d2dhip_exec_1m_pack / _unpack / _unpack_pack): block tables and
  // descriptor; and, for a stage kernel without a fused form, a pencil
  // workspace for unpack, stage, pack (allocated on first use)
  long long *d_ypack = nullptr, *d_ypack2 = nullptr;  // per-point tables (input / output side)
  size_t ypack_cap = 0, ypack2_cap = 0;                // their sizes in entries
  AxisDesc *d_ydesc = nullptr;
  std::vector<long long> h_ypack, h_ypack2;
  AxisDesc h_ydesc{};
  void *ywk = nullptr;
  size_t ywk_bytes = 0;
};

namespace {

int mkdir_one(const std::string &d) { return ::mkdir(d.c_str(), 0755); }

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

// Per-pass twiddle tables of a radix schedule: pass p (radix r, earlier
// product L) holds exp(-2 pi i s k / (L r)) at k*(r-1) + s-1, k < L, 0 < s < r.
int get_twiddles(d2dhip_plan_s *p, const std::vector<int> &radices, const TwTable **out) {
  auto it = p->tw.find(radices);
  if (it != p->tw.end()) {
    *out = &it->second;
    return D2DHIP_SUCCESS;
  }
  if ((int)radices.size() > kMaxPass)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "too many FFT passes");
  TwTable t;
  int L = 1, off = 0;
  std::vector<long double> twc, tws;
  for (int r : radices) {
    PassInfo pi;
    pi.r = r;
    pi.L = L;
    pi.tw_off = off;
    pi.invL = 1.0f / (float)L;
    t.passes.push_back(pi);
    for (int k = 0; k < L; ++k)
      for (int s = 1; s < r; ++s) {
        long double c, sn;
        unit_root((long long)s * k, (long long)L * r, c, sn);
        twc.push_back(c);
        tws.push_back(-sn);
      }
    off += L * (r - 1);
    L *= r;
  }
  const int rb = p->rb;
  t.bytes = std::max<size_t>(twc.size(), 1) * 2 * rb;
  std::vector<unsigned char> host(t.bytes, 0);
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
  if (hipMalloc(&t.d, t.bytes) != hipSuccess) return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(twiddles)");
  int st = hip_check(hipMemcpy(t.d, host.data(), t.bytes, hipMemcpyHostToDevice), "hipMemcpy(twiddles)");
  auto res = p->tw.emplace(radices, t);
  *out = &res.first->second;
  return st;
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
int choose_shape(int n, const std::vector<PassInfo> &passes, bool lines, int cb, long long C0,
                 AxisShape &s) {
  int txmax = lines ? 64 : std::max(1, 128 / cb);
  long long c0p2 = 1;
  while (c0p2 < C0 && c0p2 < txmax) c0p2 <<= 1;
  txmax = (int)std::min<long long>(txmax, c0p2);
  const int ldt_lines = (n % 2 == 0) ? n + 1 : n;
  for (int pass = 0; pass < 2; ++pass) {
    const size_t cap = pass == 0 ? 64 * 1024 : 160 * 1024;
    for (int tx = txmax; tx >= 1; tx >>= 1) {
      int nt = 64;
      for (const PassInfo &p : passes) {
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

// Bluestein tables of line length n: c_j = exp(-i pi j^2/n) (j < n) and the
// length-M DFT of w_m = exp(+i pi m^2/n) wrapped as w[m] = w[M-m] = w_m,
// evaluated in long double with j^2 reduced mod 2n exactly.
int get_bluestein(d2dhip_plan_s *p, int n, const BlueTables **out) {
  auto it = p->blue.find(n);
  if (it != p->blue.end()) {
    *out = &it->second;
    return D2DHIP_SUCCESS;
  }
  BlueTables b;
  b.M = 1;
  while (b.M < 2 * n - 1) b.M *= 2;
  const int M = b.M;
  std::vector<long double> cr(n), ci(n), wr(M, 0.0L), wi(M, 0.0L);
  for (int j = 0; j < n; ++j) {
    long double c, s;
    unit_root(((long long)j * j) % (2LL * n), 2LL * n, c, s);
    cr[j] = c;
    ci[j] = -s;
    wr[j] = c;
    wi[j] = s;
    if (j > 0) {
      wr[M - j] = c;
      wi[M - j] = s;
    }
  }
  // iterative radix-2 DFT of w (sign -1), bit-reversed input order
  for (int i = 1, j = 0; i < M; ++i) {
    int bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(wr[i], wr[j]);
      std::swap(wi[i], wi[j]);
    }
  }
  for (int len = 2; len <= M; len <<= 1) {
    for (int k = 0; k < len / 2; ++k) {
      long double c, s;
      unit_root(k, len, c, s);
      const long double tr = c, ti = -s;
      for (int i = k; i < M; i += len) {
        const int j = i + len / 2;
        const long double xr = wr[j] * tr - wi[j] * ti, xi = wr[j] * ti + wi[j] * tr;
        wr[j] = wr[i] - xr;
        wi[j] = wi[i] - xi;
        wr[i] += xr;
        wi[i] += xi;
      }
    }
  }
  const int rb = p->rb;
  std::vector<unsigned char> hc((size_t)n * 2 * rb), hw((size_t)M * 2 * rb);
  for (int j = 0; j < n; ++j) {
    if (rb == 8) {
      reinterpret_cast<double *>(hc.data())[2 * j] = (double)cr[j];
      reinterpret_cast<double *>(hc.data())[2 * j + 1] = (double)ci[j];
    } else {
      reinterpret_cast<float *>(hc.data())[2 * j] = (float)cr[j];
      reinterpret_cast<float *>(hc.data())[2 * j + 1] = (float)ci[j];
    }
  }
  for (int m = 0; m < M; ++m) {
    if (rb == 8) {
      reinterpret_cast<double *>(hw.data())[2 * m] = (double)wr[m];
      reinterpret_cast<double *>(hw.data())[2 * m + 1] = (double)wi[m];
    } else {
      reinterpret_cast<float *>(hw.data())[2 * m] = (float)wr[m];
      reinterpret_cast<float *>(hw.data())[2 * m + 1] = (float)wi[m];
    }
  }
  if (hipMalloc(&b.chirp, hc.size()) != hipSuccess || hipMalloc(&b.bhat, hw.size()) != hipSuccess)
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(Bluestein tables)");
  int st = hip_check(hipMemcpy(b.chirp, hc.data(), hc.size(), hipMemcpyHostToDevice), "hipMemcpy");
  if (!st) st = hip_check(hipMemcpy(b.bhat, hw.data(), hw.size(), hipMemcpyHostToDevice), "hipMemcpy");
  auto res = p->blue.emplace(n, b);
  *out = &res.first->second;
  return st;
}

// Fills the descriptor and launch shape of stage idx for kernel choice c
// (-1 = general kernel, else the c-th specialised candidate).
int configure(d2dhip_plan_s *p, int idx, int c) {
  StageGeom &g = p->geom[idx];
  const int cb = 2 * p->rb;
  AxisDesc &a = p->h_desc[idx];
  AxisShape &s = p->shape[idx];
  std::memset(&a, 0, sizeof(a));
  s = AxisShape();
  std::vector<int> radices;
  const FastVariant *fv = (c >= 0) ? g.cands[c] : nullptr;
  const BlueTables *bt = nullptr;
  int st;
  if (fv) {
    radices = fv->radices;
  } else {
    bool ok = true;
    radices = factor_radices(g.n, ok);
    if (!ok) {  // a prime factor above 17: Bluestein over a power-of-two length
      if ((st = get_bluestein(p, g.n, &bt))) return st;
      radices = factor_radices(bt->M, ok);
    }
  }
  const TwTable *tt = nullptr;
  st = get_twiddles(p, radices, &tt);
  if (st) return st;
  a.tw = tt->d;
  a.n = fv ? fv->n : (bt ? bt->M : g.n);
  if (bt) {
    a.blue = 1;
    a.nblue = g.n;
    a.chirp = bt->chirp;
    a.bhat = bt->bhat;
  }
  a.npass = (int)tt->passes.size();
  for (int i = 0; i < a.npass; ++i) a.pass[i] = tt->passes[i];
  a.C0r = g.C0r;
  a.C0 = (g.io == IO_C2C || (fv && fv->io == IO_DTT4)) ? g.C0r : (g.C0r + 1) / 2;
  a.C1 = g.C1;
  if (fv) {
    s.TX = fv->TX;
    s.nthreads = fv->NT;
    s.lds_bytes = fv->lds_bytes;
    s.ldj = fv->ldj;
    s.ldt = fv->ldt;
    st = hip_check(fv->prepare(), "hipFuncSetAttribute(fast kernel)");
    if (st) return st;
  } else {
    st = choose_shape(a.n, tt->passes, g.axis == 0, cb, a.C0, s);
    if (st) return st;
  }
  a.TX = s.TX;
  a.log2TX = ilog2(s.TX);
  const long long tiles0 = (a.C0 + s.TX - 1) / s.TX;
  if (tiles0 > (1LL << 30)) return set_error(D2DHIP_ERR_UNSUPPORTED, "pencil too large");
  a.tiles0 = (int)tiles0;
  a.in_s0 = g.in_s0;
  a.in_s1 = g.in_s1;
  a.in_sj = g.in_sj;
  a.out_s0 = g.out_s0;
  a.out_s1 = g.out_s1;
  a.out_sj = g.out_sj;
  a.ldj = s.ldj;
  a.ldt = s.ldt;
  a.nh = g.n / 2 + 1;
  a.dtt = g.dtt;
  a.dtt_first = g.dtt_first;
  a.dtt_len = g.dtt_len;
  a.dtt_n = g.dtt_n;
  a.dtw = (fv && fv->io == IO_DTT4) ? g.dtw4 : g.dtw;
  p->fast[idx] = fv;
  g.choice = c;
  return D2DHIP_SUCCESS;
}

int make_stage(d2dhip_plan_s *p, int idx, int axis, int io, const StageIo &sio, long long in_elems,
               long long out_elems) {
  StageGeom &g = p->geom[idx];
  g.axis = axis;
  g.io = io;
  g.n = axis == 0 ? p->nx : (axis == 1 ? p->ny : p->nz);
  g.C0r = sio.C0r;
  g.C1 = sio.C1;
  g.in_s0 = sio.in_s0;
  g.in_s1 = sio.in_s1;
  g.in_sj = sio.in_sj;
  g.out_s0 = sio.out_s0;
  g.out_s1 = sio.out_s1;
  g.out_sj = sio.out_sj;
  g.in_elems = in_elems;
  g.out_elems = out_elems;
  g.cands.clear();
  // Split-BF16 plans take the matrix-core variants of a length; a length
  // without one runs the FP32 VALU kernels (same storage, FP32 arithmetic).
  const bool split = p->prec == D2DHIP_PREC_SPLIT_BF16;
  if (!std::getenv("D2DHIP_NO_FAST")) {
    for (int pass = 0; pass < 2 && g.cands.empty(); ++pass) {
      const int want_mf = (split && pass == 0) ? 1 : 0;
      if (!split && pass == 1) break;
      for (const FastVariant &v : fast_registry())
        if (v.rb == p->rb && v.mf == want_mf && v.io == io && v.lines == (axis == 0 ? 1 : 0) &&
            v.n == g.n)
          g.cands.push_back(&v);
    }
  }
  return configure(p, idx, g.cands.empty() ? -1 : 0);
}

// (cos, sin)(pi k / (2L)), k < L, in long double rounded to the plan's
// precision: the twiddles of the half-length DCT / DST stages.
int get_dtt_twiddles(d2dhip_plan_s *p, int L, const void **out) {
  auto it = p->dtw.find(L);
  if (it != p->dtw.end()) {
    *out = it->second;
    return D2DHIP_SUCCESS;
  }
  const int rb = p->rb;
  std::vector<unsigned char> h((size_t)L * 2 * rb);
  for (int k = 0; k < L; ++k) {
    long double c, s;
    unit_root(k, 4LL * L, c, s);
    if (rb == 8) {
      reinterpret_cast<double *>(h.data())[2 * k] = (double)c;
      reinterpret_cast<double *>(h.data())[2 * k + 1] = (double)s;
    } else {
      reinterpret_cast<float *>(h.data())[2 * k] = (float)c;
      reinterpret_cast<float *>(h.data())[2 * k + 1] = (float)s;
    }
  }
  void *d = nullptr;
  if (hipMalloc(&d, h.size()) != hipSuccess) return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(DTT twiddles)");
  p->dtw[L] = d;
  *out = d;
  return hip_check(hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice), "hipMemcpy");
}

// exp(-i pi (4j+1) / (4L)), j < L/2, then exp(-i pi k / L), k < L/2: the pre-
// and post-twiddles of DCT-IV / DST-IV through an L/2-point DFT (kept under
// key -L beside the other tables).
int get_dtt4_twiddles(d2dhip_plan_s *p, int L, const void **out) {
  auto it = p->dtw.find(-L);
  if (it != p->dtw.end()) {
    *out = it->second;
    return D2DHIP_SUCCESS;
  }
  const int rb = p->rb, H = L / 2;
  std::vector<unsigned char> h((size_t)L * 2 * rb);
  for (int k = 0; k < L; ++k) {
    long double c, s;
    if (k < H)
      unit_root(4LL * k + 1, 8LL * L, c, s);
    else
      unit_root(k - H, 2LL * L, c, s);
    if (rb == 8) {
      reinterpret_cast<double *>(h.data())[2 * k] = (double)c;
      reinterpret_cast<double *>(h.data())[2 * k + 1] = (double)-s;
    } else {
      reinterpret_cast<float *>(h.data())[2 * k] = (float)c;
      reinterpret_cast<float *>(h.data())[2 * k + 1] = (float)-s;
    }
  }
  void *d = nullptr;
  if (hipMalloc(&d, h.size()) != hipSuccess) return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(DTT twiddles)");
  p->dtw[-L] = d;
  *out = d;
  return hip_check(hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice), "hipMemcpy");
}

// A real-to-real stage along `axis` of the (nx,ny,nz) real array, in place
// or not (same layout on both sides): `len` points from `first` (0-based) of
// every line go through a transform of `type`, computed through a DFT of
// length len (types 2, 3, 6, 7) or of the line's symmetric extension (the
// general kernel; Bluestein where that length has a prime factor above 17).
int make_dtt_stage(d2dhip_plan_s *p, int idx, int axis, int type, int first, int len) {
  StageGeom &g = p->geom[idx];
  const StageIo v = cols_view(axis, p->nx, p->ny, p->nz);
  g.axis = axis;
  g.io = IO_R2R;
  g.dtw = nullptr;
  g.dtw4 = nullptr;
  switch (type) {
    case 1: g.n = 2 * (len - 1); break;
    case 5: g.n = 2 * (len + 1); break;
    case 4:
    case 8: g.n = 8 * len; break;
    default: {  // 2, 3, 6, 7: the half-length algorithm
      g.n = len;
      const int st = get_dtt_twiddles(p, len, &g.dtw);
      if (st) return st;
    }
  }
  g.C0r = v.C0r;
  g.C1 = v.C1;
  g.in_s0 = v.in_s0;
  g.in_s1 = v.in_s1;
  g.in_sj = v.in_sj;
  g.out_s0 = v.out_s0;
  g.out_s1 = v.out_s1;
  g.out_sj = v.out_sj;
  g.in_elems = g.out_elems = (long long)p->nx * p->ny * p->nz;
  g.cands.clear();
  g.dtt = type;
  g.dtt_first = first;
  g.dtt_len = len;
  g.dtt_n = axis == 0 ? p->nx : (axis == 1 ? p->ny : p->nz);
  if (!std::getenv("D2DHIP_NO_FAST")) {  // specialised tiles of the stage's DFT length
    const int lines = axis == 0 ? 1 : 0;
    if ((type == 4 || type == 8) && len % 2 == 0) {  // first choice: L/2-point tiles
      for (const FastVariant &v : fast_registry())
        if (v.rb == p->rb && v.mf == 0 && v.io == IO_DTT4 && v.lines == lines && v.n == len / 2)
          g.cands.push_back(&v);
      if (!g.cands.empty()) {
        const int st = get_dtt4_twiddles(p, len, &g.dtw4);
        if (st) return st;
      }
    }
    const int want = (type == 2 || type == 6) ? IO_DTT2 : (type == 3 || type == 7) ? IO_DTT3 : IO_DTTX;
    for (const FastVariant &v : fast_registry())
      if (v.rb == p->rb && v.mf == 0 && v.io == want && v.lines == lines && v.n == g.n)
        g.cands.push_back(&v);
  }
  return configure(p, idx, g.cands.empty() ? -1 : 0);
}

int upload_chunk_descs(d2dhip_plan_s *p);
int upload_blocked_descs(d2dhip_plan_s *p);

// A captured graph holds the kernels, grids and chunk offsets of the
// configuration it was recorded with, while the descriptors it reads are
// rewritten in place; every change of configuration therefore drops it
// (graph_launch then fails until the plan is captured again).
void drop_graph(d2dhip_plan_s *p) {
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
}

// Host -> device copy of plan metadata, ordered after the work already
// queued on the plan's stream (which may still read the old values) and
// complete on return (the host copy may change right after).
int upload_sync(d2dhip_plan_s *p, void *dst, const void *src, size_t bytes, const char *what) {
  int st = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->stream), what);
  if (!st) st = hip_check(hipStreamSynchronize(p->stream), what);
  return st;
}

int upload_descs(d2dhip_plan_s *p) {
  if (p->capturing)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "plan is capturing: its configuration is fixed");
  drop_graph(p);
  int st = upload_sync(p, p->d_desc, p->h_desc, sizeof(AxisDesc) * d2dhip_plan_s::kStages,
                       "hipMemcpy(stage descriptors)");
  if (!st) st = upload_chunk_descs(p);
  if (!st) st = upload_blocked_descs(p);
  return st;
}

// Builds the seven stage descriptors of a plan and uploads them.
long long pencil_elems(const long long (&v)[3]) { return v[0] * v[1] * v[2]; }

StageIo pencil_view(int axis, const long long (&v)[3]) { return cols_view(axis, v[0], v[1], v[2]); }

// The pencils of a plan: the whole grid at one rank, else the rank's
// physical and spectral pencils.
void set_pencils(d2dhip_plan_s *p) {
  const long long nx = p->nx, ny = p->ny, nz = p->nz;
  if (!p->multi) {
    const long long sx = p->format == D2DHIP_PHYSICAL_IN_X ? nx / 2 + 1 : nx;
    const long long sz = p->format == D2DHIP_PHYSICAL_IN_X ? nz : nz / 2 + 1;
    for (int q = 0; q < 3; ++q) {
      p->pen[q][0] = nx, p->pen[q][1] = ny, p->pen[q][2] = nz;
      p->pen[3 + q][0] = sx, p->pen[3 + q][1] = ny, p->pen[3 + q][2] = sz;
    }
    return;
  }
  const d2dhip_decomp *dd[2] = {&p->ph, &p->sp};
  for (int h = 0; h < 2; ++h)
    for (int a = 0; a < 3; ++a) {
      p->pen[3 * h + 0][a] = dd[h]->xsz[a];
      p->pen[3 * h + 1][a] = dd[h]->ysz[a];
      p->pen[3 * h + 2][a] = dd[h]->zsz[a];
    }
}

// Builds the seven stage descriptors of a plan on its pencils and uploads
// them: c2c along x, y, z of the physical pencils; the r2c stage (physical
// x- or z-pencil to the spectral one); the spectral y stage; the other
// spectral c2c stage; the c2r stage.
int make_stages(d2dhip_plan_s *p) {
  set_pencils(p);
  const auto &px = p->pen[0], &py = p->pen[1], &pz = p->pen[2];
  const auto &sx = p->pen[3], &sy = p->pen[4], &sz = p->pen[5];
  int st;
  for (int a = 0; a < 3; ++a)
    if ((st = make_stage(p, a, a, IO_C2C, pencil_view(a, p->pen[a]), pencil_elems(p->pen[a]),
                         pencil_elems(p->pen[a]))))
      return st;
  if (p->format == D2DHIP_PHYSICAL_IN_X) {
    StageIo r = pencil_view(0, px);  // real lines of nx -> complex lines of nx/2+1
    r.out_s0 = sx[0];
    if ((st = make_stage(p, 3, 0, IO_R2C, r, pencil_elems(px), pencil_elems(sx)))) return st;
    if ((st = make_stage(p, 4, 1, IO_C2C, pencil_view(1, sy), pencil_elems(sy), pencil_elems(sy))))
      return st;
    if ((st = make_stage(p, 5, 2, IO_C2C, pencil_view(2, sz), pencil_elems(sz), pencil_elems(sz))))
      return st;
    StageIo c = pencil_view(0, px);
    c.in_s0 = sx[0];
    if ((st = make_stage(p, 6, 0, IO_C2R, c, pencil_elems(sx), pencil_elems(px)))) return st;
  } else {
    // real columns i + n1*j of the z axis; the complex output has the same
    // column and point strides
    if ((st = make_stage(p, 3, 2, IO_R2C, pencil_view(2, pz), pencil_elems(pz), pencil_elems(sz))))
      return st;
    if ((st = make_stage(p, 4, 1, IO_C2C, pencil_view(1, sy), pencil_elems(sy), pencil_elems(sy))))
      return st;
    if ((st = make_stage(p, 5, 0, IO_C2C, pencil_view(0, sx), pencil_elems(sx), pencil_elems(sx))))
      return st;
    if ((st = make_stage(p, 6, 2, IO_C2R, pencil_view(2, pz), pencil_elems(sz), pencil_elems(pz))))
      return st;
  }
  const size_t bytes = sizeof(AxisDesc) * d2dhip_plan_s::kStages;
  if (hipMalloc(&p->d_desc, bytes) != hipSuccess) {
    p->d_desc = nullptr;
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(stage descriptors) failed");
  }
  return upload_descs(p);
}

// ---- autotuner -----------------------------------------------------------------
// Times every kernel that can run a stage (the general kernel and each
// specialised variant) on scratch buffers of the stage's size and keeps the
// fastest.  Choices persist in a tab-separated cache file
// ($D2DHIP_TUNE_CACHE, default $HOME/.cache/d2dhip/tuning.tsv) keyed by the
// stage geometry, so later plans of the same shape skip the timing.

std::mutex g_tune_mu;
std::map<std::string, std::string> g_tune;
std::string g_tune_loaded_path;  // cache file g_tune holds; read again when the path changes
bool g_tune_loaded = false;

std::string tune_path() {
  if (const char *e = std::getenv("D2DHIP_TUNE_CACHE")) return e;
  const char *h = std::getenv("HOME");
  return std::string(h ? h : "/tmp") + "/.cache/d2dhip/tuning.tsv";
}

void tune_load() {
  const std::string path = tune_path();
  if (g_tune_loaded && path == g_tune_loaded_path) return;
  g_tune_loaded = true;
  g_tune_loaded_path = path;
  g_tune.clear();
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    char *tab = std::strchr(line, '\t');
    if (!tab) continue;
    *tab = 0;
    std::string val(tab + 1);
    while (!val.empty() && (val.back() == '\n' || val.back() == '\r')) val.pop_back();
    g_tune[line] = val;
  }
  std::fclose(f);
}

// Records a choice in memory and appends it to the cache file; a failed write
// marks the plan, whose set-up call then returns D2DHIP_WARN_TUNE_CACHE.
void tune_store(d2dhip_plan_s *p, const std::string &key, const std::string &val) {
  g_tune[key] = val;
  const std::string path = tune_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    std::string dir = path.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') (void)mkdir_one(dir.substr(0, i));
  }
  FILE *f = std::fopen(path.c_str(), "a");
  bool ok = f != nullptr;
  if (f) {
    ok = std::fprintf(f, "%s\t%s\n", key.c_str(), val.c_str()) > 0;
    ok = (std::fclose(f) == 0) && ok;
  }
  if (!ok) p->tune_write_failed = true;
}

// The key also carries a hash of the candidate kernels' names, so that a
// library with other specialised kernels for a stage tunes it again instead
// of reusing a choice made among fewer candidates.
std::string tune_key(const d2dhip_plan_s *p, int idx) {
  const StageGeom &g = p->geom[idx];
  unsigned h = 2166136261u;  // FNV-1a
  for (const FastVariant *v : g.cands) {
    for (const char *c = v->name; *c; ++c) h = (h ^ (unsigned char)*c) * 16777619u;
    h = (h ^ 0xffu) * 16777619u;
  }
  char b[288];
  std::snprintf(b, sizeof(b),
                "v2 pr=%d io=%d ax=%d n=%d C0r=%lld C1=%lld s=%lld,%lld,%lld,%lld,%lld,%lld c=%08x",
                p->prec, g.io, g.axis, g.n, g.C0r, g.C1, g.in_s0, g.in_s1, g.in_sj, g.out_s0,
                g.out_s1, g.out_sj, h);
  if (g.io == IO_R2R) {  // real-to-real stages: the transform and its points too
    char t[96];
    std::snprintf(t, sizeof(t), " dtt=%d,%d,%d,%d", g.dtt, g.dtt_first, g.dtt_len, g.dtt_n);
    return std::string(b) + t;
  }
  return b;
}

std::string choice_name(const d2dhip_plan_s *p, int idx) {
  const StageGeom &g = p->geom[idx];
  return g.choice < 0 ? std::string("general") : std::string(g.cands[g.choice]->name);
}

int time_stage(d2dhip_plan_s *p, int idx, const void *in, void *out, hipEvent_t e0, hipEvent_t e1,
               float *ms) {
  int st = upload_sync(p, p->d_desc + idx, &p->h_desc[idx], sizeof(AxisDesc), "hipMemcpy(descriptor)");
  if (st) return st;
  const AxisIo io = {in, out, 0};
  auto launch = [&]() {
    return p->fast[idx] ? p->fast[idx]->launch(p->d_desc + idx, p->h_desc[idx], io, p->stream)
                        : launch_axis(p->rb, p->geom[idx].io, p->d_desc + idx, p->h_desc[idx],
                                      io, p->shape[idx], p->stream);
  };
  if ((st = launch())) return st;
  if ((st = hip_check(hipEventRecord(e0, p->stream), "hipEventRecord"))) return st;
  const int reps = 3;
  for (int r = 0; r < reps && !st; ++r) st = launch();
  if (st) return st;
  if ((st = hip_check(hipEventRecord(e1, p->stream), "hipEventRecord"))) return st;
  if ((st = hip_check(hipEventSynchronize(e1), "hipEventSynchronize"))) return st;
  float t = 0.f;
  st = hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
  *ms = t / reps;
  return st;
}

// Whether `bytes` more of device memory can be taken for tuning scratch, with
// a sixteenth of the device left free: a plan too large to tune keeps its
// default choices (and caches nothing) instead of failing.
// D2DHIP_TUNE_MEM_LIMIT (bytes) caps the free memory this assumes.
bool tune_fits(size_t bytes) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (const char *e = std::getenv("D2DHIP_TUNE_MEM_LIMIT")) {
    const unsigned long long cap = std::strtoull(e, nullptr, 10);
    if (cap < fr) fr = (size_t)cap;
  }
  return bytes + tot / 16 <= fr;
}

int tune_stages(d2dhip_plan_s *p) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_load();
  size_t need = 0;
  bool any = false;
  for (int i = 0; i < d2dhip_plan_s::kStages; ++i) {
    const StageGeom &g = p->geom[i];
    if (g.cands.empty()) continue;
    any = true;
    const size_t bi = (size_t)g.in_elems * (g.io == IO_R2C ? 1 : 2) * p->rb;
    const size_t bo = (size_t)g.out_elems * (g.io == IO_C2R ? 1 : 2) * p->rb;
    need = std::max(need, std::max(bi, bo));
  }
  if (!any) return D2DHIP_SUCCESS;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = D2DHIP_SUCCESS;
  for (int i = 0; i < d2dhip_plan_s::kStages && !st; ++i) {
    StageGeom &g = p->geom[i];
    if (g.cands.empty()) continue;
    const std::string key = tune_key(p, i);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      int c = -1;
      for (size_t k = 0; k < g.cands.size(); ++k)
        if (it->second == g.cands[k]->name) c = (int)k;
      if (c >= 0 || it->second == "general") {
        st = configure(p, i, c);
        continue;
      }
    }
    if (!a) {
      if (!tune_fits(2 * need)) break;  // untuned stages keep their defaults
      if (hipMalloc(&a, need) != hipSuccess || hipMalloc(&b, need) != hipSuccess) {
        st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(tuning buffers) failed");
        break;
      }
      // the candidates run on noise, not zeros: arithmetic on zeros draws
      // less power, and the clock it allows flatters the heavier kernels
      st = launch_fill_uniform(a, (long long)(need / p->rb), p->rb == 8 ? D2DHIP_R8 : D2DHIP_R4, 7,
                               p->stream);
      if (!st) st = hip_check(hipMemsetAsync(b, 0, need, p->stream), "hipMemsetAsync");
      if (!st) st = hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
      if (!st) st = hip_check(hipEventCreate(&e0), "hipEventCreate");
      if (!st) st = hip_check(hipEventCreate(&e1), "hipEventCreate");
      if (st) break;
    }
    // every candidate once; then those within 6 % of the fastest twice more,
    // round robin, each keeping its best time: near-ties are common and one
    // run of each decides them by noise
    // the stages that the 3D drivers run in place on their output are timed
    // in place, on fresh noise for each candidate
    const bool inplace = i == 1 || i == 2 || i == 4 || i == 5;
    auto measure = [&](int c, float *ms) {
      int r = D2DHIP_SUCCESS;
      if (inplace)
        r = launch_fill_uniform(b, (long long)(need / p->rb), p->rb == 8 ? D2DHIP_R8 : D2DHIP_R4,
                                11 + c, p->stream);
      return r ? r : time_stage(p, i, inplace ? b : a, b, e0, e1, ms);
    };
    const int nc = (int)g.cands.size();
    std::vector<float> t(nc + 1, 1e30f);  // t[c + 1]: candidate c, -1 = general
    for (int c = -1; c < nc && !st; ++c) {
      if (configure(p, i, c)) continue;  // e.g. the general kernel cannot run this length
      st = measure(c, &t[c + 1]);
    }
    const float first = *std::min_element(t.begin(), t.end());
    std::vector<int> finalists;
    for (int c = -1; c < nc; ++c)
      if (t[c + 1] <= 1.06f * first) finalists.push_back(c);
    for (int round = 0; round < 2 && finalists.size() > 1 && !st; ++round)
      for (int c : finalists) {
        if (configure(p, i, c)) continue;
        float ms = 0.f;
        if ((st = measure(c, &ms))) break;
        t[c + 1] = std::min(t[c + 1], ms);
      }
    int best_c = g.choice;
    float best = 1e30f;
    for (int c = -1; c < nc; ++c)
      if (t[c + 1] < best) {
        best = t[c + 1];
        best_c = c;
      }
    if (!st) st = configure(p, i, best_c);
    if (!st) tune_store(p, key, choice_name(p, i));
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (!st) st = upload_descs(p);
  return st;
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
  int st = p->fast[idx]
               ? p->fast[idx]->launch(p->d_desc + idx, p->h_desc[idx], io, p->stream)
               : launch_axis(p->rb, p->geom[idx].io, p->d_desc + idx, p->h_desc[idx], io,
                             p->shape[idx], p->stream);
  if (p->prof) {
    if (!st) st = hip_check(hipEventRecord(b, p->stream), "hipEventRecord");
    p->recs.push_back({idx, a, b, 1.0});
  }
  return st;
}

// ---- plane chunking ------------------------------------------------------------

// Stage descriptor `idx` restricted to `planes` z-planes of a PHYSICAL_IN_X
// pencil: the x stages own ny lines per plane, the y stage one column set
// per plane.
AxisDesc chunk_desc(const d2dhip_plan_s *p, int idx, long long planes) {
  AxisDesc a = p->h_desc[idx];
  const StageGeom &g = p->geom[idx];
  if (g.axis == 0) {
    a.C0r = planes * p->ny;
    a.C0 = (g.io == IO_C2C) ? a.C0r : (a.C0r + 1) / 2;
    a.tiles0 = (int)((a.C0 + a.TX - 1) / a.TX);
  } else {
    a.C1 = planes;
  }
  return a;
}

// Element offsets of z-plane k0 in a stage's input and output, and the
// element sizes in bytes.
void chunk_offsets(const d2dhip_plan_s *p, int idx, long long k0, size_t *in_b, size_t *out_b) {
  const StageGeom &g = p->geom[idx];
  const size_t ie = (size_t)p->rb * (g.io == IO_R2C ? 1 : 2);
  const size_t oe = (size_t)p->rb * (g.io == IO_C2R ? 1 : 2);
  const long long in_pl = (g.axis == 0) ? (long long)p->ny * g.in_s0 : g.in_s1;
  const long long out_pl = (g.axis == 0) ? (long long)p->ny * g.out_s0 : g.out_s1;
  *in_b = (size_t)(k0 * in_pl) * ie;
  *out_b = (size_t)(k0 * out_pl) * oe;
}

// Rebuilds and uploads the eight chunk descriptors for the current kernel
// choices and chunk sizes.
int upload_chunk_descs(d2dhip_plan_s *p) {
  if (p->format != D2DHIP_PHYSICAL_IN_X) return D2DHIP_SUCCESS;
  const int stage_of[4] = {3, 4, 4, 6};
  for (int q = 0; q < 4; ++q) {
    const long long K = (q < 2) ? p->chunk_fwd : p->chunk_bwd;
    const long long tail = K > 0 ? p->nz - (p->nz - 1) / K * K : 0;
    p->h_cdesc[2 * q] = chunk_desc(p, stage_of[q], K > 0 ? K : 1);
    p->h_cdesc[2 * q + 1] = chunk_desc(p, stage_of[q], tail > 0 ? tail : 1);
  }
  if (!p->d_cdesc && hipMalloc(&p->d_cdesc, sizeof(p->h_cdesc)) != hipSuccess) {
    p->d_cdesc = nullptr;
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(chunk descriptors) failed");
  }
  drop_graph(p);
  return upload_sync(p, p->d_cdesc, p->h_cdesc, sizeof(p->h_cdesc), "hipMemcpy(chunk descriptors)");
}

// One stage launch over z-planes [k0, k0 + planes) through chunk slot q.
int run_chunk(d2dhip_plan_s *p, int q, int idx, long long k0, long long planes, long long K,
              const void *in, void *out, int conj) {
  const int slot = 2 * q + (planes == K ? 0 : 1);
  size_t ib = 0, ob = 0;
  chunk_offsets(p, idx, k0, &ib, &ob);
  const AxisIo io = {static_cast<const unsigned char *>(in) + ib, static_cast<unsigned char *>(out) + ob,
                     conj};
  hipEvent_t a = nullptr, b = nullptr;
  if (p->prof) {
    a = take_event(p);
    b = take_event(p);
    if (!a || !b) return set_error(D2DHIP_ERR_HIP, "hipEventCreate failed");
    int st = hip_check(hipEventRecord(a, p->stream), "hipEventRecord");
    if (st) return st;
  }
  const AxisDesc *dd = p->d_cdesc + slot;
  int st = p->fast[idx] ? p->fast[idx]->launch(dd, p->h_cdesc[slot], io, p->stream)
                        : launch_axis(p->rb, p->geom[idx].io, dd, p->h_cdesc[slot], io,
                                      p->shape[idx], p->stream);
  if (p->prof) {
    if (!st) st = hip_check(hipEventRecord(b, p->stream), "hipEventRecord");
    p->recs.push_back({idx, a, b, (double)planes / (double)p->nz});
  }
  return st;
}

// The stage pair (i1 then i2) over chunks of K z-planes: q1/q2 are their
// chunk slots; i1 reads in1 and writes mid, i2 reads mid and writes out2.
int run_pair_chunked(d2dhip_plan_s *p, long long K, int q1, int i1, int q2, int i2,
                     const void *in1, void *mid, void *out2, int conj) {
  for (long long k0 = 0; k0 < p->nz; k0 += K) {
    const long long pl = std::min<long long>(K, p->nz - k0);
    int st = run_chunk(p, q1, i1, k0, pl, K, in1, mid, conj);
    if (!st) st = run_chunk(p, q2, i2, k0, pl, K, mid, out2, conj);
    if (st) return st;
  }
  return D2DHIP_SUCCESS;
}

// ---- blocked spectral layout ----------------------------------------------------

// Block width (i-columns per block) of the plan's blocked layout.
int blk_w(const d2dhip_plan_s *p) { return p->layout == D2DHIP_LAYOUT_BLOCKED32 ? 32 : 16; }
long long blk_nyb(const d2dhip_plan_s *p) { return (long long)p->ny + 1; }
long long blk_bs(const d2dhip_plan_s *p) { return blk_w(p) * blk_nyb(p) * p->nz; }
// Complex elements of the blocked pencil (whole blocks).
long long blk_elems(const d2dhip_plan_s *p) {
  const int w = blk_w(p);
  return ((p->nx / 2 + 1 + w - 1) / w) * blk_bs(p);
}

bool blk_possible(const d2dhip_plan_s *p) {
  return p->format == D2DHIP_PHYSICAL_IN_X && !p->multi && !(p->flags & D2DHIP_FLAG_INPLACE);
}

// Stage idx's descriptor (its kernel, twiddles and tile width) on other
// columns: C0 x C1 columns, column c0 at c0 * s0 + (c0 >> bsh) * b0 with the
// layout's block width 2^bsh.
AxisDesc blk_cols(const d2dhip_plan_s *p, int idx, long long C0, long long C1, long long is0,
                  long long ib0, long long is1, long long isj, long long os0, long long ob0,
                  long long os1, long long osj) {
  AxisDesc a = p->h_desc[idx];
  a.bsh = ilog2(blk_w(p));
  a.C0 = a.C0r = C0;
  a.C1 = C1;
  a.tiles0 = (int)((C0 + a.TX - 1) / a.TX);
  a.in_s0 = is0;
  a.in_b0 = ib0;
  a.in_s1 = is1;
  a.in_sj = isj;
  a.out_s0 = os0;
  a.out_b0 = ob0;
  a.out_s1 = os1;
  a.out_sj = osj;
  return a;
}

// Rebuilds and uploads the blocked-layout descriptors for the current kernel
// choices (nothing while the layout is plain).
int upload_blocked_descs(d2dhip_plan_s *p) {
  if (p->layout == D2DHIP_LAYOUT_PLAIN) return D2DHIP_SUCCESS;
  const long long nxc = p->nx / 2 + 1, ny = p->ny, nz = p->nz, W = blk_w(p);
  const long long nyb = blk_nyb(p), bs = blk_bs(p), nfull = nxc / W, r = nxc % W;
  // x stages: point i of line l = j + k * nyb at (i / W) * bs + (i % W) + W l
  // (the plain x stages see the ny * nz lines as one run; here lines j of
  // plane k, so that the line index can skip the spare j row)
  for (int s = 0; s < 2; ++s) {
    AxisDesc a = p->h_desc[s ? 6 : 3];
    a.C0r = ny;
    a.C0 = (ny + 1) / 2;
    a.C1 = nz;
    a.tiles0 = (int)((a.C0 + a.TX - 1) / a.TX);
    a.in_s1 = s ? nxc * ny : (long long)p->nx * ny;
    a.out_s1 = s ? (long long)p->nx * ny : nxc * ny;
    a.pack = nullptr;
    a.npack = -1;
    a.pack_l1 = nyb;
    a.pack_bs = bs;
    a.bsh = ilog2((int)W);
    p->h_bdesc[s ? 5 : 0] = a;
  }
  // z over whole blocks: the columns (j, i % W) of a block are contiguous
  p->h_bdesc[1] = blk_cols(p, 5, W * ny, nfull > 0 ? nfull : 1, 1, 0, bs, W * nyb, 1, 0, bs,
                           W * nyb);
  // z over the last partial block (its first r columns of each row j),
  // launched on the block's first element
  p->h_bdesc[2] = blk_cols(p, 5, r > 0 ? r : 1, ny, 1, 0, W, W * nyb, 1, 0, W, W * nyb);
  // y: blocks -> plain (nx/2+1, ny, nz) forward, plain -> blocks inverse
  p->h_bdesc[3] = blk_cols(p, 4, nxc, nz, 1, bs - W, W * nyb, W, 1, 0, nxc * ny, nxc);
  p->h_bdesc[4] = blk_cols(p, 4, nxc, nz, 1, 0, nxc * ny, nxc, 1, bs - W, W * nyb, W);
  if (!p->d_bdesc && hipMalloc(&p->d_bdesc, sizeof(p->h_bdesc)) != hipSuccess) {
    p->d_bdesc = nullptr;
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(blocked descriptors) failed");
  }
  drop_graph(p);
  return upload_sync(p, p->d_bdesc, p->h_bdesc, sizeof(p->h_bdesc), "hipMemcpy(blocked descriptors)");
}

// One launch of stage idx's kernel through blocked descriptor `slot`, the
// buffers offset by in_off / out_off complex elements; profiled as stage idx
// (weight w of one stage call).
int run_blk(d2dhip_plan_s *p, int slot, int idx, const void *in, void *out, int conj,
            long long in_off = 0, long long out_off = 0, double w = 1.0) {
  const size_t cb = 2 * (size_t)p->rb;
  const AxisIo io = {static_cast<const unsigned char *>(in) + in_off * cb,
                     static_cast<unsigned char *>(out) + out_off * cb, conj};
  hipEvent_t a = nullptr, b = nullptr;
  if (p->prof) {
    a = take_event(p);
    b = take_event(p);
    if (!a || !b) return set_error(D2DHIP_ERR_HIP, "hipEventCreate failed");
    int st = hip_check(hipEventRecord(a, p->stream), "hipEventRecord");
    if (st) return st;
  }
  const AxisDesc *dd = p->d_bdesc + slot;
  int st = p->fast[idx] ? p->fast[idx]->launch(dd, p->h_bdesc[slot], io, p->stream)
                        : launch_axis(p->rb, p->geom[idx].io, dd, p->h_bdesc[slot], io,
                                      p->shape[idx], p->stream);
  if (p->prof) {
    if (!st) st = hip_check(hipEventRecord(b, p->stream), "hipEventRecord");
    p->recs.push_back({idx, a, b, w});
  }
  return st;
}

// The z stage on the blocked pencil in wk (in place): the whole blocks, then
// the last partial one.
int run_blk_z(d2dhip_plan_s *p, void *wk, int conj) {
  const long long nxc = p->nx / 2 + 1, nfull = nxc / blk_w(p), r = nxc % blk_w(p);
  int st = D2DHIP_SUCCESS;
  if (nfull > 0) st = run_blk(p, 1, 5, wk, wk, conj);
  if (!st && r > 0)
    st = run_blk(p, 2, 5, wk, wk, conj, nfull * blk_bs(p), nfull * blk_bs(p), nfull > 0 ? 0.0 : 1.0);
  return st;
}

// exec_r2c / the c2r stages in the blocked layout (wk: the plan's workspace).
int run_blocked_r2c(d2dhip_plan_s *p, const void *in_r, void *out_c) {
  int st = run_blk(p, 0, 3, in_r, p->wk, 0);
  if (!st) st = run_blk_z(p, p->wk, 0);
  if (!st) st = run_blk(p, 3, 4, p->wk, out_c, 0);
  return st;
}

int run_blocked_c2r(d2dhip_plan_s *p, const void *in_c, void *out_r) {
  int st = run_blk(p, 4, 4, in_c, p->wk, 1);
  if (!st) st = run_blk_z(p, p->wk, 1);
  if (!st) st = run_blk(p, 5, 6, p->wk, out_r, 1);
  return st;
}

// Whether an exec call runs in the blocked layout: set, whole stages, none
// skipped, and the plan's own workspace.
bool blocked_now(const d2dhip_plan_s *p, int chunk, const void *wk) {
  return p->layout != D2DHIP_LAYOUT_PLAIN && chunk == 0 && wk && wk == p->wk &&
         !(p->flags & (D2DHIP_FLAG_SKIP_X | D2DHIP_FLAG_SKIP_Y | D2DHIP_FLAG_SKIP_Z));
}

// Workspace bytes a plan needs for its layout.
size_t wk_need(const d2dhip_plan_s *p) {
  const long long nsp = (p->format == D2DHIP_PHYSICAL_IN_X)
                            ? (long long)(p->nx / 2 + 1) * p->ny * p->nz
                            : (long long)p->nx * p->ny * (p->nz / 2 + 1);
  const long long e = p->layout != D2DHIP_LAYOUT_PLAIN ? std::max(nsp, blk_elems(p)) : nsp;
  return (size_t)e * 2 * p->rb;
}

// Switches the layout, resizing the workspace to what the layout needs.
int set_layout(d2dhip_plan_s *p, int layout) {
  if (layout != D2DHIP_LAYOUT_PLAIN && layout != D2DHIP_LAYOUT_BLOCKED &&
      layout != D2DHIP_LAYOUT_BLOCKED32)
    return set_error(D2DHIP_ERR_INVALID_VALUE,
                     "layout must be D2DHIP_LAYOUT_PLAIN, _BLOCKED or _BLOCKED32");
  if (layout != D2DHIP_LAYOUT_PLAIN && !blk_possible(p))
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "blocked layout: a whole-grid PHYSICAL_IN_X plan with a workspace");
  if (p->capturing)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "plan is capturing: its layout is fixed");
  drop_graph(p);
  p->layout = layout;
  if (!(p->flags & D2DHIP_FLAG_INPLACE) && !p->multi) {
    const size_t need = wk_need(p);
    if (need > p->wk_bytes) {  // (grows only: switching back keeps the larger one)
      // the old workspace may still be read by queued work (on whichever
      // stream the plan last ran)
      int st = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      if (st) return st;
      if (p->wk) (void)hipFree(p->wk);
      p->wk = nullptr;
      p->wk_bytes = 0;
      if (hipMalloc(&p->wk, need) != hipSuccess) {
        p->wk = nullptr;
        p->layout = D2DHIP_LAYOUT_PLAIN;
        return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(workspace) failed");
      }
      p->wk_bytes = need;
    }
  }
  return upload_blocked_descs(p);
}

// Index of the candidate of stage idx named `name` (-1: "general"; -2: none).
int cand_index(const d2dhip_plan_s *p, int idx, const std::string &name) {
  if (name == "general") return -1;
  const StageGeom &g = p->geom[idx];
  for (size_t k = 0; k < g.cands.size(); ++k)
    if (name == g.cands[k]->name) return (int)k;
  return -2;
}

// Kernel choices of stages 3-6 (c[0..3]) and the layout, uploaded together.
int use_layout(d2dhip_plan_s *p, int layout, const int c[4]) {
  int st = D2DHIP_SUCCESS;
  for (int k = 0; k < 4 && !st; ++k) st = configure(p, 3 + k, c[k]);
  if (!st) st = set_layout(p, layout);
  if (!st) st = upload_descs(p);
  return st;
}

// Times the r2c + c2r round trip in each layout (plain, blocks of 16 and of
// 32 columns) on scratch buffers and keeps the fastest, each layout with its
// own kernels for stages 3-6 (a blocked layout's picked by timing every
// candidate of each stage in that layout); cached like the kernel choices, as
// "0" or "<layout>|<k3>|<k4>|<k5>|<k6>".
int tune_layout(d2dhip_plan_s *p) {
  if (!blk_possible(p) || (p->flags & (D2DHIP_FLAG_SKIP_Y | D2DHIP_FLAG_SKIP_Z)) || p->chunk_fwd ||
      p->chunk_bwd)
    return D2DHIP_SUCCESS;
  char kb[512];
  std::snprintf(kb, sizeof(kb), "v4 layout pr=%d n=%d,%d,%d k=%s|%s|%s|%s", p->prec, p->nx, p->ny,
                p->nz, choice_name(p, 3).c_str(), choice_name(p, 4).c_str(),
                choice_name(p, 5).c_str(), choice_name(p, 6).c_str());
  const std::string key(kb);
  constexpr int kLayouts = 3;  // D2DHIP_LAYOUT_PLAIN, _BLOCKED, _BLOCKED32
  int ch[kLayouts][4];
  for (int l = 0; l < kLayouts; ++l)
    for (int k = 0; k < 4; ++k) ch[l][k] = p->geom[3 + k].choice;
  auto it = g_tune.find(key);
  if (it != g_tune.end()) {
    const std::string &v = it->second;
    const int lay = v.empty() ? 0 : v[0] - '0';
    if (lay >= 1 && lay < kLayouts && v.size() > 2 && v[1] == '|') {
      size_t at = 2;
      bool ok = true;
      for (int k = 0; k < 4 && ok; ++k) {
        const size_t bar = k < 3 ? v.find('|', at) : v.size();
        ok = bar != std::string::npos;
        if (ok) {
          ch[lay][k] = cand_index(p, 3 + k, v.substr(at, bar - at));
          ok = ch[lay][k] != -2;
          at = bar + 1;
        }
      }
      if (ok) return use_layout(p, lay, ch[lay]);
    }
    return set_layout(p, D2DHIP_LAYOUT_PLAIN);
  }
  const long long nph = (long long)p->nx * p->ny * p->nz;
  const long long nsp = (long long)(p->nx / 2 + 1) * p->ny * p->nz;
  // scratch, and the workspace's growth to the larger of the blocked layouts
  const long long nxc = p->nx / 2 + 1, rows = ((long long)p->ny + 1) * p->nz;
  const size_t blk = (size_t)std::max((nxc + 15) / 16 * 16, (nxc + 31) / 32 * 32) * rows * 2 * p->rb;
  const size_t grow = blk > p->wk_bytes ? blk - p->wk_bytes : 0;
  if (!tune_fits((size_t)nph * p->rb * 2 + (size_t)nsp * 2 * p->rb + grow))
    return set_layout(p, D2DHIP_LAYOUT_PLAIN);
  void *a = nullptr, *b = nullptr, *c = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = D2DHIP_SUCCESS;
  if (hipMalloc(&a, (size_t)nph * p->rb) != hipSuccess ||
      hipMalloc(&b, (size_t)nsp * 2 * p->rb) != hipSuccess ||
      hipMalloc(&c, (size_t)nph * p->rb) != hipSuccess)
    st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(tuning buffers) failed");
  if (!st) st = launch_fill_uniform(a, nph, p->rb == 8 ? D2DHIP_R8 : D2DHIP_R4, 7, p->stream);
  if (!st) st = hip_check(hipEventCreate(&e0), "hipEventCreate");
  if (!st) st = hip_check(hipEventCreate(&e1), "hipEventCreate");
  const int prof = p->prof;
  p->prof = 0;
  // ms per call of op over reps calls, after one untimed call
  auto timed = [&](auto op, int reps, float *ms) {
    int r = op();
    if (!r) r = hip_check(hipEventRecord(e0, p->stream), "hipEventRecord");
    for (int q = 0; q < reps && !r; ++q) r = op();
    if (!r) r = hip_check(hipEventRecord(e1, p->stream), "hipEventRecord");
    if (!r) r = hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float t = 0.f;
    if (!r) r = hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
    *ms = t / reps;
    return r;
  };
  // each blocked layout's kernels, stage by stage in that layout: z (in place
  // on the workspace, refilled by a blocked r2c of the noise before each
  // candidate, so that the values stay finite), y forward, r2c x into the
  // blocks, c2r x out of them
  const int order[4] = {5, 4, 3, 6};
  for (int lay = 1; lay < kLayouts && !st; ++lay) {
    if ((st = use_layout(p, lay, ch[lay]))) break;
    for (int q = 0; q < 4 && !st; ++q) {
      const int idx = order[q];
      float best_t = 1e30f;
      int best_c = p->geom[idx].choice;
      for (int cc = -1; cc < (int)p->geom[idx].cands.size() && !st; ++cc) {
        if (configure(p, idx, cc)) continue;  // e.g. the general kernel cannot run this length
        if ((st = upload_descs(p))) break;
        if ((st = run_blocked_r2c(p, a, b))) break;
        float t = 0.f;
        st = timed([&]() {
          switch (idx) {
            case 5: return run_blk_z(p, p->wk, 0);
            case 4: return run_blk(p, 3, 4, p->wk, b, 0);
            case 3: return run_blk(p, 0, 3, a, p->wk, 0);
            default: return run_blk(p, 5, 6, p->wk, c, 1);
          }
        }, 2, &t);
        if (!st && t < best_t) {
          best_t = t;
          best_c = cc;
        }
      }
      if (!st) st = configure(p, idx, best_c);
      if (!st) st = upload_descs(p);
      ch[lay][idx - 3] = best_c;
    }
  }
  // round trips, two rounds, each layout's best kept (one run each is
  // decided by noise when two are close)
  float ms[kLayouts] = {1e30f, 1e30f, 1e30f};
  for (int round = 0; round < 2 && !st; ++round) {
    for (int lay = 0; lay < kLayouts && !st; ++lay) {
      if ((st = use_layout(p, lay, ch[lay]))) break;
      float t = 0.f;
      st = timed([&]() {
        int r = lay ? run_blocked_r2c(p, a, b) : d2dhip_exec_r2c(p, a, b);
        return r ? r : (lay ? run_blocked_c2r(p, b, c) : d2dhip_exec_c2r(p, b, c));
      }, 3, &t);
      if (!st) ms[lay] = std::min(ms[lay], t);
    }
  }
  p->prof = prof;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (c) (void)hipFree(c);
  int best = 0;
  for (int lay = 1; lay < kLayouts; ++lay)
    if (ms[lay] < ms[best]) best = lay;
  int st2 = use_layout(p, best, ch[best]);
  if (!st && !st2) {
    std::string v = "0";
    if (best) {
      v = std::to_string(best);
      for (int k = 3; k <= 6; ++k) v += "|" + choice_name(p, k);
    }
    tune_store(p, key, v);
  }
  return st ? st : st2;
}

// Times the forward pair (r2c x, y) and the inverse pair (y, c2r x) for
// chunk sizes K in {0, 2, 4, ..., 64} (K < nz) on scratch buffers and keeps
// the fastest of each; cached like the kernel choices, keyed by the grid and
// the chosen kernels.
int tune_chunks(d2dhip_plan_s *p) {
  if (p->format != D2DHIP_PHYSICAL_IN_X || p->nz < 2) return D2DHIP_SUCCESS;
  char kb[512];
  std::snprintf(kb, sizeof(kb), "v2 chunk pr=%d n=%d,%d,%d k=%s|%s|%s", p->prec, p->nx, p->ny, p->nz,
                choice_name(p, 3).c_str(), choice_name(p, 4).c_str(), choice_name(p, 6).c_str());
  const std::string key(kb);
  auto it = g_tune.find(key);
  if (it != g_tune.end() && std::sscanf(it->second.c_str(), "%d,%d", &p->chunk_fwd, &p->chunk_bwd) == 2)
    return upload_chunk_descs(p);
  const long long nph = (long long)p->nx * p->ny * p->nz;
  const long long nsp = (long long)(p->nx / 2 + 1) * p->ny * p->nz;
  const size_t need = (size_t)std::max(nph * p->rb, nsp * 2 * p->rb);
  if (!tune_fits(2 * need)) return upload_chunk_descs(p);  // whole stages
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = D2DHIP_SUCCESS;
  if (hipMalloc(&a, need) != hipSuccess || hipMalloc(&b, need) != hipSuccess)
    st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(tuning buffers) failed");
  if (!st)
    st = launch_fill_uniform(a, (long long)(need / p->rb), p->rb == 8 ? D2DHIP_R8 : D2DHIP_R4, 7,
                             p->stream);
  if (!st) st = hip_check(hipMemsetAsync(b, 0, need, p->stream), "hipMemsetAsync");
  if (!st) st = hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  if (!st) st = hip_check(hipEventCreate(&e0), "hipEventCreate");
  if (!st) st = hip_check(hipEventCreate(&e1), "hipEventCreate");
  const int prof = p->prof;
  p->prof = 0;
  int best[2] = {0, 0};
  for (int dir = 0; dir < 2 && !st; ++dir) {
    float best_ms = 1e30f;
    for (int K = 0; K <= 64 && K < p->nz && !st; K = (K == 0 ? 2 : 2 * K)) {
      (dir == 0 ? p->chunk_fwd : p->chunk_bwd) = K;
      if ((st = upload_chunk_descs(p))) break;
      auto once = [&]() {
        if (dir == 0) {
          if (K == 0) {
            int r = run(p, 3, a, b, 0);
            return r ? r : run(p, 4, b, b, 0);
          }
          return run_pair_chunked(p, K, 0, 3, 1, 4, a, b, b, 0);
        }
        if (K == 0) {
          int r = run(p, 4, a, a, 1);
          return r ? r : run(p, 6, a, b, 1);
        }
        return run_pair_chunked(p, K, 2, 4, 3, 6, a, a, b, 1);
      };
      if ((st = once())) break;
      if ((st = hip_check(hipEventRecord(e0, p->stream), "hipEventRecord"))) break;
      const int reps = 3;
      for (int r = 0; r < reps && !st; ++r) st = once();
      if (st) break;
      if ((st = hip_check(hipEventRecord(e1, p->stream), "hipEventRecord"))) break;
      if ((st = hip_check(hipEventSynchronize(e1), "hipEventSynchronize"))) break;
      float t = 0.f;
      if ((st = hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime"))) break;
      if (t / reps < best_ms) {
        best_ms = t / reps;
        best[dir] = K;
      }
    }
  }
  p->prof = prof;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  p->chunk_fwd = best[0];
  p->chunk_bwd = best[1];
  if (!st) {
    char v[64];
    std::snprintf(v, sizeof(v), "%d,%d", best[0], best[1]);
    tune_store(p, key, v);
  }
  int st2 = upload_chunk_descs(p);
  return st ? st : st2;
}

}  // namespace

extern "C" {

namespace {

// Plan creation for one rank: the whole grid (ph == nullptr), or the
// pencils of the rank whose physical and spectral decompositions are ph, sp.
int plan_create(d2dhip_plan *plan, int format, int nx, int ny, int nz, int prec, unsigned flags,
                void *stream, const d2dhip_decomp *ph, const d2dhip_decomp *sp) {
  if (!plan) return set_error(D2DHIP_ERR_NULL, "plan pointer is NULL");
  *plan = nullptr;
  if (nx <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for nx");
  if (ny <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for ny");
  if (nz <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for nz");
  if (format != D2DHIP_PHYSICAL_IN_X && format != D2DHIP_PHYSICAL_IN_Z)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for pencil");
  if (prec != D2DHIP_PREC_F64 && prec != D2DHIP_PREC_F32 && prec != D2DHIP_PREC_SPLIT_BF16)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid precision");
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
  if (ph) {
    p->ph = *ph;
    p->sp = *sp;
    p->multi = ph->p_row * ph->p_col > 1;
  }
  int st = axis_kernel_prepare(p->rb);
  if (!st) st = make_stages(p);
  const bool tune = (flags & D2DHIP_FLAG_TUNE) || std::getenv("D2DHIP_TUNE");
  if (!st && tune) st = tune_stages(p);
  if (!st && !p->multi) {
    // D2DHIP_CHUNK_PLANES="f,b" fixes the plane chunks (0 = whole stages)
    const char *e = std::getenv("D2DHIP_CHUNK_PLANES");
    int f = 0, b = 0;
    if (e && std::sscanf(e, "%d,%d", &f, &b) == 2) {
      st = d2dhip_plan_set_chunking(p, f, b);
    } else if (tune) {
      std::lock_guard<std::mutex> lock(g_tune_mu);
      tune_load();
      st = tune_chunks(p);
    }
  }
  if (!st && !p->multi && !(flags & D2DHIP_FLAG_INPLACE)) {
    const long long nsp = (format == D2DHIP_PHYSICAL_IN_X)
                              ? (long long)(nx / 2 + 1) * ny * nz
                              : (long long)nx * ny * (nz / 2 + 1);
    p->wk_bytes = (size_t)nsp * 2 * p->rb;
    if (hipMalloc(&p->wk, p->wk_bytes) != hipSuccess) {
      p->wk = nullptr;
      st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(c2r workspace) failed");
    }
  }
  if (!st && blk_possible(p)) {
    // D2DHIP_LAYOUT=plain|blocked fixes the layout; tuning times both
    const char *e = std::getenv("D2DHIP_LAYOUT");
    if (e && (std::strcmp(e, "blocked") == 0 || std::strcmp(e, "blocked32") == 0)) {
      st = set_layout(p, e[7] ? D2DHIP_LAYOUT_BLOCKED32 : D2DHIP_LAYOUT_BLOCKED);
    } else if (!(e && std::strcmp(e, "plain") == 0) && tune) {
      std::lock_guard<std::mutex> lock(g_tune_mu);
      tune_load();
      st = tune_layout(p);
    }
  }
  if (st) {
    d2dhip_plan_destroy(p);
    return st;
  }
  *plan = p;
  if (p->tune_write_failed)
    return set_error(D2DHIP_WARN_TUNE_CACHE, "tuning choices could not be written to " + tune_path());
  return D2DHIP_SUCCESS;
}

}  // namespace

int d2dhip_plan_create(d2dhip_plan *plan, int format, int nx, int ny, int nz, int prec,
                       unsigned flags, void *stream) {
  Region region(PROF_FFT, "fft_init");
  return plan_create(plan, format, nx, ny, nz, prec, flags, stream, nullptr, nullptr);
}

// decomp_2d_fft_init on a p_row x p_col process grid, for one rank: the
// physical decomposition of (nx, ny, nz) and the spectral one of
// (nx/2+1, ny, nz) or (nx, ny, nz/2+1) (src/fft_common.f90:198-216).
int d2dhip_plan_create_rank(d2dhip_plan *plan, int format, int nx, int ny, int nz, int prec,
                            int p_row, int p_col, int coord_row, int coord_col, unsigned flags,
                            void *stream) {
  Region region(PROF_FFT, "fft_init");
  if (!plan) return set_error(D2DHIP_ERR_NULL, "plan pointer is NULL");
  *plan = nullptr;
  if (format != D2DHIP_PHYSICAL_IN_X && format != D2DHIP_PHYSICAL_IN_Z)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value for pencil");
  if (nx <= 0 || ny <= 0 || nz <= 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid grid");
  d2dhip_decomp ph, sp;
  int st = d2dhip_decomp_info_init(nx, ny, nz, p_row, p_col, coord_row, coord_col, &ph);
  if (!st)
    st = format == D2DHIP_PHYSICAL_IN_X
             ? d2dhip_decomp_info_init(nx / 2 + 1, ny, nz, p_row, p_col, coord_row, coord_col, &sp)
             : d2dhip_decomp_info_init(nx, ny, nz / 2 + 1, p_row, p_col, coord_row, coord_col, &sp);
  if (st) return st;
  return plan_create(plan, format, nx, ny, nz, prec, flags, stream, &ph, &sp);
}

int d2dhip_plan_get_rank(d2dhip_plan p, d2dhip_decomp *ph, d2dhip_decomp *sp) {
  if (!p || !ph || !sp) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (p->multi) {
    *ph = p->ph;
    *sp = p->sp;
    return D2DHIP_SUCCESS;
  }
  int st = d2dhip_decomp_info_init(p->nx, p->ny, p->nz, 1, 1, 0, 0, ph);
  if (!st)
    st = p->format == D2DHIP_PHYSICAL_IN_X
             ? d2dhip_decomp_info_init(p->nx / 2 + 1, p->ny, p->nz, 1, 1, 0, 0, sp)
             : d2dhip_decomp_info_init(p->nx, p->ny, p->nz / 2 + 1, 1, 1, 0, 0, sp);
  return st;
}

int d2dhip_plan_destroy(d2dhip_plan p) {
  Region region(PROF_FFT, "fft_fin");
  if (!p) return D2DHIP_SUCCESS;
  for (auto &r : p->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : p->pool) (void)hipEventDestroy(e);
  for (auto &kv : p->tw)
    if (kv.second.d) (void)hipFree(kv.second.d);
  for (auto &kv : p->dtw) (void)hipFree(kv.second);
  for (auto &kv : p->blue) {
    if (kv.second.chirp) (void)hipFree(kv.second.chirp);
    if (kv.second.bhat) (void)hipFree(kv.second.bhat);
  }
  if (p->wk) (void)hipFree(p->wk);
  if (p->d_desc) (void)hipFree(p->d_desc);
  if (p->d_cdesc) (void)hipFree(p->d_cdesc);
  if (p->d_bdesc) (void)hipFree(p->d_bdesc);
  if (p->d_pack) (void)hipFree(p->d_pack);
  if (p->d_pdesc) (void)hipFree(p->d_pdesc);
  if (p->d_ypack) (void)hipFree(p->d_ypack);
  if (p->d_ypack2) (void)hipFree(p->d_ypack2);
  if (p->d_ydesc) (void)hipFree(p->d_ydesc);
  if (p->ywk) (void)hipFree(p->ywk);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->capturing) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(p->cap_stream, &g);
    if (g) (void)hipGraphDestroy(g);
  }
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  delete p;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_set_stream(d2dhip_plan p, void *stream) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (p->capturing)
    p->cap_keep = (hipStream_t)stream;  // takes effect when the capture ends
  else
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
  if (p && p->capturing && enable)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "plan is capturing: no stage profiling");
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  p->prof = enable ? 1 : 0;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_profile_report(d2dhip_plan p, double ms[7], int launches[7]) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  constexpr int ns = d2dhip_plan_s::kFftStages;
  for (int i = 0; i < ns; ++i) {
    if (ms) ms[i] = 0.0;
    if (launches) launches[i] = 0;
  }
  int st = D2DHIP_SUCCESS;
  double wsum[ns] = {};
  for (auto &r : p->recs) {
    if (!st) st = hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
    float t = 0.f;
    if (!st) st = hip_check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
    if (!st) {
      // real-to-real stages count with the other stages along their axis
      const int slot = r.stage < ns ? r.stage : p->geom[r.stage].axis;
      if (ms) ms[slot] += t;
      if (launches) wsum[slot] += r.w;
    }
    p->pool.push_back(r.a);
    p->pool.push_back(r.b);
  }
  p->recs.clear();
  if (launches)
    for (int i = 0; i < ns; ++i) launches[i] = (int)std::lround(wsum[i]);
  return st;
}

// Stages 7-12 exist once the plan has real-to-real transforms.
static int check_stage(const d2dhip_plan_s *p, int stage) {
  if (stage < 0 || stage >= d2dhip_plan_s::kStages ||
      (stage >= d2dhip_plan_s::kFftStages && !p->has_dtt))
    return set_error(D2DHIP_ERR_INVALID_VALUE, "stage out of range");
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_stage_kernel(d2dhip_plan p, int stage, char *name, int len) {
  if (!p || !name || len <= 0) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (int st = check_stage(p, stage)) return st;
  const std::string s = choice_name(p, stage);
  std::snprintf(name, (size_t)len, "%s", s.c_str());
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_stage_candidate(d2dhip_plan p, int stage, int idx, char *name, int len) {
  if (!p || !name || len <= 0) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (int st = check_stage(p, stage)) return st;
  const StageGeom &g = p->geom[stage];
  if (idx < 0 || idx > (int)g.cands.size())
    return set_error(D2DHIP_ERR_INVALID_VALUE, "candidate out of range");
  std::snprintf(name, (size_t)len, "%s", idx == 0 ? "general" : g.cands[idx - 1]->name);
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_set_stage_kernel(d2dhip_plan p, int stage, const char *name) {
  if (!p || !name) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (int st = check_stage(p, stage)) return st;
  if (p->capturing)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "plan is capturing: its kernels are fixed");
  const StageGeom &g = p->geom[stage];
  int c = -2;
  if (std::strcmp(name, "general") == 0) c = -1;
  for (size_t k = 0; k < g.cands.size(); ++k)
    if (std::strcmp(name, g.cands[k]->name) == 0) c = (int)k;
  if (c == -2) return set_error(D2DHIP_ERR_INVALID_VALUE, std::string("no kernel ") + name);
  int st = configure(p, stage, c);
  if (!st) st = upload_descs(p);
  return st;
}

int d2dhip_plan_set_chunking(d2dhip_plan p, int fwd_planes, int bwd_planes) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (fwd_planes < 0 || bwd_planes < 0)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "chunk size must be >= 0");
  if (p->format != D2DHIP_PHYSICAL_IN_X && (fwd_planes || bwd_planes))
    return set_error(D2DHIP_ERR_UNSUPPORTED, "plane chunking is for PHYSICAL_IN_X plans");
  if (p->capturing)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "plan is capturing: its chunking is fixed");
  if (p->multi && (fwd_planes || bwd_planes))
    return set_error(D2DHIP_ERR_UNSUPPORTED, "plane chunking is for whole-grid plans");
  p->chunk_fwd = fwd_planes >= p->nz ? 0 : fwd_planes;
  p->chunk_bwd = bwd_planes >= p->nz ? 0 : bwd_planes;
  return upload_chunk_descs(p);
}

int d2dhip_plan_set_layout(d2dhip_plan p, int layout) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  return set_layout(p, layout);
}

int d2dhip_plan_get_layout(d2dhip_plan p, int *layout) {
  if (!p || !layout) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  *layout = p->layout;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_get_chunking(d2dhip_plan p, int *fwd_planes, int *bwd_planes) {
  if (!p || !fwd_planes || !bwd_planes) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  *fwd_planes = p->chunk_fwd;
  *bwd_planes = p->chunk_bwd;
  return D2DHIP_SUCCESS;
}

long long d2dhip_plan_device_bytes(d2dhip_plan p) {
  if (!p) return 0;
  long long b = (long long)p->wk_bytes;
  for (auto &kv : p->tw) b += (long long)kv.second.bytes;
  for (auto &kv : p->dtw) b += (long long)std::abs(kv.first) * 2 * p->rb;
  return b;
}

// fft_3d_c2c: src/fft_common_3d.f90:9-116.  X-forward and Z-backward run
// x, y, z; X-backward and Z-forward run z, y, x.  Skipped axes
// (opt_skip_XYZ_c2c) are left out; the first stage that runs reads `in`.
int d2dhip_exec_c2c(d2dhip_plan p, void *in, void *out, int isign) {
  Region region(PROF_FFT, "fft_c2c");
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (p->multi)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "plan of one rank of a process grid: the 3D transform needs the exchange "
                     "between its stages (use d2dhip_exec_1m)");
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
  Region region(PROF_FFT, "fft_r2c");
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (p->multi)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "plan of one rank of a process grid: the 3D transform needs the exchange "
                     "between its stages (use d2dhip_exec_1m)");
  if (!in_r || !out_c) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  if (blocked_now(p, p->chunk_fwd, p->wk)) return run_blocked_r2c(p, in_r, out_c);
  const bool fx = p->format == D2DHIP_PHYSICAL_IN_X;
  // the r2c axis cannot be skipped (src/fft_generic.f90:222-223, :265-266)
  const int warn = (p->flags & (fx ? D2DHIP_FLAG_SKIP_X : D2DHIP_FLAG_SKIP_Z)) ? 1 : 0;
  const unsigned skip_last = fx ? D2DHIP_FLAG_SKIP_Z : D2DHIP_FLAG_SKIP_X;
  int st;
  if (fx && p->chunk_fwd > 0 && !(p->flags & D2DHIP_FLAG_SKIP_Y)) {
    if ((st = run_pair_chunked(p, p->chunk_fwd, 0, 3, 1, 4, in_r, out_c, out_c, 0))) return st;
  } else {
    if ((st = run(p, 3, in_r, out_c, 0))) return st;
    if (!(p->flags & D2DHIP_FLAG_SKIP_Y) && (st = run(p, 4, out_c, out_c, 0))) return st;
  }
  if (!(p->flags & skip_last) && (st = run(p, 5, out_c, out_c, 0))) return st;
  if (warn)
    return set_error(D2DHIP_WARN_SKIP_IGNORED, "r2c / c2r transform can not be skipped");
  return D2DHIP_SUCCESS;
}

// fft_3d_c2r: src/fft_common_3d.f90:194-296.  The c2c stages write the
// plan's workspace (or in_c itself with D2DHIP_FLAG_INPLACE), so that in_c
// is preserved as the reference's wk1 copy preserves it, without the copy.
namespace {

// The c2r stages with the spectral workspace wk (the plan's own, in_c itself,
// or the library's static scratch array).
int c2r_stages(d2dhip_plan_s *p, void *in_c, void *out_r, void *wk) {
  if (blocked_now(p, p->chunk_bwd, wk)) return run_blocked_c2r(p, in_c, out_r);
  const bool fx = p->format == D2DHIP_PHYSICAL_IN_X;
  const int warn = (p->flags & (fx ? D2DHIP_FLAG_SKIP_X : D2DHIP_FLAG_SKIP_Z)) ? 1 : 0;
  const unsigned skip_first = fx ? D2DHIP_FLAG_SKIP_Z : D2DHIP_FLAG_SKIP_X;
  const void *src = in_c;
  int st;
  if (!(p->flags & skip_first)) {  // z (X format) or x (Z format), inverse
    if ((st = run(p, 5, src, wk, 1))) return st;
    src = wk;
  }
  if (fx && p->chunk_bwd > 0 && !(p->flags & D2DHIP_FLAG_SKIP_Y)) {
    if ((st = run_pair_chunked(p, p->chunk_bwd, 2, 4, 3, 6, src, wk, out_r, 1))) return st;
  } else {
    if (!(p->flags & D2DHIP_FLAG_SKIP_Y)) {
      if ((st = run(p, 4, src, wk, 1))) return st;
      src = wk;
    }
    if ((st = run(p, 6, src, out_r, 1))) return st;
  }
  if (warn)
    return set_error(D2DHIP_WARN_SKIP_IGNORED, "r2c / c2r transform can not be skipped");
  return D2DHIP_SUCCESS;
}

// Workspace of d2dhip_exec_c2r_noalloc: one array per GPU, living as long as
// the process.  Host code never passes `scratch` itself to a kernel (on the
// host it is only a stand-in for the symbol); it uses the device address
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[D2DHIP_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device.
int scratch_address(void **ptr) {
  int dev = 0;
  int st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st) return st;
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *a = nullptr;
    st = hip_check(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (st) return st;
    if (!a) return set_error(D2DHIP_ERR_HIP, "hipGetSymbolAddress(scratch) gave NULL");
    g_scratch_ptr[dev] = a;
  }
  *ptr = g_scratch_ptr[dev];
  return D2DHIP_SUCCESS;
}

}  // namespace

int d2dhip_exec_c2r(d2dhip_plan p, void *in_c, void *out_r) {
  Region region(PROF_FFT, "fft_c2r");
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (p->multi)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "plan of one rank of a process grid: the 3D transform needs the exchange "
                     "between its stages (use d2dhip_exec_1m)");
  if (!in_c || !out_r) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  return c2r_stages(p, in_c, out_r, (p->flags & D2DHIP_FLAG_INPLACE) ? in_c : p->wk);
}

// d2dhip_exec_c2r with the workspace in the library's static scratch array:
// no allocation of any kind, in_c preserved whatever the plan's flags.
int d2dhip_exec_c2r_noalloc(d2dhip_plan p, void *in_c, void *out_r) {
  Region region(PROF_FFT, "fft_c2r");
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (p->multi)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "plan of one rank of a process grid: the 3D transform needs the exchange "
                     "between its stages (use d2dhip_exec_1m)");
  if (!in_c || !out_r) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  if (p->prof)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "stage profiling records events: not with noalloc");
  const long long nsp = (p->format == D2DHIP_PHYSICAL_IN_X)
                            ? (long long)(p->nx / 2 + 1) * p->ny * p->nz
                            : (long long)p->nx * p->ny * (p->nz / 2 + 1);
  const unsigned long long need = (unsigned long long)nsp * 2 * p->rb;
  if (need > (unsigned long long)D2DHIP_SCRATCH_BYTES)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "spectral pencil of " + std::to_string(need) + " bytes exceeds the " +
                         std::to_string((long long)D2DHIP_SCRATCH_BYTES) +
                         "-byte scratch array of d2dhip_exec_c2r_noalloc");
  void *wk = nullptr;
  const int st = scratch_address(&wk);
  if (st) return st;
  return c2r_stages(p, in_c, out_r, wk);
}

// decomp_2d_fft_engine_dtt_init (src/fft_fftw3_f03.f90:341-390) with
// dtt_assign_default (:394-420) and dtt_invert (:424-444): 3 types, or 12
// values (types, ifirst, ndismiss, ofirst; 1-based), per axis x, y, z.
int d2dhip_plan_set_dtt(d2dhip_plan p, int ndtt, const int *in_dtt) {
  if (!p || !in_dtt) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (p->multi)
    return set_error(D2DHIP_ERR_UNSUPPORTED,
                     "plan of one rank of a process grid: the 3D transform needs the exchange "
                     "between its stages (use d2dhip_exec_1m)");
  if (p->capturing)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "plan is capturing: DTT set-up allocates");
  if (ndtt != 3 && ndtt != 12) return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid argument");
  int v[15] = {};
  for (int k = 0; k < 3; ++k)
    if (in_dtt[k] < 1 || in_dtt[k] > 8)
      return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid argument: DTT types are 1 to 8");
  if (ndtt == 12) {
    for (int k = 0; k < 12; ++k) v[k] = in_dtt[k];
    for (int k = 0; k < 3; ++k)
      if (v[3 + k] != v[9 + k])
        return set_error(D2DHIP_ERR_INVALID_VALUE,
                         "Setup is not compatible with in-place r2r transforms");
  } else {
    for (int k = 0; k < 3; ++k) {
      v[k] = in_dtt[k];
      v[3 + k] = (v[k] == 5 || v[k] == 7) ? 2 : 1;  // ifirst: DST-I and DST-III start at 2
      v[6 + k] = v[k] == 1 ? 0 : (v[k] == 5 ? 2 : 1);  // ndismiss
      v[9 + k] = v[3 + k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    const int t = v[k];
    v[12 + k] = t == 2 ? 3 : t == 3 ? 2 : t == 6 ? 7 : t == 7 ? 6 : t;
  }
  const int npts[3] = {p->nx, p->ny, p->nz};
  for (int k = 0; k < 3; ++k) {
    const int len = npts[k] - v[6 + k], first = v[3 + k] - 1;
    if (len < (v[k] == 1 ? 2 : 1) || first < 0 || first + len > npts[k])
      return set_error(D2DHIP_ERR_INVALID_VALUE,
                       "DTT points out of range along axis " + std::to_string(k));
  }
  int st = D2DHIP_SUCCESS;
  for (int k = 0; k < 3 && !st; ++k) {
    const int len = npts[k] - v[6 + k];
    st = make_dtt_stage(p, 7 + k, k, v[k], v[3 + k] - 1, len);
    if (!st) st = make_dtt_stage(p, 10 + k, k, v[12 + k], v[9 + k] - 1, len);
  }
  if (!st) st = upload_descs(p);
  // plans made to be tuned time the real-to-real stages' kernels too
  if (!st && ((p->flags & D2DHIP_FLAG_TUNE) || std::getenv("D2DHIP_TUNE"))) st = tune_stages(p);
  p->has_dtt = !st;
  if (!st) std::memcpy(p->dtt, v, sizeof(v));
  if (!st && p->tune_write_failed)
    return set_error(D2DHIP_WARN_TUNE_CACHE, "tuning choices could not be written to " + tune_path());
  return st;
}

// engine%dtt after set-up, with the transform types as 1-8 (the reference
// keeps FFTW's kinds there; decomp_2d_fft_get_dtt_input,
// src/fft_fftw3_f03.f90:629-643, converts them back to these numbers).
int d2dhip_plan_get_dtt(d2dhip_plan p, int dtt[15]) {
  if (!p || !dtt) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (!p->has_dtt) return set_error(D2DHIP_ERR_INVALID_VALUE, "No DTT for the current engine");
  std::memcpy(dtt, p->dtt, sizeof(p->dtt));
  return D2DHIP_SUCCESS;
}

// decomp_2d_dtt_3d_r2r (src/fft_fftw3_f03.f90:1959-2023): PHYSICAL_IN_X
// forward and PHYSICAL_IN_Z backward run x, y, z, the others z, y, x; the
// backward direction runs the inverse types over the ofirst points.  The
// first stage writes `out` and the others work in place on it, so `in` is
// preserved (the reference transforms it in place along the first axis).
// Skipped axes (opt_skip_XYZ_c2c) are left out as r2r_1m_x/y/z leave them.
int d2dhip_exec_dtt(d2dhip_plan p, const void *in, void *out, int isign) {
  Region region(PROF_FFT, "decomp_2d_dtt_3d_r2r");
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (!in || !out) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  if (isign != D2DHIP_FORWARD && isign != D2DHIP_BACKWARD)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value");
  if (!p->has_dtt) return set_error(D2DHIP_ERR_PLAN_MISMATCH, "No DTT for the current engine");
  const bool x_first = (p->format == D2DHIP_PHYSICAL_IN_X && isign == D2DHIP_FORWARD) ||
                       (p->format == D2DHIP_PHYSICAL_IN_Z && isign == D2DHIP_BACKWARD);
  const int order[3] = {x_first ? 0 : 2, 1, x_first ? 2 : 0};
  const unsigned skipbit[3] = {D2DHIP_FLAG_SKIP_X, D2DHIP_FLAG_SKIP_Y, D2DHIP_FLAG_SKIP_Z};
  const int base = isign == D2DHIP_FORWARD ? 7 : 10;
  const void *src = in;
  int ran = 0;
  for (int s = 0; s < 3; ++s) {
    const int ax = order[s];
    if (p->flags & skipbit[ax]) continue;
    int st = run(p, base + ax, src, out, 0);
    if (st) return st;
    src = out;
    ++ran;
  }
  if (!ran && in != out) {
    const size_t bytes = (size_t)p->nx * p->ny * p->nz * p->rb;
    return hip_check(hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, p->stream),
                     "hipMemcpyAsync");
  }
  return D2DHIP_SUCCESS;
}

namespace {

// The x stage (idx 3: r2c, idx 6: c2r) of a PHYSICAL_IN_X plan with its
// complex side in the x-pencil send / receive-buffer order of process grid
// d: blocks of x1dist points at x1disp (or m * x1count with even), the
// mem_split_xy / mem_merge_yx maps (src/transpose_x_to_y.f90:266-330,
// src/transpose_y_to_x.f90:384-441) applied by the stage's stores / loads.
int run_x_stage_packed(d2dhip_plan_s *p, int idx, const void *in, void *out,
                       const d2dhip_decomp *d, int even) {
  if (p->capturing)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "fused pack / unpack calls are not captured into a graph");
  if (p->format != D2DHIP_PHYSICAL_IN_X)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "fused pack / unpack needs a PHYSICAL_IN_X plan");
  const long long nxc = p->nx / 2 + 1;
  if (d->xsz[0] != p->pen[3][0] || d->xsz[1] != p->pen[3][1] || d->xsz[2] != p->pen[3][2])
    return set_error(D2DHIP_ERR_INVALID_VALUE, "decomposition x-pencil is not the plan's spectral x-pencil");
  const int np = d->p_row;
  if (np < 1 || np > D2DHIP_MAX_PROCS) return set_error(D2DHIP_ERR_INVALID_VALUE, "p_row out of range");
  long long b1 = 0;
  for (int m = 0; m < np; ++m) {
    if (d->x1dist[m] < 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "negative block size");
    b1 += d->x1dist[m];
  }
  if (b1 != nxc) return set_error(D2DHIP_ERR_INVALID_VALUE, "decomposition does not match the pencil");
  // {base - b1, w} of point i's block, for every complex point i of a line
  p->h_pack.assign(2 * (size_t)nxc, 0);
  b1 = 0;
  for (int m = 0; m < np; ++m) {
    const long long base = even ? (long long)m * d->x1count : d->x1disp[m];
    for (long long i = b1; i < b1 + d->x1dist[m]; ++i) {
      p->h_pack[2 * i] = base - b1;
      p->h_pack[2 * i + 1] = d->x1dist[m];
    }
    b1 += d->x1dist[m];
  }
  if (p->pack_cap < p->h_pack.size()) {
    if (p->d_pack) (void)hipFree(p->d_pack);
    p->d_pack = nullptr;
    p->pack_cap = 0;
    if (hipMalloc(&p->d_pack, p->h_pack.size() * sizeof(long long)) != hipSuccess) {
      p->d_pack = nullptr;
      return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(pack table)");
    }
    p->pack_cap = p->h_pack.size();
  }
  if (!p->d_pdesc && hipMalloc(&p->d_pdesc, sizeof(AxisDesc)) != hipSuccess) {
    p->d_pdesc = nullptr;
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(pack descriptor)");
  }
  p->h_pdesc = p->h_desc[idx];
  p->h_pdesc.pack = p->d_pack;
  p->h_pdesc.npack = np;
  p->h_pdesc.pack_l1 = p->h_pdesc.C0r;  // lines j + k * ny of the x-pencil
  int st = upload_sync(p, p->d_pack, p->h_pack.data(), sizeof(long long) * p->h_pack.size(),
                       "hipMemcpy(pack table)");
  if (!st) st = upload_sync(p, p->d_pdesc, &p->h_pdesc, sizeof(AxisDesc), "hipMemcpy(pack descriptor)");
  if (st) return st;
  const AxisIo io = {in, out, 0};
  return p->fast[idx] ? p->fast[idx]->launch(p->d_pdesc, p->h_pdesc, io, p->stream)
                      : launch_axis(p->rb, p->geom[idx].io, p->d_pdesc, p->h_pdesc, io,
                                    p->shape[idx], p->stream);
}

// The y stage of a plan (1: on the physical y-pencil, 4: on the spectral
// one) with its output in the send order of exchange dir (D2DHIP_Y_TO_Z:
// mem_split_yz, src/transpose_y_to_z.f90:370-428; D2DHIP_Y_TO_X:
// mem_split_yx, src/transpose_y_to_x.f90:266-323) or its input in the
// receive order of dir (D2DHIP_Z_TO_Y: mem_merge_zy,
// src/transpose_z_to_y.f90:489-547; D2DHIP_X_TO_Y: mem_merge_xy,
// src/transpose_x_to_y.f90:384-441), and the z stages likewise with
// mem_split_zy / mem_merge_yz.  All keep natural order inside blocks of the
// stage's own axis (disp_m = m * count under EVEN).  The stage kernel's
// fused form (" ypk" / " yun" / " yup") or the general kernel applies the
// map in its stores / loads; a specialised kernel without a fused form runs
// unfused through a pencil workspace.  Results equal the stage and
// d2dhip_pack / d2dhip_unpack bit for bit.
//
// stage_table: the per-point table ({dj, dc1} for every point of the
// stage's axis) of map `dir` on the output (pack) or input (unpack) side of
// stage `stage`.
// y stages: the four y-pencil maps; z stages (c2c_1m_z on the physical
// z-pencil, and the spectral one under PHYSICAL_IN_X): mem_split_zy after,
// mem_merge_yz before.  Every one of these maps cuts its pencil into blocks
// along the stage's own axis.
int stage_table(const d2dhip_plan_s *p, int stage, int dir, const d2dhip_decomp *d, int even,
                bool pack, std::vector<long long> &tab) {
  const bool ystage = stage == 1 || stage == 4;
  const bool zstage = stage == 2 || (stage == 5 && p->format == D2DHIP_PHYSICAL_IN_X);
  if (!ystage && !zstage)
    return set_error(D2DHIP_ERR_INVALID_VALUE,
                     "fused pack / unpack: a y stage (1, 4) or a z stage (2; 5 under PHYSICAL_IN_X)");
  const bool two = ystage ? (pack ? dir == D2DHIP_Y_TO_Z : dir == D2DHIP_Z_TO_Y)
                          : (pack ? dir == D2DHIP_Z_TO_Y : dir == D2DHIP_Y_TO_Z);
  const bool one = ystage && (pack ? dir == D2DHIP_Y_TO_X : dir == D2DHIP_X_TO_Y);
  if (!two && !one)
    return set_error(D2DHIP_ERR_INVALID_VALUE,
                     ystage ? (pack ? "pack after a y stage: D2DHIP_Y_TO_Z or D2DHIP_Y_TO_X"
                                    : "unpack before a y stage: D2DHIP_Z_TO_Y or D2DHIP_X_TO_Y")
                            : (pack ? "pack after a z stage: D2DHIP_Z_TO_Y"
                                    : "unpack before a z stage: D2DHIP_Y_TO_Z"));
  const auto &pen = p->pen[stage];
  const int *dsz = ystage ? d->ysz : d->zsz;
  if (dsz[0] != pen[0] || dsz[1] != pen[1] || dsz[2] != pen[2])
    return set_error(D2DHIP_ERR_INVALID_VALUE,
                     ystage ? "decomposition y-pencil is not the stage's pencil"
                            : "decomposition z-pencil is not the stage's pencil");
  const int np = two ? d->p_col : d->p_row;
  if (np < 1 || np > D2DHIP_MAX_PROCS) return set_error(D2DHIP_ERR_INVALID_VALUE, "process count out of range");
  // block m of the stage's axis (b1 <= j < b1 + w) of the pencil starts at
  // disp_m in the buffer, and its point (i, j, k) sits at
  //   y-pencil: disp_m + i + (j - b1) n1 + k w n1      (blocks along j)
  //   z-pencil: disp_m + i + j n1 + (k - b1) n1 n2     (blocks along k)
  // i.e. at the plain pencil offset plus table[2] + c1 table[1] (c1 = k for
  // a y stage, 0 for a z stage, whose columns cover the (i, j) plane)
  const int *dist = ystage ? (two ? d->y2dist : d->y1dist) : d->z2dist;
  const long long *disp = ystage ? (two ? d->y2disp : d->y1disp) : d->z2disp;
  const long long count = ystage ? (two ? d->y2count : d->y1count) : d->z2count;
  const long long n1 = pen[0], naxis = ystage ? pen[1] : pen[2];
  const long long row = ystage ? n1 : n1 * pen[1];
  long long b1 = 0;
  for (int m = 0; m < np; ++m) {
    if (dist[m] < 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "negative block size");
    b1 += dist[m];
  }
  if (b1 != naxis) return set_error(D2DHIP_ERR_INVALID_VALUE, "decomposition does not match the pencil");
  // one entry {dj, dc1} per point j of the stage's axis
  tab.assign(2 * (size_t)naxis, 0);
  b1 = 0;
  for (int m = 0; m < np; ++m) {
    const long long dj = (even ? (long long)m * count : disp[m]) - b1 * row;
    const long long dc1 = ystage ? n1 * (dist[m] - naxis) : 0;
    for (long long j = b1; j < b1 + dist[m]; ++j) {
      tab[2 * j] = dj;
      tab[2 * j + 1] = dc1;
    }
    b1 += dist[m];
  }
  return D2DHIP_SUCCESS;
}

// A y or z stage with its input in the receive order of `dir_in` (-1: plain)
// and its output in the send order of `dir_out` (-1: plain): packj 2, 1 or
// (both) 3.
int run_stage_packed(d2dhip_plan_s *p, int stage, const void *in, void *out, int dir_in,
                     int dir_out, const d2dhip_decomp *d, int even, int isign) {
  if (p->capturing)
    return set_error(D2DHIP_ERR_UNSUPPORTED, "fused pack / unpack calls are not captured into a graph");
  if (isign != D2DHIP_FORWARD && isign != D2DHIP_BACKWARD)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value of isign");
  const bool unpack = dir_in >= 0, pack = dir_out >= 0;
  int st;
  if (unpack && (st = stage_table(p, stage, dir_in, d, even, false, p->h_ypack))) return st;
  if (pack && (st = stage_table(p, stage, dir_out, d, even, true, unpack ? p->h_ypack2 : p->h_ypack)))
    return st;
  const int packj = unpack ? (pack ? 3 : 2) : 1;
  const int conj = isign == D2DHIP_BACKWARD ? 1 : 0;
  const int esize = 2 * p->rb;
  const auto &pen = p->pen[stage];
  // the specialised kernel's fused form, if it has one
  const int io = packj == 1 ? IO_YPK : packj == 2 ? IO_YUN : IO_YUP;
  const FastVariant *v = p->fast[stage], *twin = nullptr;
  if (v) {
    std::string want = v->name;
    const size_t at = want.find(" c2c");
    if (at != std::string::npos) {
      want.replace(at, 4, packj == 1 ? " ypk" : packj == 2 ? " yun" : " yup");
      for (const FastVariant &c : fast_registry())
        if (c.rb == p->rb && c.io == io && want == c.name) twin = &c;
    }
  }
  if (v && !twin) {  // unpack, stage, pack through a workspace
    const size_t need = (size_t)pencil_elems(pen) * esize;
    if (p->ywk_bytes < need) {
      if (p->ywk) (void)hipFree(p->ywk);
      p->ywk = nullptr;
      p->ywk_bytes = 0;
      if (hipMalloc(&p->ywk, need) != hipSuccess) return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(stage workspace)");
      p->ywk_bytes = need;
    }
    const void *src = in;
    if (unpack) {
      if ((st = launch_pack(dir_in, 0, esize, in, p->ywk, d, even, p->stream))) return st;
      src = p->ywk;
    }
    if (!pack) return run(p, stage, src, out, conj);
    if ((st = run(p, stage, src, p->ywk, conj))) return st;
    return launch_pack(dir_out, 1, esize, p->ywk, out, d, even, p->stream);
  }
  auto fit = [&](long long *&buf, size_t &cap, size_t n) {
    if (cap >= n) return D2DHIP_SUCCESS;
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    if (hipMalloc(&buf, n * sizeof(long long)) != hipSuccess) {
      buf = nullptr;
      return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(pack table)");
    }
    cap = n;
    return D2DHIP_SUCCESS;
  };
  if ((st = fit(p->d_ypack, p->ypack_cap, p->h_ypack.size()))) return st;
  if (packj == 3 && (st = fit(p->d_ypack2, p->ypack2_cap, p->h_ypack2.size()))) return st;
  if (!p->d_ydesc && hipMalloc(&p->d_ydesc, sizeof(AxisDesc)) != hipSuccess) {
    p->d_ydesc = nullptr;
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(pack descriptor)");
  }
  p->h_ydesc = p->h_desc[stage];
  p->h_ydesc.pack = p->d_ypack;
  p->h_ydesc.npack = (int)(p->h_ypack.size() / 2);
  p->h_ydesc.pack2 = packj == 3 ? p->d_ypack2 : nullptr;
  p->h_ydesc.npack2 = packj == 3 ? (int)(p->h_ypack2.size() / 2) : 0;
  p->h_ydesc.packj = packj;
  st = upload_sync(p, p->d_ypack, p->h_ypack.data(), sizeof(long long) * p->h_ypack.size(),
                   "hipMemcpy(pack table)");
  if (!st && packj == 3)
    st = upload_sync(p, p->d_ypack2, p->h_ypack2.data(), sizeof(long long) * p->h_ypack2.size(),
                     "hipMemcpy(pack table)");
  if (!st) st = upload_sync(p, p->d_ydesc, &p->h_ydesc, sizeof(AxisDesc), "hipMemcpy(pack descriptor)");
  if (st) return st;
  const AxisIo aio = {in, out, conj};
  if (twin) {
    if ((st = hip_check(twin->prepare(), "hipFuncSetAttribute(fast kernel)"))) return st;
    return twin->launch(p->d_ydesc, p->h_ydesc, aio, p->stream);
  }
  return launch_axis(p->rb, IO_C2C, p->d_ydesc, p->h_ydesc, aio, p->shape[stage], p->stream);
}

}  // namespace

// One 1D stage of a plan on its pencil: c2c_1m_x / _y / _z, r2c_1m_x / _z,
// c2r_1m_x / _z (src/fft_generic.f90:112-384, FFTW plans
// src/fft_fftw3_f03.f90:1414-1552).
int d2dhip_exec_1m(d2dhip_plan p, int stage, const void *in, void *out, int isign) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (!in || !out) return set_error(D2DHIP_ERR_NULL, "NULL data pointer");
  if (stage < 0 || stage >= d2dhip_plan_s::kFftStages)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "stage out of range (0-6)");
  if (isign != D2DHIP_FORWARD && isign != D2DHIP_BACKWARD)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "Invalid value of isign");
  if ((stage == 3 && isign != D2DHIP_FORWARD) || (stage == 6 && isign != D2DHIP_BACKWARD))
    return set_error(D2DHIP_ERR_INVALID_VALUE, "r2c is forward, c2r backward");
  if ((stage == 3 || stage == 6) && in == out)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "r2c / c2r stages are not in place");
  const int conj = stage == 6 ? 1 : (stage == 3 ? 0 : (isign == D2DHIP_BACKWARD ? 1 : 0));
  return run(p, stage, in, out, conj);
}

int d2dhip_exec_1m_pack(d2dhip_plan p, int stage, const void *in, void *send, int dir,
                        const d2dhip_decomp *d, int even, int isign) {
  if (!p || !in || !send || !d) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  return run_stage_packed(p, stage, in, send, -1, dir, d, even, isign);
}

int d2dhip_exec_1m_unpack(d2dhip_plan p, int stage, const void *recv, void *out, int dir,
                          const d2dhip_decomp *d, int even, int isign) {
  if (!p || !recv || !out || !d) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  return run_stage_packed(p, stage, recv, out, dir, -1, d, even, isign);
}

int d2dhip_exec_1m_unpack_pack(d2dhip_plan p, int stage, const void *recv, void *send,
                               int dir_in, int dir_out, const d2dhip_decomp *d, int even,
                               int isign) {
  if (!p || !recv || !send || !d) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (stage != 1 && stage != 4)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "unpack and pack around one stage: a y stage (1, 4)");
  if (dir_in < 0 || dir_out < 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "bad exchange direction");
  return run_stage_packed(p, stage, recv, send, dir_in, dir_out, d, even, isign);
}

// ---- hipGraph of a plan's exec calls --------------------------------------------
// Between capture_begin and capture_end the plan launches on a capturing
// stream of its own, so that every exec call (with the buffers given to it)
// is recorded; graph_launch replays the recording on the plan's stream.
// The exec paths issue only kernels and device-to-device copies, no memset
// and no host-synchronous call, so a replay repeats them exactly.
int d2dhip_plan_capture_begin(d2dhip_plan p) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (p->capturing) return set_error(D2DHIP_ERR_INVALID_VALUE, "plan is already capturing");
  if (p->prof) return set_error(D2DHIP_ERR_INVALID_VALUE, "stage profiling is on: no capture");
  if (!p->cap_stream) {
    const int st = hip_check(hipStreamCreate(&p->cap_stream), "hipStreamCreate");
    if (st) return st;
  }
  // the capture starts after the work already queued on the plan's stream
  hipEvent_t e = nullptr;
  int st = hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  if (!st) st = hip_check(hipEventRecord(e, p->stream), "hipEventRecord");
  if (!st) st = hip_check(hipStreamWaitEvent(p->cap_stream, e, 0), "hipStreamWaitEvent");
  if (e) (void)hipEventDestroy(e);
  if (!st)
    st = hip_check(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal),
                   "hipStreamBeginCapture");
  if (st) return st;
  p->cap_keep = p->stream;
  p->stream = p->cap_stream;
  p->capturing = true;
  return D2DHIP_SUCCESS;
}

int d2dhip_plan_capture_end(d2dhip_plan p) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (!p->capturing) return set_error(D2DHIP_ERR_INVALID_VALUE, "plan is not capturing");
  hipGraph_t g = nullptr;
  int st = hip_check(hipStreamEndCapture(p->cap_stream, &g), "hipStreamEndCapture");
  p->stream = p->cap_keep;
  p->capturing = false;
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph = g;
  p->graph_exec = nullptr;
  return hip_check(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
}

int d2dhip_plan_graph_launch(d2dhip_plan p) {
  if (!p) return set_error(D2DHIP_ERR_NULL, "plan is NULL");
  if (!p->graph_exec) return set_error(D2DHIP_ERR_INVALID_VALUE, "no captured graph");
  return hip_check(hipGraphLaunch(p->graph_exec, p->stream), "hipGraphLaunch");
}

// The forward x stage of a PHYSICAL_IN_X plan (r2c along x) writing the
// send buffers of transpose_x_to_y: r2c_1m_x followed by mem_split_xy
// (src/fft_common_3d.f90:133-140, src/transpose_x_to_y.f90:93-101) in one
// pass over the data.
int d2dhip_exec_r2c_pack(d2dhip_plan p, const void *in_r, void *send, const d2dhip_decomp *d,
                         int even) {
  if (!p || !in_r || !send || !d) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  return run_x_stage_packed(p, 3, in_r, send, d, even);
}

// The inverse x stage of a PHYSICAL_IN_X plan (c2r along x) reading the
// receive buffers of transpose_y_to_x: mem_merge_yx followed by c2r_1m_x
// (src/transpose_y_to_x.f90:131-133, src/fft_common_3d.f90:259-266).
int d2dhip_exec_c2r_unpack(d2dhip_plan p, const void *recv, void *out_r, const d2dhip_decomp *d,
                           int even) {
  if (!p || !recv || !out_r || !d) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  return run_x_stage_packed(p, 6, recv, out_r, d, even);
}

}  // extern "C"

// ---- many grids in one launch -------------------------------------------------
// The fft_multiple_grids use (examples/fft_multiple_grids/fft_multiple_grids.f90):
// the c2c stages of many grids as stage phases, through one persistent
// kernel fed by device queues or through one launch per phase and kernel.

// Work per persistent queue item (complex points, whole tiles): larger items
// lengthen the wait at each phase boundary (the last items of a phase run
// while other workgroups wait), smaller ones add queue traffic.  Measured on
// 64 uneven FP64 grids with specialised tiles: 1024 to 4096 points within
// 2 %, 16384 points 5 % slower.
constexpr long long kBatchItemPoints = 2048;
// Default cluster budget (bytes of a cluster's input and output buffers):
// none, every grid's stage k before any grid's stage k+1.  Measured on 64
// uneven FP64 grids (32..256 per dimension, grouped executor): one cluster
// 5.09 ms, 192 MiB clusters 5.28 ms, 64 MiB 5.41 ms, one grid per cluster
// 5.56 ms -- fewer, larger launches gain more than the cache reuse.
constexpr long long kBatchClusterBytes = 0;
// Consecutive items a persistent queue takes at a time (neighbouring tiles
// of strided columns share cache lines, and one queue serves one XCD).
constexpr size_t kBatchQueueChunk = 4;

struct d2dhip_batch_s {
  int rb = 8;
  hipStream_t stream = nullptr;
  std::vector<d2dhip_plan_s *> plans;
  AxisDesc *d_descs = nullptr;   // 3 general-kernel c2c descriptors per plan (x, y, z)
  std::vector<AxisDesc> h_descs;
  // The persistent kernel on specialised tiles (when every c2c stage of
  // every plan has one): 3 descriptors and tile kinds per plan.
  bool fast = false;
  AxisDesc *d_fdescs = nullptr;
  std::vector<AxisDesc> h_fdescs;
  std::vector<int> kinds;
  size_t flds = 0;
  int fblocks = 0;
  BatchGroup *d_fgroups = nullptr;
  unsigned long long *d_fphase_items = nullptr;  // items of each phase and queue
  // the queues (general / specialised groups): entries and each queue's share
  QueueEntry *d_entries = nullptr, *d_fentries = nullptr;
  long long qoff[kBatchQueues] = {}, qlen[kBatchQueues] = {};
  long long fqoff[kBatchQueues] = {}, fqlen[kBatchQueues] = {};
  long long item_points = kBatchItemPoints;  // complex points per queue item
  BatchGroup *d_groups = nullptr;
  int ngroups = 0;
  unsigned long long *d_phase_items = nullptr;  // items of each phase and queue
  int nphase = 0;
  unsigned long long *d_head = nullptr;  // the queue counters
  unsigned long long *d_done = nullptr;  // completed items of each phase and queue
  unsigned int *d_err = nullptr;
  long long wait_ticks = 0;              // phase-wait limit in wall-clock ticks
  int nthreads = 64;
  size_t lds = 0;
  int blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool bound = false;
  int mode = D2DHIP_BATCH_PERSISTENT;
  // Grouped executor: per phase, the grids partitioned by the stage kernel
  // their own plans run (as of bind); each part is one launch of that kernel
  // over the tiles of its grids, with the plans' own descriptors.
  struct Part {
    const FastVariant *fv = nullptr;  // nullptr: the general kernel
    int nthreads = 64;
    size_t lds = 0;
    long long m0 = 0;                 // first entry in d_multi
    int ng = 0;
    long long tiles = 0;
  };
  std::vector<std::vector<Part>> parts;  // per phase
  MultiGroup *d_multi = nullptr;
  std::vector<const FastVariant *> bound_fast;  // the plans' c2c kernels at bind (3 per plan)
  // Clusters: consecutive grids whose buffers (in + out) total at most this
  // many bytes run all their phases before the next cluster starts, so that
  // a cluster's later stages find its data in the memory-side cache;
  // 0 = one cluster (phase-major over the whole batch).
  long long cluster_bytes = kBatchClusterBytes;
  int nclusters = 0;
};

namespace {

// Widest power-of-two tile (TX columns) of a general-kernel stage that needs
// at most nt threads and lds bytes of LDS, capped at the stage's column
// count: in a batch every workgroup has the batch's block size and LDS, so
// stages with short lines take more columns per tile instead of idling.
void reshape_for_batch(AxisDesc &a, bool lines, int cb, int nt, size_t lds) {
  const int ldt_lines = (a.n % 2 == 0) ? a.n + 1 : a.n;
  long long c0p2 = 1;
  while (c0p2 < a.C0 && c0p2 < 1024) c0p2 <<= 1;
  int best = 0;
  for (int tx = 1; tx <= c0p2; tx <<= 1) {
    int need = 64;
    for (int i = 0; i < a.npass; ++i) {
      const int r = a.pass[i].r, per = kRegCap / r;
      need = std::max<long long>(need, ((long long)tx * (a.n / r) + per - 1) / per);
    }
    const size_t l = lines ? (size_t)tx * ldt_lines * cb : (size_t)tx * a.n * cb;
    if (need > nt || l > lds) break;
    best = tx;
  }
  if (best == 0) return;  // not even one column fits: keep the planned tile
  a.TX = best;
  a.log2TX = ilog2(best);
  a.tiles0 = (int)((a.C0 + best - 1) / best);
  if (lines) {
    a.ldj = 1;
    a.ldt = ldt_lines;
  } else {
    a.ldj = best;
    a.ldt = 1;
  }
}

}  // namespace

extern "C" {

int d2dhip_batch_create(d2dhip_batch *out, int nplans, const d2dhip_plan *plans, void *stream) {
  if (!out || !plans) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  *out = nullptr;
  if (nplans < 1) return set_error(D2DHIP_ERR_INVALID_VALUE, "nplans < 1");
  d2dhip_batch_s *b = new (std::nothrow) d2dhip_batch_s();
  if (!b) return set_error(D2DHIP_ERR_ALLOC, "host allocation failed");
  b->stream = (hipStream_t)stream;
  int st = D2DHIP_SUCCESS;
  for (int i = 0; i < nplans && !st; ++i) {
    d2dhip_plan_s *p = plans[i];
    if (!p) st = set_error(D2DHIP_ERR_NULL, "NULL plan in batch");
    else if (p->multi) st = set_error(D2DHIP_ERR_UNSUPPORTED, "batches take whole-grid plans");
    else if (i > 0 && p->rb != b->rb) st = set_error(D2DHIP_ERR_INVALID_VALUE, "mixed precisions in batch");
    if (st) break;
    b->rb = p->rb;
    b->plans.push_back(p);
    for (int a = 0; a < 3 && !st; ++a) {
      // the general persistent kernel: configure the plan's stage for the
      // general kernel, copy the descriptor, restore the plan's choice
      const int keep = p->geom[a].choice;
      st = configure(p, a, -1);
      if (!st) {
        b->h_descs.push_back(p->h_desc[a]);
        if (p->h_desc[a].blue) st = set_error(D2DHIP_ERR_UNSUPPORTED, "Bluestein lengths are not batched");
      }
      if (!st) st = configure(p, a, keep);
    }
  }
  // the persistent kernel on specialised tiles: the plan's candidate with
  // the kind's kernel body, configured the same way
  b->fast = !st;
  const std::vector<PersistKind> &pk = persist_kinds(b->rb);
  for (size_t i = 0; i < b->plans.size() && !st && b->fast; ++i) {
    d2dhip_plan_s *p = b->plans[i];
    for (int a = 0; a < 3 && !st && b->fast; ++a) {
      const StageGeom &g = p->geom[a];
      int kind = -1, cand = -1;
      for (size_t k = 0; k < pk.size() && kind < 0; ++k)
        if (pk[k].n == g.n && pk[k].lines == (a == 0 ? 1 : 0)) kind = (int)k;
      if (kind >= 0)
        for (size_t c = 0; c < g.cands.size() && cand < 0; ++c)
          if (pk[kind].name == g.cands[c]->name && g.cands[c]->lines == pk[kind].lines) cand = (int)c;
      if (cand < 0) {
        b->fast = false;
        break;
      }
      const int keep = g.choice;
      st = configure(p, a, cand);
      if (!st) b->h_fdescs.push_back(p->h_desc[a]);
      if (!st) st = configure(p, a, keep);
      b->kinds.push_back(kind);
      b->flds = std::max(b->flds, pk[kind].lds_bytes);
    }
  }
  if (!st && b->fast) {
    const size_t bytes = sizeof(AxisDesc) * b->h_fdescs.size();
    if (hipMalloc(&b->d_fdescs, bytes) != hipSuccess) st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(batch)");
    if (!st) st = hip_check(hipMemcpy(b->d_fdescs, b->h_fdescs.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy");
    int bpc = 0, dev = 0, ncu = 0;
    if (!st) st = batch_fast_prepare(b->rb, b->flds, &bpc);
    if (!st) st = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (!st) st = hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    b->fblocks = std::max(1, bpc) * std::max(1, ncu);
  }
  if (!st) {
    // Block shape of the batch: the general kernel runs at most 8 waves per
    // CU (its register use), so 2-wave blocks with a quarter of the LDS keep
    // 4 independent tiles per CU; stages re-tile to fill that block.  A stage
    // whose single column needs more keeps its planned tile, and the batch
    // grows to fit it.
    int nt = 128;
    size_t lds = 40 * 1024;
    for (size_t i = 0; i < b->h_descs.size(); ++i) {
      AxisDesc &a = b->h_descs[i];
      reshape_for_batch(a, i % 3 == 0, 2 * b->rb, nt, lds);
    }
    b->nthreads = nt;
    b->lds = 0;
    for (size_t i = 0; i < b->h_descs.size(); ++i) {
      const AxisDesc &a = b->h_descs[i];
      const bool lines = (i % 3 == 0);
      const int ldt_lines = (a.n % 2 == 0) ? a.n + 1 : a.n;
      const size_t l = lines ? (size_t)a.TX * ldt_lines * 2 * b->rb : (size_t)a.TX * a.n * 2 * b->rb;
      int need = 64;
      for (int k = 0; k < a.npass; ++k) {
        const int r = a.pass[k].r, per = kRegCap / r;
        need = std::max<long long>(need, ((long long)a.TX * (a.n / r) + per - 1) / per);
      }
      b->nthreads = std::max(b->nthreads, (need + 63) / 64 * 64);
      b->lds = std::max(b->lds, l);
    }
    const size_t bytes = sizeof(AxisDesc) * b->h_descs.size();
    if (hipMalloc(&b->d_descs, bytes) != hipSuccess) st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(batch)");
    if (!st) st = hip_check(hipMemcpy(b->d_descs, b->h_descs.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy");
  }
  if (!st && b->lds > 160 * 1024 - 64)
    st = set_error(D2DHIP_ERR_UNSUPPORTED, "batched tile does not fit LDS");
  if (!st) {
    int bpc = 0, dev = 0, ncu = 0;
    st = batch_prepare(b->rb, b->nthreads, b->lds, &bpc);
    if (!st) st = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (!st) st = hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    b->blocks = std::max(1, bpc) * std::max(1, ncu);
    int khz = 0;  // s_memrealtime rate
    if (!st) st = hip_check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
    b->wait_ticks = 5000LL * (long long)std::max(khz, 1000);  // 5 s
  }
  if (!st && (hipMalloc(&b->d_head, sizeof(unsigned long long) * kBatchQueues * kHeadStride) != hipSuccess ||
              hipMalloc(&b->d_err, sizeof(unsigned int)) != hipSuccess))
    st = set_error(D2DHIP_ERR_ALLOC, "hipMalloc(batch counters)");
  // the error flag starts clear: only the persistent launches reset and set
  // it, and check reads it whatever the executor
  if (!st) {
    const unsigned int zero = 0;
    st = hip_check(hipMemcpy(b->d_err, &zero, sizeof(zero), hipMemcpyHostToDevice), "hipMemcpy(batch flag)");
  }
  if (st) {
    d2dhip_batch_destroy(b);
    return st;
  }
  *out = b;
  return D2DHIP_SUCCESS;
}

// Lays out the work queue for the given buffers.  mode -1 / +1: one
// direction; 0: forward then backward (out = unnormalised round trip).
int d2dhip_batch_bind_c2c(d2dhip_batch b, void *const *in, void *const *out, int mode) {
  if (!b || !in || !out) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  if (mode != -1 && mode != 0 && mode != 1) return set_error(D2DHIP_ERR_INVALID_VALUE, "mode");
  std::vector<BatchGroup> groups;
  std::vector<unsigned long long> phase_items;
  // phase-major order: stage 0 of every grid, then stage 1, ...; phases in
  // which every grid skips its axis are left out
  const int nph = mode == 0 ? 6 : 3;
  const int np = (int)b->plans.size();
  // clusters of consecutive grids within the byte budget (at least one grid)
  std::vector<int> cl_begin;
  {
    long long acc = 0;
    for (int g = 0; g < np; ++g) {
      const d2dhip_plan_s *p = b->plans[g];
      const long long pts = (long long)p->nx * p->ny * p->nz;
      const long long bytes = pts * 2 * b->rb * (in[g] == out[g] ? 1 : 2);
      if (g == 0 || (b->cluster_bytes > 0 && acc + bytes > b->cluster_bytes)) {
        cl_begin.push_back(g);
        acc = 0;
      }
      acc += bytes;
    }
    cl_begin.push_back(np);
  }
  const int ncl = (int)cl_begin.size() - 1;
  std::vector<char> started(np, 0);
  // the same groups for the persistent kernel on specialised tiles
  std::vector<BatchGroup> fgroups;
  std::vector<std::vector<d2dhip_batch_s::Part>> parts;
  std::vector<MultiGroup> multi;
  size_t general_lds = 0;
  for (int cph = 0; cph < ncl * nph; ++cph) {
    const int cl = cph / nph, ph = cph % nph;
    unsigned long long items = 0;
    // grouped executor: grids of this phase by their plan's stage kernel
    std::vector<const FastVariant *> keys;
    std::vector<std::vector<MultiGroup>> members;
    std::vector<d2dhip_batch_s::Part> pp;
    for (int g = cl_begin[cl]; g < cl_begin[cl + 1]; ++g) {
      d2dhip_plan_s *p = b->plans[g];
      const int isign = (mode == 0) ? (ph < 3 ? -1 : 1) : mode;
      const bool x_first = (p->format == D2DHIP_PHYSICAL_IN_X && isign == -1) ||
                           (p->format == D2DHIP_PHYSICAL_IN_Z && isign == 1);
      const int k = ph % 3;
      const int ax = (k == 1) ? 1 : ((k == 0) == x_first ? 0 : 2);
      const unsigned skipbit[3] = {D2DHIP_FLAG_SKIP_X, D2DHIP_FLAG_SKIP_Y, D2DHIP_FLAG_SKIP_Z};
      if (p->flags & skipbit[ax]) continue;
      // a queue group: all tiles of this stage of grid g, about
      // item_points complex points per queue item
      auto make_group = [&](const AxisDesc *dd, const AxisDesc &h, int kind) {
        BatchGroup G;
        G.d = dd;
        G.in = started[g] ? out[g] : in[g];
        G.out = out[g];
        G.conj = isign == 1 ? 1 : 0;
        G.phase = (int)phase_items.size();
        G.kind = kind;
        G.ntiles = (long long)h.tiles0 * h.C1;
        const long long tile_pts = (long long)h.TX * h.n;
        G.per_item = (int)std::max<long long>(1, std::min<long long>(64, b->item_points / tile_pts));
        // strided columns: at least one 128-byte line of each row per item
        if (h.ldt == 1) {
          const long long row_bytes = (long long)h.TX * 2 * b->rb;
          G.per_item = (int)std::max<long long>(G.per_item, (128 + row_bytes - 1) / row_bytes);
        }
        return G;
      };
      const BatchGroup G = make_group(b->d_descs + 3 * g + ax, b->h_descs[3 * g + ax], 0);
      groups.push_back(G);
      items += (unsigned long long)((G.ntiles + G.per_item - 1) / G.per_item);
      if (b->fast)
        fgroups.push_back(make_group(b->d_fdescs + 3 * g + ax, b->h_fdescs[3 * g + ax],
                                     b->kinds[3 * g + ax]));
      // the same stage as the plan itself runs it
      const FastVariant *fv = p->fast[ax];
      size_t q = 0;
      while (q < keys.size() && keys[q] != fv) ++q;
      if (q == keys.size()) {
        keys.push_back(fv);
        members.emplace_back();
        pp.emplace_back();
        pp[q].fv = fv;
      }
      const AxisDesc &ph_d = p->h_desc[ax];
      MultiGroup M;
      M.d = p->d_desc + ax;
      M.in = started[g] ? out[g] : in[g];
      M.out = out[g];
      M.conj = G.conj;
      M.tile0 = pp[q].tiles;
      members[q].push_back(M);
      pp[q].tiles += (long long)ph_d.tiles0 * ph_d.C1;
      pp[q].nthreads = std::max(pp[q].nthreads, p->shape[ax].nthreads);
      pp[q].lds = std::max(pp[q].lds, p->shape[ax].lds_bytes);
      started[g] = 1;
    }
    if (items) {
      phase_items.push_back(items);
      for (size_t k = 0; k < pp.size(); ++k) {
        pp[k].m0 = (long long)multi.size();
        pp[k].ng = (int)members[k].size();
        multi.insert(multi.end(), members[k].begin(), members[k].end());
        if (!pp[k].fv) general_lds = std::max(general_lds, pp[k].lds);
      }
      parts.push_back(pp);
    }
  }
  for (int g = 0; g < np; ++g)
    if (!started[g] && in[g] != out[g])
      return set_error(D2DHIP_ERR_UNSUPPORTED, "every axis skipped: nothing to batch");
  if (b->d_groups) (void)hipFree(b->d_groups);
  if (b->d_done) (void)hipFree(b->d_done);
  if (b->d_phase_items) (void)hipFree(b->d_phase_items);
  if (b->d_multi) (void)hipFree(b->d_multi);
  if (b->d_fgroups) (void)hipFree(b->d_fgroups);
  if (b->d_fphase_items) (void)hipFree(b->d_fphase_items);
  if (b->d_entries) (void)hipFree(b->d_entries);
  if (b->d_fentries) (void)hipFree(b->d_fentries);
  b->d_entries = nullptr;
  b->d_fentries = nullptr;
  b->d_groups = nullptr;
  b->d_done = nullptr;
  b->d_phase_items = nullptr;
  b->d_multi = nullptr;
  b->d_fgroups = nullptr;
  b->d_fphase_items = nullptr;
  b->parts = parts;
  b->nclusters = ncl;
  b->bound_fast.assign(3 * np, nullptr);
  for (int g = 0; g < np; ++g)
    for (int a = 0; a < 3; ++a) b->bound_fast[3 * g + a] = b->plans[g]->fast[a];
  b->ngroups = (int)groups.size();
  b->nphase = (int)phase_items.size();
  b->bound = false;
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  if (groups.empty()) {  // in == out everywhere: nothing to do
    b->bound = true;
    return D2DHIP_SUCCESS;
  }
  if (hipMalloc(&b->d_groups, sizeof(BatchGroup) * groups.size()) != hipSuccess ||
      hipMalloc(&b->d_done, sizeof(unsigned long long) * phase_items.size() * kBatchQueues *
                                kHeadStride) != hipSuccess ||
      hipMalloc(&b->d_multi, sizeof(MultiGroup) * multi.size()) != hipSuccess)
    return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(batch queue)");
  int st = hip_check(hipMemcpy(b->d_groups, groups.data(), sizeof(BatchGroup) * groups.size(),
                               hipMemcpyHostToDevice),
                     "hipMemcpy(batch queue)");
  if (!st)
    st = hip_check(hipMemcpy(b->d_multi, multi.data(), sizeof(MultiGroup) * multi.size(),
                             hipMemcpyHostToDevice),
                   "hipMemcpy(batch groups)");
  // the queues: each phase's items (contiguous, phase-major) dealt to the
  // kBatchQueues queues in chunks of kBatchQueueChunk, round robin
  auto build_queues = [&](const std::vector<BatchGroup> &gs, QueueEntry **dst, long long *off,
                          long long *len, unsigned long long **dphase) {
    std::vector<std::vector<QueueEntry>> qs(kBatchQueues);
    std::vector<unsigned long long> qphase;  // items of phase p in queue q
    size_t k = 0;
    while (k < gs.size()) {
      const int ph = gs[k].phase;
      std::vector<QueueEntry> items;
      for (; k < gs.size() && gs[k].phase == ph; ++k) {
        const long long n = (gs[k].ntiles + gs[k].per_item - 1) / gs[k].per_item;
        for (long long i = 0; i < n; ++i) items.push_back(QueueEntry{(int)k, (int)i, ph, 0});
      }
      const size_t qp0 = qphase.size();
      qphase.resize(qp0 + kBatchQueues, 0ull);
      for (size_t i = 0; i < items.size(); ++i) {
        const size_t q = (i / kBatchQueueChunk) % kBatchQueues;
        qs[q].push_back(items[i]);
        ++qphase[qp0 + q];
      }
    }
    std::vector<QueueEntry> all;
    for (int q = 0; q < kBatchQueues; ++q) {
      off[q] = (long long)all.size();
      len[q] = (long long)qs[q].size();
      all.insert(all.end(), qs[q].begin(), qs[q].end());
    }
    if (hipMalloc(dst, sizeof(QueueEntry) * std::max<size_t>(1, all.size())) != hipSuccess ||
        hipMalloc(dphase, sizeof(unsigned long long) * qphase.size()) != hipSuccess)
      return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(batch queue)");
    int e = hip_check(hipMemcpy(*dst, all.data(), sizeof(QueueEntry) * all.size(), hipMemcpyHostToDevice),
                      "hipMemcpy(batch queue)");
    if (!e)
      e = hip_check(hipMemcpy(*dphase, qphase.data(), sizeof(unsigned long long) * qphase.size(),
                              hipMemcpyHostToDevice),
                    "hipMemcpy(batch phases)");
    return e;
  };
  if (!st) st = build_queues(groups, &b->d_entries, b->qoff, b->qlen, &b->d_phase_items);
  if (!st && b->fast)
    st = build_queues(fgroups, &b->d_fentries, b->fqoff, b->fqlen, &b->d_fphase_items);
  if (!st && b->fast) {
    if (hipMalloc(&b->d_fgroups, sizeof(BatchGroup) * fgroups.size()) != hipSuccess)
      return set_error(D2DHIP_ERR_ALLOC, "hipMalloc(batch queue)");
    st = hip_check(hipMemcpy(b->d_fgroups, fgroups.data(), sizeof(BatchGroup) * fgroups.size(),
                             hipMemcpyHostToDevice),
                   "hipMemcpy(batch queue)");
  }
  // kernel attributes now, so that a captured launch makes no such call
  if (!st && general_lds) st = general_multi_prepare(b->rb, general_lds);
  for (size_t ph = 0; ph < parts.size() && !st; ++ph)
    for (const auto &q : parts[ph])
      if (!st && q.fv) st = hip_check(q.fv->prepare_multi(), "hipFuncSetAttribute(multi kernel)");
  b->bound = (st == D2DHIP_SUCCESS);
  return st;
}

int d2dhip_batch_launch(d2dhip_batch b) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  if (!b->bound) return set_error(D2DHIP_ERR_INVALID_VALUE, "batch not bound");
  if (b->ngroups == 0) return D2DHIP_SUCCESS;
  // the default executor runs grouped launches when some length has no
  // persistent tile: the general kernel's tiles in one persistent launch
  // are about 3x slower than the plans' own kernels
  if (b->mode == D2DHIP_BATCH_GROUPED || (b->mode == D2DHIP_BATCH_PERSISTENT && !b->fast)) {
    for (size_t g = 0; g < b->plans.size(); ++g)
      for (int a = 0; a < 3; ++a)
        if (b->plans[g]->fast[a] != b->bound_fast[3 * g + a])
          return set_error(D2DHIP_ERR_INVALID_VALUE, "a plan's stage kernel changed since bind: bind again");
    for (const auto &pp : b->parts)
      for (const auto &q : pp) {
        const MultiGroup *g = b->d_multi + q.m0;
        const int st = q.fv ? q.fv->launch_multi(g, q.ng, q.tiles, b->stream)
                            : launch_general_multi(b->rb, g, q.ng, q.tiles, q.nthreads, q.lds,
                                                   b->stream);
        if (st) return st;
      }
    return D2DHIP_SUCCESS;
  }
  const bool fast = b->mode == D2DHIP_BATCH_PERSISTENT && b->fast;
  BatchQueue q;
  q.groups = fast ? b->d_fgroups : b->d_groups;
  q.entries = fast ? b->d_fentries : b->d_entries;
  for (int k = 0; k < kBatchQueues; ++k) {
    q.off[k] = fast ? b->fqoff[k] : b->qoff[k];
    q.len[k] = fast ? b->fqlen[k] : b->qlen[k];
  }
  q.phase_items = fast ? b->d_fphase_items : b->d_phase_items;
  q.nphase = b->nphase;
  q.heads = b->d_head;
  q.done = b->d_done;
  q.err = b->d_err;
  q.wait_ticks = b->wait_ticks;
  if (fast) return launch_batch_fast(b->rb, q, b->fblocks, b->flds, b->stream);
  return launch_batch(b->rb, q, b->blocks, b->nthreads, b->lds, b->stream);
}

// Captures d2dhip_batch_launch (counter reset + persistent kernel, no
// memset nodes) into a hipGraph owned by the batch.
int d2dhip_batch_capture_graph(d2dhip_batch b) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  if (!b->bound) return set_error(D2DHIP_ERR_INVALID_VALUE, "batch not bound");
  hipStream_t cs = nullptr;
  int st = hip_check(hipStreamCreate(&cs), "hipStreamCreate");
  if (st) return st;
  hipStream_t keep = b->stream;
  b->stream = cs;
  st = hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (!st) {
    const int lst = d2dhip_batch_launch(b);
    hipGraph_t g = nullptr;
    st = hip_check(hipStreamEndCapture(cs, &g), "hipStreamEndCapture");
    if (!st && lst) st = lst;
    if (!st) {
      if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->graph = g;
      st = hip_check(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
  }
  b->stream = keep;
  (void)hipStreamDestroy(cs);
  return st;
}

int d2dhip_batch_graph_launch(d2dhip_batch b) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  if (!b->graph_exec) return set_error(D2DHIP_ERR_INVALID_VALUE, "no captured graph");
  return hip_check(hipGraphLaunch(b->graph_exec, b->stream), "hipGraphLaunch");
}

// Waits for the batch's stream and reports a dependency-wait timeout.
int d2dhip_batch_check(d2dhip_batch b) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  int st = hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  if (st) return st;
  unsigned int err = 0;
  st = hip_check(hipMemcpy(&err, b->d_err, sizeof(err), hipMemcpyDeviceToHost), "hipMemcpy");
  if (!st && err) st = set_error(D2DHIP_ERR_HIP, "batched phase wait timed out (5 s)");
  return st;
}

int d2dhip_batch_set_cluster_bytes(d2dhip_batch b, long long bytes) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  if (bytes < 0) return set_error(D2DHIP_ERR_INVALID_VALUE, "cluster bytes < 0");
  b->cluster_bytes = bytes;
  b->bound = false;  // the queue order changes: bind again
  return D2DHIP_SUCCESS;
}

int d2dhip_batch_set_item_points(d2dhip_batch b, long long points) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  if (points < 1) return set_error(D2DHIP_ERR_INVALID_VALUE, "item points < 1");
  b->item_points = points;
  b->bound = false;  // the queue changes: bind again
  return D2DHIP_SUCCESS;
}

int d2dhip_batch_persistent_tiles(d2dhip_batch b, int *specialised) {
  if (!b || !specialised) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  *specialised = b->fast ? 1 : 0;
  return D2DHIP_SUCCESS;
}

int d2dhip_batch_get_clusters(d2dhip_batch b, int *nclusters, int *nphases) {
  if (!b || !nclusters || !nphases) return set_error(D2DHIP_ERR_NULL, "NULL argument");
  *nclusters = b->nclusters;
  *nphases = b->nphase;
  return D2DHIP_SUCCESS;
}

int d2dhip_batch_set_mode(d2dhip_batch b, int mode) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  if (mode != D2DHIP_BATCH_PERSISTENT && mode != D2DHIP_BATCH_GROUPED &&
      mode != D2DHIP_BATCH_PERSISTENT_GENERAL)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "batch mode");
  if (mode != b->mode) {
    b->mode = mode;
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
  }
  return D2DHIP_SUCCESS;
}

int d2dhip_batch_set_stream(d2dhip_batch b, void *stream) {
  if (!b) return set_error(D2DHIP_ERR_NULL, "batch is NULL");
  b->stream = (hipStream_t)stream;
  return D2DHIP_SUCCESS;
}

int d2dhip_batch_destroy(d2dhip_batch b) {
  if (!b) return D2DHIP_SUCCESS;
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_descs) (void)hipFree(b->d_descs);
  if (b->d_fdescs) (void)hipFree(b->d_fdescs);
  if (b->d_entries) (void)hipFree(b->d_entries);
  if (b->d_fentries) (void)hipFree(b->d_fentries);
  if (b->d_fgroups) (void)hipFree(b->d_fgroups);
  if (b->d_fphase_items) (void)hipFree(b->d_fphase_items);
  if (b->d_groups) (void)hipFree(b->d_groups);
  if (b->d_done) (void)hipFree(b->d_done);
  if (b->d_phase_items) (void)hipFree(b->d_phase_items);
  if (b->d_multi) (void)hipFree(b->d_multi);
  if (b->d_head) (void)hipFree(b->d_head);
  if (b->d_err) (void)hipFree(b->d_err);
  delete b;
  return D2DHIP_SUCCESS;
}

}  // extern "C"
