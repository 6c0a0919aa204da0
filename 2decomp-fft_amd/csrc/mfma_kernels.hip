// mfma_kernels.hip -- split-BF16 axis stages: radix-16 passes on the gfx950
// matrix cores (v_mfma_f32_16x16x32_bf16), FP32 storage and FP32 twiddles.
//
// A radix-16 butterfly is the product of the 16x16 DFT matrix W with its 16
// twiddled inputs.  Over the reals, with the inputs interleaved as
// (re_0, im_0, re_1, im_1, ...), the real and imaginary outputs are two
// 16x32 by 32x1 products, yr = Ar x and yi = Ai x, where for t = 2 pi i s/16
//   Ar[i][2s] = cos t, Ar[i][2s+1] = sin t, Ai[i][2s] = -sin t, Ai[i][2s+1] = cos t.
// One MFMA evaluates 16 butterflies (the 16 columns of B).
//
// Precision: every FP32 operand is split into three bf16 parts,
// x = xh + xm + xl, each the round-to-nearest-even of what the earlier parts
// leave (x - xh and xm's residual are exact in FP32), so the parts hold the
// full 24-bit significand.  The product keeps the six terms whose size is at
// least 2^-16 of the leading one, hh + hm + mh + hl + lh + mm, and
// accumulates them in FP32 with the small terms first; the dropped terms
// (ml, lm, ll) are below 2^-24 relative.  W is split the same way from its
// correctly rounded FP64 value.  A pass therefore carries about FP32
// rounding error, and a stage about that of the FP32 VALU kernels.
//
// Lane layout of v_mfma_f32_16x16x32_bf16 (wave of 64 lanes, lane l):
//   A (16x32): row l%16, k = 8*(l/16) .. 8*(l/16)+7   (8 bf16 per lane)
//   B (32x16): col l%16, k = 8*(l/16) .. 8*(l/16)+7
//   D (16x16): col l%16, rows 4*(l/16) .. 4*(l/16)+3  (4 f32 per lane)
// Lane l therefore works on butterfly l%16 of a group of 16: it loads the
// inputs s = 4*(l/16) .. +3 (k = 2s - 8*(l/16) + {0: re, 1: im}), twiddles
// and splits them itself, and receives the outputs 4*(l/16) .. +3.
//
// Passes with another radix (the 4 of 1024 = 16*16*4, the 2 of
// 512 = 16*16*2) run on the VALU with the FP32 codelets, one butterfly per
// thread, as in fast_kernels.hip.
#include <hip/hip_runtime.h>

#include "codelets.h"
#include "fast_common.h"
#include "fast_kernels.h"
#include "internal.h"

namespace d2dhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Split3 {
  bf16x8 h, m, l;
};

// cos and sin of 2 pi m / 16 (correctly rounded FP64).
struct W16Table {
  double c[16], s[16];
  constexpr W16Table() : c(), s() {
    for (int m = 0; m < 16; ++m) {
      c[m] = kRootCos[kRootOff[16] + m];
      s[m] = kRootSin[kRootOff[16] + m];
    }
  }
};
__constant__ W16Table kW16 = W16Table();

template <int E>
__device__ __forceinline__ void split_f32(float x, Split3 &o) {
  const __bf16 h = (__bf16)x;
  const float r = x - (float)h;  // exact
  const __bf16 m = (__bf16)r;
  const __bf16 l = (__bf16)(r - (float)m);  // the difference is exact
  o.h[E] = h;
  o.m[E] = m;
  o.l[E] = l;
}

template <int E>
__device__ __forceinline__ void split_f64(double x, Split3 &o) {
  const __bf16 h = (__bf16)(float)x;
  const double r = x - (double)(float)h;  // exact
  const __bf16 m = (__bf16)(float)r;
  const double r2 = r - (double)(float)m;  // exact
  o.h[E] = h;
  o.m[E] = m;
  o.l[E] = (__bf16)(float)r2;
}

// This lane's parts of Ar and Ai (rows l%16, k = 8*(l/16) .. +7).
__device__ __forceinline__ void w16_operands(int lane, Split3 &ar, Split3 &ai) {
  const int i = lane & 15, q = lane >> 4;
  static_for<0, 4>([&](auto ec) __attribute__((always_inline)) {
    constexpr int e = decltype(ec)::value;
    const int m = (i * (4 * q + e)) & 15;
    const double c = kW16.c[m], s = kW16.s[m];
    split_f64<2 * e>(c, ar);
    split_f64<2 * e + 1>(s, ar);
    split_f64<2 * e>(-s, ai);
    split_f64<2 * e + 1>(c, ai);
  });
}

__device__ __forceinline__ f32x4 mfma6(const Split3 &a, const Split3 &b) {
  f32x4 lo = {0.f, 0.f, 0.f, 0.f}, hh = {0.f, 0.f, 0.f, 0.f};
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.m, lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.h, lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.m, lo, 0, 0, 0);
  hh = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, hh, 0, 0, 0);
  return hh + lo;
}

// Radix-16 DFT of 16 butterflies spread over the wave: x holds this lane's
// inputs 4*(l/16) .. +3 (already twiddled), y receives its outputs
// 4*(l/16) .. +3.
__device__ __forceinline__ void mf_dft16(const cpx<float> (&x)[4], cpx<float> (&y)[4],
                                         const Split3 &ar, const Split3 &ai) {
  Split3 b;
  static_for<0, 4>([&](auto ec) __attribute__((always_inline)) {
    constexpr int e = decltype(ec)::value;
    split_f32<2 * e>(x[e].x, b);
    split_f32<2 * e + 1>(x[e].y, b);
  });
  const f32x4 yr = mfma6(ar, b);
  const f32x4 yi = mfma6(ai, b);
#pragma unroll
  for (int e = 0; e < 4; ++e) y[e] = cpx<float>{yr[e], yi[e]};
}

template <int N, int TX, int NT, int IO, bool LINES, class RL>
__global__ void __launch_bounds__(NT)
    mf_axis_kernel(const AxisDesc *__restrict__ d, const void *in, void *out, int conj) {
  using T = float;
  using C = FastCfg<T, N, TX, NT, LINES, RL, 0>;
  constexpr int P = RL::P, RPT = C::RPT, LDT = C::LDT, LDJ = C::LDJ;
  constexpr int W = NT / 64;
  static_assert(NT % 64 == 0, "whole waves");
  static_assert(P >= 2, "at least two passes");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(lds_raw);

  const long long tiles0 = d->tiles0;
  const long long tile =
      xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, tiles0 * d->C1);
  if (tile >= tiles0 * d->C1) return;  // whole block, before any barrier
  const long long c1 = tile / tiles0;
  const long long c0b = (tile - c1 * tiles0) * TX;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(d->tw);

  // column t of the tile: first pencil column, validity and element bases
  auto col = [&](int t, long long &ca, bool &cvalid, long long &ib, long long &ob)
      __attribute__((always_inline)) {
    const long long c0 = c0b + t;
    ca = 2 * c0;
    cvalid = (IO == IO_C2C) ? (c0 < d->C0) : (ca < d->C0r);
    ib = (IO == IO_R2C ? ca : c0) * d->in_s0 + c1 * d->in_s1;
    ob = (IO == IO_C2R ? ca : c0) * d->out_s0 + c1 * d->out_s1;
  };

  Split3 ar, ai;
  w16_operands(lane, ar, ai);

  if constexpr (IO == IO_C2R) {
    // conj(Z) of the merged half spectra, all N points, into LDS
    constexpr int NH = N / 2 + 1;
    const cpx<T> *src = reinterpret_cast<const cpx<T> *>(in);
    // all loads first (compile-time trip count), then the LDS writes, so a
    // thread has every load of the phase in flight at once
    constexpr int ITER = (TX * NH + NT - 1) / NT;
    cpx<T> A[ITER], B[ITER];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int w = tid + it * NT;
      A[it] = cpx<T>{T(0), T(0)};
      B[it] = cpx<T>{T(0), T(0)};
      if (w < TX * NH) {
        const int tt = LINES ? w / NH : w % TX;
        const int k = LINES ? w % NH : w / TX;
        const long long cc = 2 * (c0b + tt);
        const long long base = cc * d->in_s0 + c1 * d->in_s1 + (long long)k * d->in_sj;
        if (cc < d->C0r) A[it] = src[base];
        if (cc + 1 < d->C0r) B[it] = src[base + d->in_s0];
      }
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int w = tid + it * NT;
      if (w < TX * NH) {
        const int tt = LINES ? w / NH : w % TX;
        const int k = LINES ? w % NH : w / TX;
        cpx<T> a = A[it], b = B[it];
        if (k == 0 || 2 * k == N) {
          a.y = T(0);
          b.y = T(0);
        }
        lds[C::row(k) * LDJ + tt * LDT] = cpx<T>{a.x - b.y, -(a.y + b.x)};
        if (k > 0 && 2 * k < N) lds[C::row(N - k) * LDJ + tt * LDT] = cpx<T>{a.x + b.y, a.y - b.x};
      }
    }
  }

  static_for<0, P>([&](auto pc) __attribute__((always_inline)) {
    constexpr int p = decltype(pc)::value;
    constexpr int R = RL::r[p], NBR = N / R, L = RL::L(p);
    constexpr bool src_lds = (p > 0) || IO == IO_C2R;
    constexpr bool dst_lds = (p < P - 1) || IO == IO_R2C;
    if constexpr (src_lds) __syncthreads();
    if constexpr (R == 16) {
      constexpr int G = TX * NBR / 16;  // groups of 16 butterflies in the tile
      static_assert(G % W == 0, "groups must divide over the waves");
      constexpr int GW = G / W;
      const int s0 = 4 * (lane >> 4);
      cpx<T> y[GW][4];
#pragma unroll
      for (int q = 0; q < GW; ++q) {
        const int u = (wave + W * q) * 16 + (lane & 15);
        const int t = LINES ? u / NBR : u % TX;
        const int b = LINES ? u % NBR : u / TX;
        cpx<T> x[4];
        if constexpr (src_lds) {
#pragma unroll
          for (int e = 0; e < 4; ++e) x[e] = lds[C::row(b + (s0 + e) * NBR) * LDJ + t * LDT];
        } else {
          long long ca, ib, ob;
          bool cv;
          col(t, ca, cv, ib, ob);
#pragma unroll
          for (int e = 0; e < 4; ++e) x[e] = fload<T, IO>(d, in, ib, cv, true, ca, b + (s0 + e) * NBR, conj);
        }
        if constexpr (L > 1) {
          const int k = b % L;
          const cpx<T> *tp = tw + RL::tw_off(p) + k * (R - 1);
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (s0 + e > 0) x[e] = cmul(x[e], tp[s0 + e - 1]);
        }
        mf_dft16(x, y[q], ar, ai);
      }
      if constexpr (src_lds && dst_lds) __syncthreads();
#pragma unroll
      for (int q = 0; q < GW; ++q) {
        const int u = (wave + W * q) * 16 + (lane & 15);
        const int t = LINES ? u / NBR : u % TX;
        const int b = LINES ? u % NBR : u / TX;
        const int k = b % L, g = b / L;
        if constexpr (dst_lds) {
#pragma unroll
          for (int e = 0; e < 4; ++e) lds[C::row(g * L * 16 + k + (s0 + e) * L) * LDJ + t * LDT] = y[q][e];
        } else {
          long long ca, ib, ob;
          bool cv;
          col(t, ca, cv, ib, ob);
#pragma unroll
          for (int e = 0; e < 4; ++e) fstore<T, IO>(d, out, ob, cv, ca, g * L * 16 + k + (s0 + e) * L, y[q][e], conj);
        }
      }
    } else {
      // one butterfly per thread on the VALU
      constexpr int NB = C::nb(p);
      const int t = LINES ? tid / RPT : tid % TX;
      const int r0 = LINES ? tid % RPT : tid / TX;
      long long ca, ib, ob;
      bool cv;
      col(t, ca, cv, ib, ob);
      cpx<T> v[NB * R];
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int b = r0 + q * RPT;
        if (C::exact(p) || b < NBR) {
#pragma unroll
          for (int s = 0; s < R; ++s) {
            if constexpr (src_lds)
              v[q * R + s] = lds[C::row(b + s * NBR) * LDJ + t * LDT];
            else
              v[q * R + s] = fload<T, IO>(d, in, ib, cv, true, ca, b + s * NBR, conj);
          }
          if constexpr (L > 1) {
            const int k = b % L;
            const cpx<T> *tp = tw + RL::tw_off(p) + k * (R - 1);
#pragma unroll
            for (int s = 1; s < R; ++s) v[q * R + s] = cmul(v[q * R + s], tp[s - 1]);
          }
          dft<T, R>(&v[q * R]);
        }
      }
      if constexpr (src_lds && dst_lds) __syncthreads();
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int b = r0 + q * RPT;
        if (C::exact(p) || b < NBR) {
          const int k = b % L, g = b / L;
#pragma unroll
          for (int s = 0; s < R; ++s) {
            if constexpr (dst_lds)
              lds[C::row(g * L * R + k + s * L) * LDJ + t * LDT] = v[q * R + s];
            else
              fstore<T, IO>(d, out, ob, cv, ca, g * L * R + k + s * L, v[q * R + s], conj);
          }
        }
      }
    }
  });

  if constexpr (IO == IO_R2C) {
    __syncthreads();
    constexpr int NH = N / 2 + 1;
    cpx<T> *dst = reinterpret_cast<cpx<T> *>(out);
    const T half = T(0.5);
    for (int w = tid; w < TX * NH; w += NT) {
      const int tt = LINES ? w / NH : w % TX;
      const int k = LINES ? w % NH : w / TX;
      const long long cc = 2 * (c0b + tt);
      if (cc >= d->C0r) continue;
      const cpx<T> zk = lds[C::row(k) * LDJ + tt * LDT];
      const int km = (k == 0) ? 0 : N - k;
      const cpx<T> zm = lds[C::row(km) * LDJ + tt * LDT];
      const long long base = cc * d->out_s0 + c1 * d->out_s1 + (long long)k * d->out_sj;
      dst[base] = cpx<T>{(zk.x + zm.x) * half, (zk.y - zm.y) * half};
      if (cc + 1 < d->C0r) dst[base + d->out_s0] = cpx<T>{(zk.y + zm.y) * half, (zm.x - zk.x) * half};
    }
  }
}

// ---- registry -----------------------------------------------------------------

template <int N, int TX, int NT, int IO, bool LINES, class RL>
static int launch_mf(const AxisDesc *dd, const AxisDesc &h, const AxisIo &io, hipStream_t st) {
  using C = FastCfg<float, N, TX, NT, LINES, RL, 0>;
  const long long ntiles = xcd_grid((long long)h.tiles0 * h.C1);
  if (ntiles <= 0) return D2DHIP_SUCCESS;
  const long long gx = ntiles < (1LL << 20) ? ntiles : (1LL << 20);
  const long long gy = (ntiles + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "axis stage has too many tiles");
  hipLaunchKernelGGL((mf_axis_kernel<N, TX, NT, IO, LINES, RL>), dim3((unsigned)gx, (unsigned)gy),
                     dim3(NT), C::lds_bytes(), st, dd, io.in, io.out, io.conj);
  return hip_check(hipGetLastError(), "mf_axis_kernel launch");
}

template <int N, int TX, int NT, int IO, bool LINES, class RL>
static hipError_t prepare_mf() {
  using C = FastCfg<float, N, TX, NT, LINES, RL, 0>;
  return hipFuncSetAttribute((const void *)mf_axis_kernel<N, TX, NT, IO, LINES, RL>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::lds_bytes());
}

template <int... Rs>
static std::vector<int> mf_rvec(RList<Rs...>) {
  return {Rs...};
}

template <int N, int TX, int NT, int IO, bool LINES, class RL>
static FastVariant make_mf(const char *name) {
  using C = FastCfg<float, N, TX, NT, LINES, RL, 0>;
  FastVariant v;
  v.rb = 4;
  v.mf = 1;
  v.io = IO;
  v.lines = LINES ? 1 : 0;
  v.n = N;
  v.TX = TX;
  v.NT = NT;
  v.lds_bytes = C::lds_bytes();
  v.ldj = C::LDJ;
  v.ldt = C::LDT;
  v.radices = mf_rvec(RL{});
  v.launch = &launch_mf<N, TX, NT, IO, LINES, RL>;
  v.prepare = &prepare_mf<N, TX, NT, IO, LINES, RL>;
  v.name = name;
  return v;
}

#define D2D_MF_VARIANTS(N, TX, NT, LINES, RL)                                          \
  make_mf<N, TX, NT, IO_C2C, LINES, RL>("mfma n=" #N " tx=" #TX " nt=" #NT " c2c"),   \
      make_mf<N, TX, NT, IO_R2C, LINES, RL>("mfma n=" #N " tx=" #TX " nt=" #NT " r2c"), \
      make_mf<N, TX, NT, IO_C2R, LINES, RL>("mfma n=" #N " tx=" #TX " nt=" #NT " c2r")

using M16_16_4 = RList<16, 16, 4>;
using M16_16_2 = RList<16, 16, 2>;

std::vector<FastVariant> mf_variants() {
  return {
      D2D_MF_VARIANTS(1024, 16, 1024, false, M16_16_4),
      D2D_MF_VARIANTS(1024, 8, 512, false, M16_16_4),
      D2D_MF_VARIANTS(1024, 4, 256, true, M16_16_4),
      D2D_MF_VARIANTS(1024, 8, 512, true, M16_16_4),
      D2D_MF_VARIANTS(512, 16, 512, false, M16_16_2),
      D2D_MF_VARIANTS(512, 8, 256, true, M16_16_2),
  };
}

}  // namespace d2dhip
