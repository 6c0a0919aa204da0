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
// accumulates them in one FP32 accumulator, small terms first; the dropped terms
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
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

struct Split3 {
  bf16x8 h, m, l;
};

// Round-to-nearest-even FP32 -> bf16 bits, for finite values (compile time).
constexpr unsigned bf16_rne_bits(float f) {
  const unsigned u = __builtin_bit_cast(unsigned, f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
constexpr double bf16_value(unsigned b) { return (double)__builtin_bit_cast(float, b << 16); }

// The three bf16 parts of (a, b), each part one packed word (a in the low
// half): part k is the round-to-nearest of what parts 0..k-1 leave, taken
// from the FP64 value.
constexpr void split3_pair(double a, double b, unsigned (&w)[3]) {
  for (int k = 0; k < 3; ++k) {
    const unsigned ha = bf16_rne_bits((float)a), hb = bf16_rne_bits((float)b);
    w[k] = ha | (hb << 16);
    a -= bf16_value(ha);
    b -= bf16_value(hb);
  }
}

// Ar and Ai of the radix-16 DFT in the MFMA A-operand layout, split into
// bf16 parts: lane l holds row i = l%16, k = 8*(l/16) .. +7, i.e. for
// e = 0..3 the pair (Ar[i][2s], Ar[i][2s+1]) = (cos t, sin t) and
// (Ai[i][2s], Ai[i][2s+1]) = (-sin t, cos t), s = 4*(l/16) + e,
// t = 2 pi i s / 16; w[l][0: Ar, 1: Ai][part][e].  Built at compile time
// from the correctly rounded FP64 roots.
struct W16Parts {
  unsigned w[64][2][3][4];
  constexpr W16Parts() : w() {
    for (int l = 0; l < 64; ++l) {
      const int i = l & 15, q = l >> 4;
      for (int e = 0; e < 4; ++e) {
        const int m = (i * (4 * q + e)) & 15;
        const double c = kRootCos[kRootOff[16] + m], sn = kRootSin[kRootOff[16] + m];
        unsigned pr[3] = {}, pi[3] = {};
        split3_pair(c, sn, pr);
        split3_pair(-sn, c, pi);
        for (int k = 0; k < 3; ++k) {
          w[l][0][k][e] = pr[k];
          w[l][1][k][e] = pi[k];
        }
      }
    }
  }
};
__constant__ W16Parts kW16P = W16Parts();

// This lane's parts of Ar and Ai.
__device__ __forceinline__ void w16_operands(int lane, Split3 &ar, Split3 &ai) {
  const u32x4 *p = reinterpret_cast<const u32x4 *>(&kW16P.w[lane][0][0][0]);
  ar.h = __builtin_bit_cast(bf16x8, p[0]);
  ar.m = __builtin_bit_cast(bf16x8, p[1]);
  ar.l = __builtin_bit_cast(bf16x8, p[2]);
  ai.h = __builtin_bit_cast(bf16x8, p[3]);
  ai.m = __builtin_bit_cast(bf16x8, p[4]);
  ai.l = __builtin_bit_cast(bf16x8, p[5]);
}

// The six products, small terms first, in one FP32 accumulator.
__device__ __forceinline__ f32x4 mfma6(const Split3 &a, const Split3 &b) {
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.m, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.h, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.m, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, acc, 0, 0, 0);
}

// The bf16 pair of one 32-bit word back to FP32 (element 0 in the low half).
__device__ __forceinline__ float bf_lo(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(unsigned w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ unsigned bf_pack(float a, float b) {
  const float __attribute__((ext_vector_type(2))) v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

// (re, im) -> the three bf16 parts of both, each part one packed word (one
// v_cvt_pk_bf16_f32 per part).
__device__ __forceinline__ void split_pair(float re, float im, unsigned &h, unsigned &m, unsigned &l) {
  h = bf_pack(re, im);
  const float r1 = re - bf_lo(h), i1 = im - bf_hi(h);  // exact
  m = bf_pack(r1, i1);
  l = bf_pack(r1 - bf_lo(m), i1 - bf_hi(m));  // exact differences
}

// Radix-16 DFT of 16 butterflies spread over the wave: x holds this lane's
// inputs 4*(l/16) .. +3 (already twiddled), y receives its outputs
// 4*(l/16) .. +3.
__device__ __forceinline__ void mf_dft16(const cpx<float> (&x)[4], cpx<float> (&y)[4],
                                         const Split3 &ar, const Split3 &ai) {
  u32x4 h, m, l;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    unsigned a, b, c;
    split_pair(x[e].x, x[e].y, a, b, c);
    h[e] = a;
    m[e] = b;
    l[e] = c;
  }
  Split3 bs;
  bs.h = __builtin_bit_cast(bf16x8, h);
  bs.m = __builtin_bit_cast(bf16x8, m);
  bs.l = __builtin_bit_cast(bf16x8, l);
  const f32x4 yr = mfma6(ar, bs);
  const f32x4 yi = mfma6(ai, bs);
#pragma unroll
  for (int e = 0; e < 4; ++e) y[e] = cpx<float>{yr[e], yi[e]};
}

// Where element i of a thread's registers sits in pass p: butterfly (t, b),
// butterfly point s, and whether it exists.  Matrix-core passes (radix 16)
// hold groups of 16 butterflies per wave, lane l taking points
// 4*(l/16) .. +3 of butterfly l%16 of each group (i = 4 q + e); VALU passes
// hold whole butterflies, one per thread and q (i = q R + s).
template <int N, int TX, int NT, bool LINES, class RL, int p>
struct MfPass {
  static constexpr int R = RL::r[p], NBR = N / R, L = RL::L(p);
  static constexpr bool MF = (R == 16);
  static constexpr int W = NT / 64;
  static constexpr int RPT = NT / TX;
  static constexpr int G = MF ? TX * NBR / 16 : 0;  // groups of 16 butterflies
  static constexpr int GW = MF ? G / W : 0;
  static constexpr int NB = MF ? 0 : (NBR + RPT - 1) / RPT;
  static constexpr bool EXACT = MF || (NBR % RPT == 0);
  static constexpr int COUNT = MF ? 4 * GW : NB * R;
  static_assert(!MF || G % W == 0, "groups must divide over the waves");
  __device__ static __forceinline__ void elem(int tid, int i, int &t, int &b, int &s, bool &valid) {
    if constexpr (MF) {
      const int lane = tid & 63, wave = tid >> 6, q = i >> 2;
      const int u = (wave + W * q) * 16 + (lane & 15);
      t = LINES ? u / NBR : u % TX;
      b = LINES ? u % NBR : u / TX;
      s = 4 * (lane >> 4) + (i & 3);
      valid = true;
    } else {
      const int q = i / R;
      s = i % R;
      t = LINES ? tid / RPT : tid % TX;
      b = (LINES ? tid % RPT : tid / TX) + q * RPT;
      valid = EXACT || b < NBR;
    }
  }
  __device__ static __forceinline__ int jin(int b, int s) { return b + s * NBR; }
  __device__ static __forceinline__ int jout(int b, int s) {
    const int k = b % L, g = b / L;
    return g * L * R + k + s * L;
  }
};

template <int N, int TX, int NT, bool LINES, class RL, int p = 0>
constexpr int mf_regs() {
  if constexpr (p == RL::P) {
    return 0;
  } else {
    constexpr int c = MfPass<N, TX, NT, LINES, RL, p>::COUNT;
    constexpr int r = mf_regs<N, TX, NT, LINES, RL, p + 1>();
    return c > r ? c : r;
  }
}

// One tile of a stage (tile < tiles0 * C1), by the whole workgroup.
template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM>
__device__ __forceinline__ void mf_tile(DescPtr d, const void *in, void *out,
                                        int conj, long long tile) {
  using T = float;
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  constexpr int P = RL::P, LDT = C::LDT, LDJ = C::LDJ;
  constexpr int E = mf_regs<N, TX, NT, LINES, RL>();
  static_assert(NT % 64 == 0, "whole waves");
  static_assert(P >= 2, "at least two passes");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(lds_raw);
  T *ldr = reinterpret_cast<T *>(lds_raw);

  const long long tiles0 = d->tiles0;
  const long long c1 = tile / tiles0;
  const long long c0b = (tile - c1 * tiles0) * TX;
  const int tid = threadIdx.x;
  const void *tw = d->tw;

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
  w16_operands(tid & 63, ar, ai);
  cpx<T> v[E];

  // ---- inputs of pass 0 -----------------------------------------------------
  using P0 = MfPass<N, TX, NT, LINES, RL, 0>;
  if constexpr (IO == IO_C2R) {
    // conj(Z) of the merged half spectra into LDS, then the inputs of pass 0
    constexpr int NH = N / 2 + 1;
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
        long long ia = cc * d->in_s0 + c1 * d->in_s1 + (long long)k * d->in_sj;
        long long ib = ia + d->in_s0;
        if (LINES && d->npack) {  // receive-buffer order (fused unpack; x stages)
          ia = pack_index(d, k, cc + c1 * d->C0r);
          ib = pack_index(d, k, cc + 1 + c1 * d->C0r);
        }
        if (cc < d->C0r) A[it] = gload<T>(in, ia);
        if (cc + 1 < d->C0r) B[it] = gload<T>(in, ib);
      }
    }
#pragma unroll
    for (int half = 0; half < (XM ? 2 : 1); ++half) {
      if (half) __syncthreads();
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
          const cpx<T> zk = {a.x - b.y, -(a.y + b.x)}, zm = {a.x + b.y, a.y - b.x};
          const bool mirror = k > 0 && 2 * k < N;
          if constexpr (XM == 0) {
            lds[C::row(k) * LDJ + tt * LDT] = zk;
            if (mirror) lds[C::row(N - k) * LDJ + tt * LDT] = zm;
          } else {
            ldr[C::row(k) * LDJ + tt * LDT] = half ? zk.y : zk.x;
            if (mirror) ldr[C::row(N - k) * LDJ + tt * LDT] = half ? zm.y : zm.x;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < P0::COUNT; ++i) {
        int t, b, sp;
        bool ok;
        P0::elem(tid, i, t, b, sp, ok);
        if (ok) {
          const int a = C::row(P0::jin(b, sp)) * LDJ + t * LDT;
          if constexpr (XM == 0) {
            v[i] = lds[a];
          } else {
            if (half) v[i].y = ldr[a]; else v[i].x = ldr[a];
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < P0::COUNT; ++i) {
      int t, b, sp;
      bool ok;
      P0::elem(tid, i, t, b, sp, ok);
      if (ok) {
        long long ca, ib, ob;
        bool cv;
        col(t, ca, cv, ib, ob);
        v[i] = fload<T, IO>(d, in, ib, cv, true, ca, P0::jin(b, sp), conj);
      }
    }
  }

  // ---- passes -------------------------------------------------------------
  static_for<0, P>([&](auto pc) __attribute__((always_inline)) {
    constexpr int p = decltype(pc)::value;
    using M = MfPass<N, TX, NT, LINES, RL, p>;
    constexpr int R = M::R, L = M::L;
    if constexpr (p > 0) {
      // outputs of pass p-1 -> LDS -> inputs of pass p (real parts, then
      // imaginary parts, for the split exchange)
      using Q = MfPass<N, TX, NT, LINES, RL, p - 1>;
      if constexpr (p > 1 || IO == IO_C2R) __syncthreads();
#pragma unroll
      for (int half = 0; half < (XM ? 2 : 1); ++half) {
        if (half) __syncthreads();
#pragma unroll
        for (int i = 0; i < Q::COUNT; ++i) {
          int t, b, sp;
          bool ok;
          Q::elem(tid, i, t, b, sp, ok);
          if (ok) {
            const int a = C::row(Q::jout(b, sp)) * LDJ + t * LDT;
            if constexpr (XM == 0)
              lds[a] = v[i];
            else
              ldr[a] = half ? v[i].y : v[i].x;
          }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < M::COUNT; ++i) {
          int t, b, sp;
          bool ok;
          M::elem(tid, i, t, b, sp, ok);
          if (ok) {
            const int a = C::row(M::jin(b, sp)) * LDJ + t * LDT;
            if constexpr (XM == 0) {
              v[i] = lds[a];
            } else {
              if (half) v[i].y = ldr[a]; else v[i].x = ldr[a];
            }
          }
        }
      }
    }
    // twiddles, then the butterflies
    if constexpr (L > 1) {
#pragma unroll
      for (int i = 0; i < M::COUNT; ++i) {
        int t, b, sp;
        bool ok;
        M::elem(tid, i, t, b, sp, ok);
        if (ok && sp > 0) v[i] = cmul(v[i], gload<T>(tw, RL::tw_off(p) + (b % L) * (R - 1) + sp - 1));
      }
    }
    if constexpr (M::MF) {
#pragma unroll
      for (int q = 0; q < M::GW; ++q) {
        cpx<T> x[4] = {v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]}, y[4];
        mf_dft16(x, y, ar, ai);
#pragma unroll
        for (int e = 0; e < 4; ++e) v[4 * q + e] = y[e];
      }
    } else {
#pragma unroll
      for (int q = 0; q < M::NB; ++q) {
        int t, b, sp;
        bool ok;
        M::elem(tid, q * R, t, b, sp, ok);
        if (ok) dft<T, R>(&v[q * R]);
      }
    }
  });

  // ---- outputs of the last pass ---------------------------------------------
  using Z = MfPass<N, TX, NT, LINES, RL, P - 1>;
  if constexpr (IO == IO_R2C) {
    constexpr int NH = N / 2 + 1;
    constexpr int ITER = (TX * NH + NT - 1) / NT;
    cpx<T> oa[ITER], ob[ITER];
    const T half_ = T(0.5);
#pragma unroll
    for (int half = 0; half < (XM ? 2 : 1); ++half) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < Z::COUNT; ++i) {
        int t, b, sp;
        bool ok;
        Z::elem(tid, i, t, b, sp, ok);
        if (ok) {
          const int a = C::row(Z::jout(b, sp)) * LDJ + t * LDT;
          if constexpr (XM == 0)
            lds[a] = v[i];
          else
            ldr[a] = half ? v[i].y : v[i].x;
        }
      }
      __syncthreads();
#pragma unroll
      for (int it = 0; it < ITER; ++it) {
        const int w = tid + it * NT;
        if (w < TX * NH) {
          const int tt = LINES ? w / NH : w % TX;
          const int k = LINES ? w % NH : w / TX;
          const int km = (k == 0) ? 0 : N - k;
          const int a1 = C::row(k) * LDJ + tt * LDT, a2 = C::row(km) * LDJ + tt * LDT;
          if constexpr (XM == 0) {
            const cpx<T> zk = lds[a1], zm = lds[a2];
            oa[it] = cpx<T>{(zk.x + zm.x) * half_, (zk.y - zm.y) * half_};
            ob[it] = cpx<T>{(zk.y + zm.y) * half_, (zm.x - zk.x) * half_};
          } else {
            const T zk = ldr[a1], zm = ldr[a2];
            if (half == 0) {
              oa[it].x = (zk + zm) * half_;
              ob[it].y = (zm - zk) * half_;
            } else {
              oa[it].y = (zk - zm) * half_;
              ob[it].x = (zk + zm) * half_;
            }
          }
        }
      }
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int w = tid + it * NT;
      if (w < TX * NH) {
        const int tt = LINES ? w / NH : w % TX;
        const int k = LINES ? w % NH : w / TX;
        const long long cc = 2 * (c0b + tt);
        if (cc < d->C0r) {
          long long oa_i = cc * d->out_s0 + c1 * d->out_s1 + (long long)k * d->out_sj;
          long long ob_i = oa_i + d->out_s0;
          if (LINES && d->npack) {  // send-buffer order (fused pack; x stages)
            oa_i = pack_index(d, k, cc + c1 * d->C0r);
            ob_i = pack_index(d, k, cc + 1 + c1 * d->C0r);
          }
          gstore<T>(out, oa_i, oa[it]);
          if (cc + 1 < d->C0r) gstore<T>(out, ob_i, ob[it]);
        }
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < Z::COUNT; ++i) {
      int t, b, sp;
      bool ok;
      Z::elem(tid, i, t, b, sp, ok);
      if (ok) {
        long long ca, ib, ob;
        bool cv;
        col(t, ca, cv, ib, ob);
        fstore<T, IO>(d, out, ob, cv, ca, Z::jout(b, sp), v[i], conj);
      }
    }
  }
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
    mf_axis_kernel(const AxisDesc *__restrict__ d, const void *in, void *out, int conj) {
  const long long ntiles = (long long)d->tiles0 * d->C1;
  const long long tile = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, ntiles);
  if (tile >= ntiles) return;  // whole block, before any barrier
  mf_tile<N, TX, NT, IO, LINES, RL, XM>(desc_ptr(d), in, out, conj, tile);
}

// The same stage kernel over the tiles of several grids.
template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
    mf_multi_kernel(const MultiGroup *__restrict__ g, int ng, long long tiles) {
  const long long it = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, tiles);
  if (it >= tiles) return;  // whole block, before any barrier
  const int gi = find_multi_group(g, ng, it);
  mf_tile<N, TX, NT, IO, LINES, RL, XM>(desc_ptr(g[gi].d), g[gi].in, g[gi].out, g[gi].conj,
                                        it - g[gi].tile0);
}

// ---- registry -----------------------------------------------------------------

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static int launch_mf(const AxisDesc *dd, const AxisDesc &h, const AxisIo &io, hipStream_t st) {
  using C = FastCfg<float, N, TX, NT, LINES, RL, XM>;
  const long long ntiles = xcd_grid((long long)h.tiles0 * h.C1);
  if (ntiles <= 0) return D2DHIP_SUCCESS;
  const long long gx = ntiles < (1LL << 20) ? ntiles : (1LL << 20);
  const long long gy = (ntiles + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "axis stage has too many tiles");
  hipLaunchKernelGGL((mf_axis_kernel<N, TX, NT, IO, LINES, RL, XM, WPE>), dim3((unsigned)gx, (unsigned)gy),
                     dim3(NT), C::lds_bytes(), st, dd, io.in, io.out, io.conj);
  return hip_check(hipGetLastError(), "mf_axis_kernel launch");
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static hipError_t prepare_mf() {
  using C = FastCfg<float, N, TX, NT, LINES, RL, XM>;
  return hipFuncSetAttribute((const void *)mf_axis_kernel<N, TX, NT, IO, LINES, RL, XM, WPE>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::lds_bytes());
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static int launch_mf_multi(const MultiGroup *g, int ng, long long tiles, hipStream_t st) {
  using C = FastCfg<float, N, TX, NT, LINES, RL, XM>;
  const long long nb = xcd_grid(tiles);
  if (nb <= 0) return D2DHIP_SUCCESS;
  const long long gx = nb < (1LL << 20) ? nb : (1LL << 20);
  const long long gy = (nb + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "launch has too many tiles");
  hipLaunchKernelGGL((mf_multi_kernel<N, TX, NT, IO, LINES, RL, XM, WPE>), dim3((unsigned)gx, (unsigned)gy),
                     dim3(NT), C::lds_bytes(), st, g, ng, tiles);
  return hip_check(hipGetLastError(), "mf_multi_kernel launch");
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static hipError_t prepare_mf_multi() {
  using C = FastCfg<float, N, TX, NT, LINES, RL, XM>;
  return hipFuncSetAttribute((const void *)mf_multi_kernel<N, TX, NT, IO, LINES, RL, XM, WPE>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::lds_bytes());
}

template <int... Rs>
static std::vector<int> mf_rvec(RList<Rs...>) {
  return {Rs...};
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static FastVariant make_mf(const char *name) {
  using C = FastCfg<float, N, TX, NT, LINES, RL, XM>;
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
  v.launch = &launch_mf<N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.prepare = &prepare_mf<N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.launch_multi = &launch_mf_multi<N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.prepare_multi = &prepare_mf_multi<N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.name = name;
  return v;
}

// c2c, r2c and c2r of one shape; XM = 1 adds " xs", WPE > 1 " wN" to the name
#define D2D_MF_VARIANTS(N, TX, NT, LINES, RL, XM, WPE, TAG)                                      \
  make_mf<N, TX, NT, IO_C2C, LINES, RL, XM, WPE>("mfma n=" #N " tx=" #TX " nt=" #NT " c2c" TAG),   \
      make_mf<N, TX, NT, IO_R2C, LINES, RL, XM, WPE>("mfma n=" #N " tx=" #TX " nt=" #NT " r2c" TAG), \
      make_mf<N, TX, NT, IO_C2R, LINES, RL, XM, WPE>("mfma n=" #N " tx=" #TX " nt=" #NT " c2r" TAG)

using M16_16_4 = RList<16, 16, 4>;
using M16_16_2 = RList<16, 16, 2>;

std::vector<FastVariant> mf_variants() {
  return {
      // 1024 = 16 * 16 * 4: strided columns
      D2D_MF_VARIANTS(1024, 16, 1024, false, M16_16_4, 0, 4, ""),
      D2D_MF_VARIANTS(1024, 8, 512, false, M16_16_4, 0, 4, ""),
      D2D_MF_VARIANTS(1024, 8, 512, false, M16_16_4, 1, 5, " xs w5"),
      D2D_MF_VARIANTS(1024, 4, 256, false, M16_16_4, 1, 5, " xs w5"),
      D2D_MF_VARIANTS(1024, 16, 1024, false, M16_16_4, 1, 4, " xs"),
      // two groups per wave (8 values per lane): 64-74 VGPRs, 6-8 waves per SIMD
      D2D_MF_VARIANTS(1024, 4, 512, false, M16_16_4, 1, 6, " xs w6"),
      D2D_MF_VARIANTS(1024, 2, 256, false, M16_16_4, 1, 6, " xs w6"),
      D2D_MF_VARIANTS(1024, 8, 1024, false, M16_16_4, 1, 8, " xs w8"),
      D2D_MF_VARIANTS(1024, 8, 1024, false, M16_16_4, 0, 4, ""),
      // contiguous lines
      D2D_MF_VARIANTS(1024, 4, 256, true, M16_16_4, 0, 4, ""),
      D2D_MF_VARIANTS(1024, 8, 512, true, M16_16_4, 0, 4, ""),
      D2D_MF_VARIANTS(1024, 4, 256, true, M16_16_4, 1, 5, " xs w5"),
      D2D_MF_VARIANTS(1024, 8, 512, true, M16_16_4, 1, 5, " xs w5"),
      D2D_MF_VARIANTS(1024, 4, 512, true, M16_16_4, 1, 6, " xs w6"),
      // 512 = 16 * 16 * 2
      D2D_MF_VARIANTS(512, 16, 512, false, M16_16_2, 0, 4, ""),
      D2D_MF_VARIANTS(512, 8, 256, true, M16_16_2, 0, 4, ""),
  };
}

}  // namespace d2dhip
