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
// thread, as in fast_impl.h.
//
// What the instruction count comes to: the lane maps below make every LDS
// and HBM address a lane base plus a compile-time offset (a first version
// that derived them per value issued 2560 VALU instructions per lane and
// tile against 806 for the FP32 kernel; now about 860), the split takes 7
// VALU instructions per complex value (v_cvt_pk_bf16_f32 and exact
// v_dot2c_f32_bf16 residuals), and the HBM loads of a tile are all issued
// before the first butterfly.
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

__device__ __forceinline__ unsigned bf_pack(float a, float b) {
  const float __attribute__((ext_vector_type(2))) v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

// The three bf16 parts of (re, im), each part one packed word (re in the low
// half).  The residual of a part is taken with v_dot2c_f32_bf16 against
// (-1, -0) or (0, -1): the products with the packed part are exact and so
// is their sum with the value (the difference of a value and its
// round-to-nearest bf16 fits FP32), one instruction per residual instead of
// an unpack and a subtraction.  (-1, -0) rather than (-1, 0): the compiler
// encodes the latter as the inline constant -1.0, whose two halves the
// hardware reads the other way round.
__device__ __forceinline__ float res_lo(unsigned w, float x) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w),
                                         __builtin_bit_cast(bf16x2, 0x8000bf80u), x, false);
}
__device__ __forceinline__ float res_hi(unsigned w, float x) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w),
                                         __builtin_bit_cast(bf16x2, 0xbf800000u), x, false);
}
__device__ __forceinline__ void split3(float re, float im, unsigned &h, unsigned &m, unsigned &l) {
  h = bf_pack(re, im);
  const float r1 = res_lo(h, re), i1 = res_hi(h, im);
  m = bf_pack(r1, i1);
  l = bf_pack(res_lo(m, r1), res_hi(m, i1));
}

// The parts of (im, -re) from those of (re, im): halves exchanged, the new
// high half negated (bf16 rounding commutes with negation).
__device__ __forceinline__ unsigned swap_neg(unsigned w) {
  return __builtin_amdgcn_alignbit(w, w, 16) ^ 0x80000000u;
}

// The workgroup's LDS copy of the split DFT-16 operands (SW >= 2), part
// major so that a wave's read of one part is 64 consecutive 16-byte words:
// entry (which * 3 + part) * 64 + lane, which 0 = Ar, 1 = Ai, part 0 = h,
// 1 = m, 2 = l.
constexpr int kW16LdsBytes = 6 * 64 * 16;

// The six products with A read from that LDS copy next to its MFMAs, one
// part (4 VGPRs) live at a time, small terms first as far as the part order
// allows: l.h, m.m, m.h, h.l, h.m, h.h (the terms of size 2^-16 of the
// leading one are all added before it).  The opaque zero keeps the compiler
// from merging the reads of different groups into registers held across
// the pass, which would give back the registers this form saves.
__device__ __forceinline__ f32x4 mfma6_lds(const u32x4 *tab, int which, const Split3 &b) {
  int z = 0;
  asm volatile("" : "+s"(z));
  const u32x4 *p = tab + z + which * 3 * 64;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const bf16x8 al = __builtin_bit_cast(bf16x8, p[2 * 64]);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b.h, acc, 0, 0, 0);
  const bf16x8 am = __builtin_bit_cast(bf16x8, p[64]);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, b.m, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, b.h, acc, 0, 0, 0);
  const bf16x8 ah = __builtin_bit_cast(bf16x8, p[0]);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b.l, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b.m, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b.h, acc, 0, 0, 0);
}

// Radix-16 DFT of 16 butterflies spread over the wave: x holds this lane's
// inputs 4*(l/16) .. +3 (already twiddled), y receives its outputs
// 4*(l/16) .. +3.  SW = 0: yr = Ar x, yi = Ai x.  SW = 1: Ai is not held;
// yi = Ar x' with x'_s = (im_s, -re_s), since Ai[i][2s] = -Ar[i][2s+1] and
// Ai[i][2s+1] = Ar[i][2s] (12 fewer registers, 12 more VALU operations per
// group).  SW = 2 / 3: as 0 / 1 with A read from the LDS copy `tab` (this
// lane's entry of part 0) instead of held in registers.
template <int SW>
__device__ __forceinline__ void mf_dft16(const cpx<float> (&x)[4], cpx<float> (&y)[4],
                                         const Split3 &ar, const Split3 &ai,
                                         const u32x4 *tab = nullptr) {
  u32x4 h, m, l;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    unsigned a, b, c;
    split3(x[e].x, x[e].y, a, b, c);
    h[e] = a;
    m[e] = b;
    l[e] = c;
  }
  Split3 bs;
  bs.h = __builtin_bit_cast(bf16x8, h);
  bs.m = __builtin_bit_cast(bf16x8, m);
  bs.l = __builtin_bit_cast(bf16x8, l);
  const f32x4 yr = SW >= 2 ? mfma6_lds(tab, 0, bs) : mfma6(ar, bs);
  f32x4 yi;
  if constexpr (SW == 1 || SW == 3) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      h[e] = swap_neg(h[e]);
      m[e] = swap_neg(m[e]);
      l[e] = swap_neg(l[e]);
    }
    bs.h = __builtin_bit_cast(bf16x8, h);
    bs.m = __builtin_bit_cast(bf16x8, m);
    bs.l = __builtin_bit_cast(bf16x8, l);
    yi = SW == 3 ? mfma6_lds(tab, 0, bs) : mfma6(ar, bs);
  } else {
    yi = SW == 2 ? mfma6_lds(tab, 1, bs) : mfma6(ai, bs);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) y[e] = cpx<float>{yr[e], yi[e]};
}

// Lane maps.  A stage of length N = 16 * 16 * R2 (R2 <= 8) runs two
// radix-16 passes on the matrix cores (L = 1, 16) and one radix-R2 pass on
// the VALU (L = 256).
//
// Matrix-core passes: a wave holds groups of 16 butterflies, lane l taking
// butterfly l%16 of each group and its points s = 4*(l/16) + e, e < 4.  The
// q-th group of a lane has butterfly b = b0 + BQ*q of tile column
// t = t0 + TQ*q, with (b0, t0) fixed per lane and BQ, TQ compile-time:
//   columns: t0 = (l%16) % TX, b0 = wave*16/TX + (l%16)/TX, BQ = W*16/TX, TQ = 0
//   lines:   t0 = wave / NG,  b0 = (wave%NG)*16 + l%16,      BQ = 0, TQ = W/NG
// (NG = N/256 groups per line).  With BQ a multiple of 16, a point index
// j = b + s*N/16 and its LDS row j + j/16 are affine in (q, e), and b%16 is
// the same for every group of the lane, so every LDS address is a lane base
// plus an immediate and the second pass's twiddles are four loads per lane.
//
// VALU pass: one butterfly per thread and q, as in fast_impl.h:
// t = tid%TX (columns) or tid/RPT (lines), b = (tid/TX or tid%RPT) + q*RPT.
template <int N, int TX, int NT, bool LINES>
struct MfMap {
  static constexpr int W = NT / 64;
  static constexpr int NG = N / 256;
  static constexpr int E = TX * N / NT;  // values per lane
  static constexpr int GW = E / 4;       // groups per lane in a matrix-core pass
  static constexpr int BQ = LINES ? 0 : W * 16 / TX;
  static constexpr int TQ = LINES ? W / NG : 0;
  static_assert(E * NT == TX * N && E % 4 == 0, "whole groups per lane");
  static_assert(LINES ? (W % NG == 0) : (16 % TX == 0 && W % TX == 0), "affine lane map");
};

// Compile-time shape of a split-BF16 stage tile: the values a lane fetches
// from HBM (the pass-0 inputs, or for c2r the two half-spectrum points of
// each ITER slot), and the LDS layout (the tile, then optionally the
// twiddles of passes 1 and 2).
template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM>
struct MfShape {
  using C = FastCfg<float, N, TX, NT, LINES, RL, XM>;
  using M = MfMap<N, TX, NT, LINES>;
  static constexpr int E = M::E, NH = N / 2 + 1, ITER = (TX * NH + NT - 1) / NT;
  static constexpr int F = IO == IO_C2R ? 2 * ITER : E;
  static constexpr int TW0 = RL::tw_off(1), TWN = RL::tw_off(3) - RL::tw_off(1);
  static constexpr size_t tw_lds_off = (C::lds_bytes() + 15) / 16 * 16;
  static constexpr size_t lds_bytes(bool twl) {
    return twl ? tw_lds_off + (size_t)TWN * sizeof(cpx<float>) : C::lds_bytes();
  }
  // the DFT-16 operand table of the SW >= 2 kernels, after the tile
  static constexpr size_t w16_lds_off = tw_lds_off;
  static constexpr size_t lds_bytes_sw(int sw) {
    return sw >= 2 ? w16_lds_off + kW16LdsBytes : C::lds_bytes();
  }
};

// The HBM loads of one tile (tile < tiles0 * C1) into f, issued and not
// waited for.
template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM>
__device__ __forceinline__ void mf_fetch(DescPtr d, const void *in, int conj, long long tile,
                                         cpx<float> (&f)[MfShape<N, TX, NT, IO, LINES, RL, XM>::F]) {
  using T = float;
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  using M = MfMap<N, TX, NT, LINES>;
  constexpr int LDT = C::LDT, LDJ = C::LDJ, RPT = C::RPT, E = M::E, GW = M::GW;
  constexpr int R2 = RL::r[2], NB2 = N / R2, L2 = RL::L(2);
  static_assert(RL::P == 3 && RL::r[0] == 16 && RL::r[1] == 16 && R2 <= 8, "16 x 16 x R2");
  static_assert(NB2 % RPT == 0 && (NB2 / RPT) * R2 == E, "VALU pass covers the tile");
  static_assert(NT % 64 == 0, "whole waves");

  const long long tiles0 = d->tiles0;
  const long long c1 = tile / tiles0;
  const long long c0b = (tile - c1 * tiles0) * TX;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, i16 = lane & 15, sq = lane >> 4;
  const int t0 = LINES ? wave / M::NG : i16 % TX;
  const int b0 = LINES ? (wave % M::NG) * 16 + i16 : wave * (16 / TX) + i16 / TX;
  const int tv = LINES ? tid / RPT : tid % TX;
  const int rv = LINES ? tid % RPT : tid / TX;
  const void *tw = d->tw;

  // column t of the tile: first pencil column, validity and element bases
  auto col = [&](int t, long long &ca, bool &cvalid, long long &ib, long long &ob)
      __attribute__((always_inline)) {
    const long long c0 = c0b + t;
    ca = 2 * c0;
    cvalid = (IO == IO_C2C) ? (c0 < d->C0) : (ca < d->C0r);
    ib = (IO == IO_R2C ? ca : c0) * d->in_s0 + c1 * d->in_s1 +
         (IO == IO_C2C && !LINES ? (c0 >> d->bsh) * d->in_b0 : 0);
    ob = (IO == IO_C2R ? ca : c0) * d->out_s0 + c1 * d->out_s1 +
         (IO == IO_C2C && !LINES ? (c0 >> d->bsh) * d->out_b0 : 0);
  };
  // LDS element of point j of tile column t
  auto at = [&](int j, int t) __attribute__((always_inline)) { return C::row(j) * LDJ + t * LDT; };

  constexpr int NB16 = N / 16;
  if constexpr (IO == IO_C2R) {
    constexpr int NH = N / 2 + 1;
    constexpr int ITER = (TX * NH + NT - 1) / NT;
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int w = tid + it * NT;
      f[it] = cpx<T>{T(0), T(0)};
      f[ITER + it] = cpx<T>{T(0), T(0)};
      if (w < TX * NH) {
        const int tt = LINES ? w / NH : w % TX;
        const int k = LINES ? w % NH : w / TX;
        const long long cc = 2 * (c0b + tt);
        long long ia = cc * d->in_s0 + c1 * d->in_s1 + (long long)k * d->in_sj;
        long long ib = ia + d->in_s0;
        if (LINES && d->npack) {  // receive-buffer order (fused unpack; x stages)
          ia = pack_index(d, k, cc + c1 * d->pack_l1);
          ib = pack_index(d, k, cc + 1 + c1 * d->pack_l1);
        }
        if (cc < d->C0r) f[it] = gload<T>(in, ia);
        if (cc + 1 < d->C0r) f[ITER + it] = gload<T>(in, ib);
      }
    }
  } else {
    static_assert(IO == IO_C2C || IO == IO_R2C, "c2c, r2c or c2r");
    // Point j = b0 + 4 sq N/16 (lane part) + BQ q + e N/16 (compile-time part)
    // of column t: the lane part of the offset is computed once, the rest is a
    // uniform multiple of the point stride.  A column past the pencil's edge
    // reads column 0 of the tile instead (always present) and its values are
    // never stored, so the loads need no branches; the second real column
    // of an r2c pair is zeroed when absent.
    const long long sj = d->in_sj;
    const int jl = b0 + 4 * sq * NB16;
#pragma unroll
    for (int q = 0; q < GW; ++q) {
      long long ca, ib, ob;
      bool cv;
      const int t = t0 + M::TQ * q;
      col(t, ca, cv, ib, ob);
      bool second = (IO == IO_R2C) && (ca + 1 < d->C0r);
      if (!cv) {
        col(0, ca, cv, ib, ob);
        second = false;
      }
      const long long lb = ib + (long long)jl * sj;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long long o = lb + (long long)(M::BQ * q + e * NB16) * sj;
        cpx<T> x;
        if constexpr (IO == IO_R2C) {
          x.x = gload_r<T>(in, o);
          x.y = gload_r<T>(in, o + (second ? d->in_s0 : 0));
          if (!second) x.y = T(0);
        } else {
          x = gload<T>(in, o);
          if (conj) x.y = -x.y;
        }
        f[4 * q + e] = x;
      }
    }
  }

  (void)conj;
}

// The rest of one tile from its fetched values f: (c2r: the half spectra
// into LDS,) three passes with two LDS exchanges, the stores.  TWL: the
// twiddles of passes 1 and 2 are read from LDS past the tile (loaded there
// by the caller) instead of from HBM.
template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int SW, bool TWL>
__device__ __forceinline__ void mf_run(DescPtr d, void *out, int conj, long long tile,
                                       cpx<float> (&f)[MfShape<N, TX, NT, IO, LINES, RL, XM>::F],
                                       const Split3 &ar, const Split3 &ai) {
  using S = MfShape<N, TX, NT, IO, LINES, RL, XM>;
  using T = float;
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  using M = MfMap<N, TX, NT, LINES>;
  constexpr int LDT = C::LDT, LDJ = C::LDJ, RPT = C::RPT, E = M::E, GW = M::GW;
  constexpr int R2 = RL::r[2], NB2 = N / R2, L2 = RL::L(2);
  static_assert(RL::P == 3 && RL::r[0] == 16 && RL::r[1] == 16 && R2 <= 8, "16 x 16 x R2");
  static_assert(NB2 % RPT == 0 && (NB2 / RPT) * R2 == E, "VALU pass covers the tile");
  static_assert(NT % 64 == 0, "whole waves");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(lds_raw);
  T *ldr = reinterpret_cast<T *>(lds_raw);

  const long long tiles0 = d->tiles0;
  const long long c1 = tile / tiles0;
  const long long c0b = (tile - c1 * tiles0) * TX;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, i16 = lane & 15, sq = lane >> 4;
  const int t0 = LINES ? wave / M::NG : i16 % TX;
  const int b0 = LINES ? (wave % M::NG) * 16 + i16 : wave * (16 / TX) + i16 / TX;
  const int tv = LINES ? tid / RPT : tid % TX;
  const int rv = LINES ? tid % RPT : tid / TX;
  const void *tw = d->tw;

  // column t of the tile: first pencil column, validity and element bases
  auto col = [&](int t, long long &ca, bool &cvalid, long long &ib, long long &ob)
      __attribute__((always_inline)) {
    const long long c0 = c0b + t;
    ca = 2 * c0;
    cvalid = (IO == IO_C2C) ? (c0 < d->C0) : (ca < d->C0r);
    ib = (IO == IO_R2C ? ca : c0) * d->in_s0 + c1 * d->in_s1 +
         (IO == IO_C2C && !LINES ? (c0 >> d->bsh) * d->in_b0 : 0);
    ob = (IO == IO_C2R ? ca : c0) * d->out_s0 + c1 * d->out_s1 +
         (IO == IO_C2C && !LINES ? (c0 >> d->bsh) * d->out_b0 : 0);
  };
  // LDS element of point j of tile column t
  auto at = [&](int j, int t) __attribute__((always_inline)) { return C::row(j) * LDJ + t * LDT; };

  const cpx<T> *ldtw = reinterpret_cast<const cpx<T> *>(lds_raw + S::tw_lds_off);
  static_assert(SW < 2 || !TWL, "the operand table and LDS twiddles share the space after the tile");
  const u32x4 *w16tab = reinterpret_cast<const u32x4 *>(lds_raw + S::w16_lds_off) + lane;
  constexpr int NB16 = N / 16;
  cpx<T> v[E];
  if constexpr (IO == IO_C2R) {
    constexpr int NH = N / 2 + 1;
    constexpr int ITER = (TX * NH + NT - 1) / NT;
#pragma unroll
    for (int half = 0; half < (XM ? 2 : 1); ++half) {
      if (half) __syncthreads();
#pragma unroll
      for (int it = 0; it < ITER; ++it) {
        const int w = tid + it * NT;
        if (w < TX * NH) {
          const int tt = LINES ? w / NH : w % TX;
          const int k = LINES ? w % NH : w / TX;
          cpx<T> a = f[it], b = f[ITER + it];
          if (k == 0 || 2 * k == N) {
            a.y = T(0);
            b.y = T(0);
          }
          const cpx<T> zk = {a.x - b.y, -(a.y + b.x)}, zm = {a.x + b.y, a.y - b.x};
          const bool mirror = k > 0 && 2 * k < N;
          if constexpr (XM == 0) {
            lds[at(k, tt)] = zk;
            if (mirror) lds[at(N - k, tt)] = zm;
          } else {
            ldr[at(k, tt)] = half ? zk.y : zk.x;
            if (mirror) ldr[at(N - k, tt)] = half ? zm.y : zm.x;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < GW; ++q) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int a = at(b0 + M::BQ * q + (4 * sq + e) * NB16, t0 + M::TQ * q);
          if constexpr (XM == 0) {
            v[4 * q + e] = lds[a];
          } else {
            if (half) v[4 * q + e].y = ldr[a]; else v[4 * q + e].x = ldr[a];
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = f[i];
  }

  // ---- pass 0: radix 16, L = 1 (no twiddles) ---------------------------------
#pragma unroll
  for (int q = 0; q < GW; ++q) {
    cpx<T> x[4] = {v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]}, y[4];
    mf_dft16<SW>(x, y, ar, ai, w16tab);
#pragma unroll
    for (int e = 0; e < 4; ++e) v[4 * q + e] = y[e];
  }

  // ---- exchange 0 -> 1: outputs at 16 b + s, inputs at b + s N/16 -------------
  if constexpr (IO == IO_C2R) __syncthreads();
#pragma unroll
  for (int half = 0; half < (XM ? 2 : 1); ++half) {
    if (half) __syncthreads();
#pragma unroll
    for (int q = 0; q < GW; ++q) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int a = at(16 * (b0 + M::BQ * q) + 4 * sq + e, t0 + M::TQ * q);
        if constexpr (XM == 0) lds[a] = v[4 * q + e];
        else ldr[a] = half ? v[4 * q + e].y : v[4 * q + e].x;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < GW; ++q) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int a = at(b0 + M::BQ * q + (4 * sq + e) * NB16, t0 + M::TQ * q);
        if constexpr (XM == 0) {
          v[4 * q + e] = lds[a];
        } else {
          if (half) v[4 * q + e].y = ldr[a]; else v[4 * q + e].x = ldr[a];
        }
      }
    }
  }

  // ---- pass 1: radix 16, L = 16; butterfly b of every group has b%16 = b0%16 --
  {
    const int k1 = b0 & 15;
    cpx<T> w1[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int s = 4 * sq + e;
      const int ti = RL::tw_off(1) + k1 * 15 + (s > 0 ? s - 1 : 0);
      const cpx<T> wl = TWL ? ldtw[ti - S::TW0] : gload<T>(tw, ti);
      w1[e] = s > 0 ? wl : cpx<T>{T(1), T(0)};
    }
#pragma unroll
    for (int q = 0; q < GW; ++q) {
      cpx<T> x[4], y[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) x[e] = cmul(v[4 * q + e], w1[e]);
      mf_dft16<SW>(x, y, ar, ai, w16tab);
#pragma unroll
      for (int e = 0; e < 4; ++e) v[4 * q + e] = y[e];
    }
  }

  // ---- columns: the last pass straight from the matrix-core layout ------------
  // With groups of 16 butterflies 16 apart (BQ = 16) and as many groups per
  // lane as the last radix, group q of a lane is radix-16 butterfly
  // b0 + 16 q, whose output s sits at 256 q + b0 + 16 s: the last pass's
  // butterfly b0 + 16 s takes output s of the lane's groups q = 0 .. R2-1,
  // all held by the lane.  No exchange; each lane runs the last-pass
  // butterflies b0 + 16 (4 sq + e) and stores their points + 256 q.
  constexpr bool DIRECT = !LINES && GW == R2 && M::BQ == 16 && IO != IO_R2C;
  if constexpr (DIRECT) {
    cpx<T> z[4 * R2];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int bp = b0 + 16 * (4 * sq + e);  // < L2 = 256
      cpx<T> x[R2];
#pragma unroll
      for (int q = 0; q < R2; ++q) x[q] = v[4 * q + e];
#pragma unroll
      for (int s = 1; s < R2; ++s)
        x[s] = cmul(x[s], gload<T>(tw, RL::tw_off(2) + bp * (R2 - 1) + s - 1));
      dft<T, R2>(x);
#pragma unroll
      for (int s = 0; s < R2; ++s) z[e * R2 + s] = x[s];
    }
    long long ca, ib, ob;
    bool cv;
    col(t0, ca, cv, ib, ob);
    if (cv) {
      const long long sj = d->out_sj;
      const long long lb = ob + (long long)(b0 + 64 * sq) * sj;
      const bool second = (IO == IO_C2R) && (ca + 1 < d->C0r);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
#pragma unroll
        for (int s = 0; s < R2; ++s) {
          const long long o = lb + (long long)(16 * e + L2 * s) * sj;
          cpx<T> y = z[e * R2 + s];
          if constexpr (IO == IO_C2R) {
            gstore_r<T>(out, o, y.x);
            if (second) gstore_r<T>(out, o + d->out_s0, -y.y);
          } else {
            if (conj) y.y = -y.y;
            gstore<T>(out, o, y);
          }
        }
      }
    }
    return;
  }

  // ---- exchange 1 -> 2: outputs at 256 (b/16) + b%16 + 16 s; VALU-map inputs --
  __syncthreads();
#pragma unroll
  for (int half = 0; half < (XM ? 2 : 1); ++half) {
    if (half) __syncthreads();
#pragma unroll
    for (int q = 0; q < GW; ++q) {
      const int b = b0 + M::BQ * q;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int a = at(256 * (b >> 4) + (b & 15) + 16 * (4 * sq + e), t0 + M::TQ * q);
        if constexpr (XM == 0) lds[a] = v[4 * q + e];
        else ldr[a] = half ? v[4 * q + e].y : v[4 * q + e].x;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB2 / RPT; ++q) {
#pragma unroll
      for (int s = 0; s < R2; ++s) {
        const int a = at(rv + q * RPT + s * NB2, tv);
        if constexpr (XM == 0) {
          v[q * R2 + s] = lds[a];
        } else {
          if (half) v[q * R2 + s].y = ldr[a]; else v[q * R2 + s].x = ldr[a];
        }
      }
    }
  }

  // ---- pass 2: radix R2 on the VALU, L = 256 ----------------------------------
#pragma unroll
  for (int q = 0; q < NB2 / RPT; ++q) {
    const int b = rv + q * RPT;  // < NB2 = L2: k = b, g = 0
#pragma unroll
    for (int s = 1; s < R2; ++s)
{
      const int ti = RL::tw_off(2) + (b % L2) * (R2 - 1) + s - 1;
      v[q * R2 + s] = cmul(v[q * R2 + s], TWL ? ldtw[ti - S::TW0] : gload<T>(tw, ti));
    }
    dft<T, R2>(&v[q * R2]);
  }

  // ---- outputs: point b + s*256 of column tv ------------------------------------
  if constexpr (IO == IO_R2C) {
    constexpr int NH = N / 2 + 1;
    constexpr int ITER = (TX * NH + NT - 1) / NT;
    cpx<T> oa[ITER], ob[ITER];
    const T half_ = T(0.5);
#pragma unroll
    for (int half = 0; half < (XM ? 2 : 1); ++half) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NB2 / RPT; ++q) {
#pragma unroll
        for (int s = 0; s < R2; ++s) {
          const int a = at(rv + q * RPT + s * L2, tv);
          if constexpr (XM == 0) lds[a] = v[q * R2 + s];
          else ldr[a] = half ? v[q * R2 + s].y : v[q * R2 + s].x;
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
          const int a1 = at(k, tt), a2 = at(km, tt);
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
            oa_i = pack_index(d, k, cc + c1 * d->pack_l1);
            ob_i = pack_index(d, k, cc + 1 + c1 * d->pack_l1);
          }
          gstore<T>(out, oa_i, oa[it]);
          if (cc + 1 < d->C0r) gstore<T>(out, ob_i, ob[it]);
        }
      }
    }
  } else {
    // point j = rv + q RPT + s L2 of column tv: one branch for the column,
    // the lane part of the offset once, the rest uniform
    long long ca, ib, ob;
    bool cv;
    col(tv, ca, cv, ib, ob);
    if (cv) {
      const long long sj = d->out_sj;
      const long long lb = ob + (long long)rv * sj;
      const bool second = (IO == IO_C2R) && (ca + 1 < d->C0r);
#pragma unroll
      for (int q = 0; q < NB2 / RPT; ++q) {
#pragma unroll
        for (int s = 0; s < R2; ++s) {
          const long long o = lb + (long long)(q * RPT + s * L2) * sj;
          cpx<T> y = v[q * R2 + s];
          if constexpr (IO == IO_C2R) {
            gstore_r<T>(out, o, y.x);
            if (second) gstore_r<T>(out, o + d->out_s0, -y.y);
          } else {
            if (conj) y.y = -y.y;
            gstore<T>(out, o, y);
          }
        }
      }
    }
  }
}


// One tile of a stage, by the whole workgroup: every load issued, then the
// passes and stores.
template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int SW>
__device__ __forceinline__ void mf_tile(DescPtr d, const void *in, void *out, int conj, long long tile) {
  cpx<float> f[MfShape<N, TX, NT, IO, LINES, RL, XM>::F];
  mf_fetch<N, TX, NT, IO, LINES, RL, XM>(d, in, conj, tile, f);
  // every load of the tile in flight before the first butterfly: without
  // this the scheduler sinks each load next to its first use, and the stage
  // loses most of its memory-level parallelism
  __builtin_amdgcn_sched_barrier(0);
  Split3 ar, ai;
  if constexpr (SW >= 2) {
    // the operand table into LDS (read before the first butterfly, which
    // the barrier orders)
    extern __shared__ __align__(16) unsigned char lds_raw[];
    using S = MfShape<N, TX, NT, IO, LINES, RL, XM>;
    u32x4 *tab = reinterpret_cast<u32x4 *>(lds_raw + S::w16_lds_off);
    for (int e = threadIdx.x; e < 6 * 64; e += NT) {
      const int part = e / 64, lane = e % 64;
      tab[e] = *reinterpret_cast<const u32x4 *>(&kW16P.w[lane][part / 3][part % 3][0]);
    }
    __syncthreads();
  } else {
    w16_operands(threadIdx.x & 63, ar, ai);
  }
  mf_run<N, TX, NT, IO, LINES, RL, XM, SW, false>(d, out, conj, tile, f, ar, ai);
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE, int SW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
    mf_axis_kernel(const AxisDesc *__restrict__ d, const void *in, void *out, int conj) {
  const long long ntiles = (long long)d->tiles0 * d->C1;
  const long long tile = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, ntiles);
  if (tile >= ntiles) return;  // whole block, before any barrier
  mf_tile<N, TX, NT, IO, LINES, RL, XM, SW>(desc_ptr(d), in, out, conj, tile);
}

// The same stage kernel over the tiles of several grids.
template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE, int SW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
    mf_multi_kernel(const MultiGroup *__restrict__ g, int ng, long long tiles) {
  const long long it = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, tiles);
  if (it >= tiles) return;  // whole block, before any barrier
  const int gi = find_multi_group(g, ng, it);
  mf_tile<N, TX, NT, IO, LINES, RL, XM, SW>(desc_ptr(g[gi].d), g[gi].in, g[gi].out, g[gi].conj,
                                        it - g[gi].tile0);
}

// ---- registry -----------------------------------------------------------------

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE, int SW>
static int launch_mf(const AxisDesc *dd, const AxisDesc &h, const AxisIo &io, hipStream_t st) {
  using S = MfShape<N, TX, NT, IO, LINES, RL, XM>;
  const long long ntiles = xcd_grid((long long)h.tiles0 * h.C1);
  if (ntiles <= 0) return D2DHIP_SUCCESS;
  const long long gx = ntiles < (1LL << 20) ? ntiles : (1LL << 20);
  const long long gy = (ntiles + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "axis stage has too many tiles");
  hipLaunchKernelGGL((mf_axis_kernel<N, TX, NT, IO, LINES, RL, XM, WPE, SW>), dim3((unsigned)gx, (unsigned)gy),
                     dim3(NT), S::lds_bytes_sw(SW), st, dd, io.in, io.out, io.conj);
  return hip_check(hipGetLastError(), "mf_axis_kernel launch");
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE, int SW>
static hipError_t prepare_mf() {
  using S = MfShape<N, TX, NT, IO, LINES, RL, XM>;
  return hipFuncSetAttribute((const void *)mf_axis_kernel<N, TX, NT, IO, LINES, RL, XM, WPE, SW>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::lds_bytes_sw(SW));
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE, int SW>
static int launch_mf_multi(const MultiGroup *g, int ng, long long tiles, hipStream_t st) {
  using S = MfShape<N, TX, NT, IO, LINES, RL, XM>;
  const long long nb = xcd_grid(tiles);
  if (nb <= 0) return D2DHIP_SUCCESS;
  const long long gx = nb < (1LL << 20) ? nb : (1LL << 20);
  const long long gy = (nb + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "launch has too many tiles");
  hipLaunchKernelGGL((mf_multi_kernel<N, TX, NT, IO, LINES, RL, XM, WPE, SW>), dim3((unsigned)gx, (unsigned)gy),
                     dim3(NT), S::lds_bytes_sw(SW), st, g, ng, tiles);
  return hip_check(hipGetLastError(), "mf_multi_kernel launch");
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE, int SW>
static hipError_t prepare_mf_multi() {
  using S = MfShape<N, TX, NT, IO, LINES, RL, XM>;
  return hipFuncSetAttribute((const void *)mf_multi_kernel<N, TX, NT, IO, LINES, RL, XM, WPE, SW>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::lds_bytes_sw(SW));
}

template <int... Rs>
static std::vector<int> mf_rvec(RList<Rs...>) {
  return {Rs...};
}

template <int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE, int SW>
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
  v.lds_bytes = MfShape<N, TX, NT, IO, LINES, RL, XM>::lds_bytes_sw(SW);
  v.ldj = C::LDJ;
  v.ldt = C::LDT;
  v.radices = mf_rvec(RL{});
  v.launch = &launch_mf<N, TX, NT, IO, LINES, RL, XM, WPE, SW>;
  v.prepare = &prepare_mf<N, TX, NT, IO, LINES, RL, XM, WPE, SW>;
  v.launch_multi = &launch_mf_multi<N, TX, NT, IO, LINES, RL, XM, WPE, SW>;
  v.prepare_multi = &prepare_mf_multi<N, TX, NT, IO, LINES, RL, XM, WPE, SW>;
  v.name = name;
  return v;
}

// c2c, r2c and c2r of one shape; XM = 1 adds " xs", WPE > 1 " wN", SW = 1
// or 3 " sw", SW = 2 or 3 " al" to the name
#define D2D_MF_VARIANTS(N, TX, NT, LINES, RL, XM, WPE, SW, TAG)                                         \
  make_mf<N, TX, NT, IO_C2C, LINES, RL, XM, WPE, SW>("mfma n=" #N " tx=" #TX " nt=" #NT " c2c" TAG),   \
      make_mf<N, TX, NT, IO_R2C, LINES, RL, XM, WPE, SW>("mfma n=" #N " tx=" #TX " nt=" #NT " r2c" TAG), \
      make_mf<N, TX, NT, IO_C2R, LINES, RL, XM, WPE, SW>("mfma n=" #N " tx=" #TX " nt=" #NT " c2r" TAG)

using M16_16_4 = RList<16, 16, 4>;
using M16_16_2 = RList<16, 16, 2>;

std::vector<FastVariant> mf_variants() {
  return {
      // 1024 = 16 * 16 * 4: strided columns
      D2D_MF_VARIANTS(1024, 8, 512, false, M16_16_4, 0, 4, 0, ""),
      D2D_MF_VARIANTS(1024, 8, 512, false, M16_16_4, 1, 6, 1, " xs w6 sw"),
      D2D_MF_VARIANTS(1024, 8, 1024, false, M16_16_4, 0, 8, 1, " w8 sw"),
      D2D_MF_VARIANTS(1024, 16, 1024, false, M16_16_4, 1, 4, 0, " xs"),
      // the DFT-16 operands read from LDS next to their MFMAs (al)
      D2D_MF_VARIANTS(1024, 16, 1024, false, M16_16_4, 1, 8, 3, " xs w8 sw al"),
      D2D_MF_VARIANTS(1024, 16, 1024, false, M16_16_4, 1, 8, 2, " xs w8 al"),
      D2D_MF_VARIANTS(1024, 8, 512, false, M16_16_4, 1, 6, 2, " xs w6 al"),
      // contiguous lines
      D2D_MF_VARIANTS(1024, 4, 256, true, M16_16_4, 0, 4, 0, ""),
      D2D_MF_VARIANTS(1024, 4, 256, true, M16_16_4, 1, 6, 1, " xs w6 sw"),
      D2D_MF_VARIANTS(1024, 4, 512, true, M16_16_4, 0, 8, 1, " w8 sw"),
      D2D_MF_VARIANTS(1024, 8, 512, true, M16_16_4, 1, 6, 1, " xs w6 sw"),
      // 512 = 16 * 16 * 2
      D2D_MF_VARIANTS(512, 16, 1024, false, M16_16_2, 1, 8, 1, " xs w8 sw"),
      D2D_MF_VARIANTS(512, 8, 512, false, M16_16_2, 0, 4, 0, ""),
      D2D_MF_VARIANTS(512, 8, 256, true, M16_16_2, 0, 4, 0, ""),
      D2D_MF_VARIANTS(512, 8, 512, true, M16_16_2, 0, 8, 1, " w8 sw"),
  };
}

}  // namespace d2dhip
