// fast_impl.h -- the specialised axis-stage kernels (templates), their
// launchers and the building blocks of the variant registry.  Included by
// fast_kernels.hip (the persistent batch tiles and the registry) and by the
// fast_variants_*.hip files, each of which instantiates its own share of the
// registered variants, so that they compile in parallel.
#pragma once

// Compile-time specialised axis stages for hot lengths.
//
// Same stage contract as axis_fft_kernel (fft_kernels.hip): one workgroup
// owns a tile of TX columns of length N, the first Stockham pass reads HBM,
// the last writes HBM, the passes in between exchange through LDS.  Here N,
// TX, the block size NT and the radix sequence are template parameters, so
// every register index, LDS offset, twiddle offset and k = b mod L / g = b
// div L is resolved at compile time, each thread keeps all of its values of
// a pass in registers, and LDS is addressed with ds_read/ds_write.
//
// Thread mapping: strided axes (COLS) put consecutive threads on
// consecutive tile columns (t fastest: one 16*TX-byte run per point row);
// contiguous lines (LINES) put consecutive threads on consecutive points of
// one line (b fastest: 64 consecutive points per wave).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "batch_loop.h"
#include "codelets.h"
#include "fast_common.h"
#include "fast_kernels.h"
#include "internal.h"

namespace d2dhip {

// Cross-lane word swaps of gfx950 (v_permlane32_swap_b32,
// v_permlane16_swap_b32), on every 32-bit word of a complex value:
// swap32 trades lanes 32-63 of a for lanes 0-31 of b; swap16 trades the odd
// 16-lane rows of a (lanes 16-31, 48-63) for the even rows of b (0-15,
// 32-47).  Register moves only, no LDS.
template <typename T>
__device__ __forceinline__ void swap32(cpx<T> &a, cpx<T> &b) {
  constexpr int W = sizeof(cpx<T>) / 4;
  unsigned wa[W], wb[W];
  __builtin_memcpy(wa, &a, sizeof(a));
  __builtin_memcpy(wb, &b, sizeof(b));
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const auto r = __builtin_amdgcn_permlane32_swap(wa[i], wb[i], false, false);
    wa[i] = r[0];
    wb[i] = r[1];
  }
  __builtin_memcpy(&a, wa, sizeof(a));
  __builtin_memcpy(&b, wb, sizeof(b));
}
template <typename T>
__device__ __forceinline__ void swap16(cpx<T> &a, cpx<T> &b) {
  constexpr int W = sizeof(cpx<T>) / 4;
  unsigned wa[W], wb[W];
  __builtin_memcpy(wa, &a, sizeof(a));
  __builtin_memcpy(wb, &b, sizeof(b));
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const auto r = __builtin_amdgcn_permlane16_swap(wa[i], wb[i], false, false);
    wa[i] = r[0];
    wb[i] = r[1];
  }
  __builtin_memcpy(&a, wa, sizeof(a));
  __builtin_memcpy(&b, wb, sizeof(b));
}

// The four lanes l, l+16, l+32, l+48 (rows 0-3 of the wave) each hold 16
// values as four blocks of 4; afterwards row r holds block r of every row,
// row g's block in its register block g: u[4g + i] = (row g's v)[4r + i].
// Two rounds of swaps, 8 value pairs each, the same register pattern in
// every lane.
// Then value q of block s is input s of butterfly q: reordered to v[4q + s].
template <typename T, int E>
__device__ __forceinline__ void block_transpose4(cpx<T> (&v)[E]) {
  if constexpr (E == 16) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      swap32(v[i], v[8 + i]);       // rows 2,3 of block 0 <-> rows 0,1 of block 2
      swap32(v[4 + i], v[12 + i]);  // rows 2,3 of block 1 <-> rows 0,1 of block 3
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      swap16(v[i], v[4 + i]);       // odd rows of block 0 <-> even rows of block 1
      swap16(v[8 + i], v[12 + i]);  // odd rows of block 2 <-> even rows of block 3
    }
    cpx<T> u[16];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int s = 0; s < 4; ++s) u[4 * q + s] = v[4 * s + q];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = u[i];
  }
}

// WPE: minimum waves per SIMD the register allocation must allow (1 = no
// constraint beyond the block size).
// One tile of a stage (tile < tiles0 * C1), by the whole workgroup.
template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM>
__device__ __forceinline__ void fast_tile(DescPtr d, const void *in, void *out,
                                          int conj, long long tile) {
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  constexpr int P = C::P, RPT = C::RPT, LDT = C::LDT, LDJ = C::LDJ, E = C::emax();
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(lds_raw);

  const long long tiles0 = d->tiles0;
  const long long c1 = tile / tiles0;
  const long long c0b = (tile - c1 * tiles0) * TX;
  const int tid = threadIdx.x;
  const int t = LINES ? tid / RPT : tid % TX;
  const int r0 = LINES ? tid % RPT : tid / TX;
  const long long c0 = c0b + t;
  const long long ca = 2 * c0;  // first real column of a pair (real IO)
  constexpr bool DTT = IO == IO_DTT2 || IO == IO_DTT3 || IO == IO_DTTX || IO == IO_DTT4;
  constexpr bool PAIRED = IO != IO_C2C && IO != IO_DTT4 && IO != IO_YPK && IO != IO_YUN &&
                          IO != IO_YUP;  // two real columns per tile column
  static_assert(!DTT || XM == 0, "real-to-real stages use the plain exchange");
  const bool cvalid = PAIRED ? (ca < d->C0r) : (c0 < d->C0);
  // real-to-real stages: the transformed points start at dtt_first
  const long long f = DTT ? d->dtt_first : 0;
  // (blocked columns: c2c column stages of the blocked spectral layout)
  const long long in_base =
      ((IO == IO_R2C || DTT) && PAIRED ? ca : c0) * d->in_s0 + c1 * d->in_s1 + f * d->in_sj +
      (IO == IO_C2C && !LINES ? (c0 >> d->bsh) * d->in_b0 : 0);
  const long long out_base =
      ((IO == IO_C2R || DTT) && PAIRED ? ca : c0) * d->out_s0 + c1 * d->out_s1 + f * d->out_sj +
      (IO == IO_C2C && !LINES ? (c0 >> d->bsh) * d->out_b0 : 0);
  const void *tw = d->tw;
  cpx<T> v[E];

  // XM = 2 (columns of 1024 = 16 x 16 x 4, 16 values per thread): the first
  // exchange as XM = 1, the last inside the wave.  The radix-4 butterfly
  // b = k + 16 s' (k < 16, s' < 16) takes output s' of the radix-16
  // butterflies k + 16 g, g < 4; those four run in lanes l, l+16, l+32,
  // l+48 of one wave (rows gq = 0..3; k from the wave and the row within
  // its quarter, spread so that the pass-1 LDS reads hit distinct banks), and
  // block_transpose4 hands each its four butterflies 16 (4 gq + q) + k.
  constexpr bool SHUF = XM == 2;
  constexpr int HQ = SHUF ? 64 / TX / 4 : 1;  // tile rows per quarter wave
  static_assert(!SHUF || (!LINES && P == 3 && RL::r[0] == 16 && RL::r[1] == 16 && RL::r[2] == 4 &&
                          RPT == 64 && (HQ == 1 || HQ == 2 || HQ == 4) &&
                          (IO == IO_C2C || IO == IO_YPK || IO == IO_YUN || IO == IO_YUP)),
                "in-wave last exchange: c2c columns of 1024 points, 4 <= TX <= 16");
  const int wq = tid >> 6, rq = (tid & 63) / TX;
  const int gq = rq / HQ, kq = (wq & 3) + 4 * HQ * (wq >> 2) + 4 * (rq % HQ);

  // ---- inputs of pass 0 -------------------------------------------------
  if constexpr (IO == IO_DTT4) {
    // v_J = (x_{2J} + i x_{2N-1-2J}) exp(-i pi (4J+1) / (8N)) of one line of 2N
    // points; DST-IV takes (-1)^j x_j
    const bool alt = d->dtt == 8;
    constexpr int R = RL::r[0], NBR = N / R;
#pragma unroll
    for (int q = 0; q < C::nb(0); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(0) || b < NBR) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int J = b + s * NBR;
          cpx<T> x = {T(0), T(0)};
          if (cvalid) {
            x.x = gload_r<T>(in, in_base + (long long)(2 * J) * d->in_sj);
            x.y = gload_r<T>(in, in_base + (long long)(2 * N - 1 - 2 * J) * d->in_sj);
            if (alt) x.y = -x.y;
          }
          v[q * R + s] = cmul(x, gload<T>(d->dtw, J));
        }
      }
    }
  } else if constexpr (IO == IO_DTTX) {
    // the extension e of the line pair, gathered: e[J] = +-x_src or 0
    const int type = d->dtt, Lx = d->dtt_len;
    constexpr int R = RL::r[0], NBR = N / R;
#pragma unroll
    for (int q = 0; q < C::nb(0); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(0) || b < NBR) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int J = b + s * NBR;
          int src = -1;
          bool neg = false;
          if (type == 1) {  // e[j] = e[N-j] = x_j
            src = J < Lx ? J : N - J;
          } else if (type == 5) {  // e[j+1] = -e[N-j-1] = x_j
            if (J >= 1 && J <= Lx) src = J - 1;
            else if (J >= Lx + 2) { src = N - 1 - J; neg = true; }
          } else if (J & 1) {  // types 4, 8: e[2j+1] = +-e[N-2j-1] = x_j
            if (J < 2 * Lx) src = J >> 1;
            else if (J > N - 2 * Lx) { src = (N - 1 - J) >> 1; neg = type == 8; }
          }
          cpx<T> x = {T(0), T(0)};
          if (src >= 0) {
            const long long o = in_base + (long long)src * d->in_sj;
            if (ca < d->C0r) x.x = gload_r<T>(in, o);
            if (ca + 1 < d->C0r) x.y = gload_r<T>(in, o + d->in_s0);
            if (neg) x = cpx<T>{-x.x, -x.y};
          }
          v[q * R + s] = x;
        }
      }
    }
  } else if constexpr (IO == IO_DTT2) {
    // v[J] = x_{2J} (J < ceil(N/2)), x_{2N-1-2J} (otherwise); DST-II takes
    // (-1)^j x_j
    const bool alt = d->dtt == 6;
    constexpr int R = RL::r[0], NBR = N / R;
#pragma unroll
    for (int q = 0; q < C::nb(0); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(0) || b < NBR) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int J = b + s * NBR;
          const int j = (2 * J < N) ? 2 * J : 2 * N - 1 - 2 * J;
          const long long o = in_base + (long long)j * d->in_sj;
          cpx<T> x = {T(0), T(0)};
          if (ca < d->C0r) x.x = gload_r<T>(in, o);
          if (ca + 1 < d->C0r) x.y = gload_r<T>(in, o + d->in_s0);
          if (alt && (j & 1)) x = cpx<T>{-x.x, -x.y};
          v[q * R + s] = x;
        }
      }
    }
  } else if constexpr (IO == IO_DTT3) {
    // conj(Z_j) = w_j conj(u_j), u_j = (Xa_j + Xb_{N-j}, Xb_j - Xa_{N-j})
    // (X_N = 0; DST-III reads X backwards), for j and N - j together
    constexpr int NJ = N / 2 + 1;
    constexpr int ITER = (TX * NJ + NT - 1) / NT;
    const bool rev = d->dtt == 7;
    cpx<T> A[ITER], B[ITER];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int w = tid + it * NT;
      A[it] = cpx<T>{T(0), T(0)};
      B[it] = cpx<T>{T(0), T(0)};
      if (w < TX * NJ) {
        const int tt = LINES ? w / NJ : w % TX;
        const int j = LINES ? w % NJ : w / TX;
        const long long cc = 2 * (c0b + tt);
        const long long base = cc * d->in_s0 + c1 * d->in_s1 + f * d->in_sj;
        const long long o1 = base + (long long)(rev ? N - 1 - j : j) * d->in_sj;
        const long long o2 = base + (long long)(rev ? j - 1 : N - j) * d->in_sj;
        if (cc < d->C0r) {
          A[it].x = gload_r<T>(in, o1);
          if (j > 0) B[it].x = gload_r<T>(in, o2);
        }
        if (cc + 1 < d->C0r) {
          A[it].y = gload_r<T>(in, o1 + d->in_s0);
          if (j > 0) B[it].y = gload_r<T>(in, o2 + d->in_s0);
        }
      }
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int w = tid + it * NT;
      if (w < TX * NJ) {
        const int tt = LINES ? w / NJ : w % TX;
        const int j = LINES ? w % NJ : w / TX;
        const int jm = N - j;
        const cpx<T> a = A[it], b = B[it];
        const cpx<T> cs = gload<T>(d->dtw, j);
        const T ux = a.x + b.y, uy = a.y - b.x;
        lds[C::row(j) * LDJ + tt * LDT] = cpx<T>{cs.x * ux - cs.y * uy, -(cs.x * uy + cs.y * ux)};
        if (j > 0 && j < jm) {
          const cpx<T> cm = gload<T>(d->dtw, jm);
          const T vx = b.x + a.y, vy = b.y - a.x;
          lds[C::row(jm) * LDJ + tt * LDT] =
              cpx<T>{cm.x * vx - cm.y * vy, -(cm.x * vy + cm.y * vx)};
        }
      }
    }
    __syncthreads();
    constexpr int R = RL::r[0], NBR = N / R;
#pragma unroll
    for (int q = 0; q < C::nb(0); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(0) || b < NBR) {
#pragma unroll
        for (int s = 0; s < R; ++s) v[q * R + s] = lds[C::row(b + s * NBR) * LDJ + t * LDT];
      }
    }
  } else if constexpr (IO == IO_C2R) {
    // conj(Z) of the merged half spectra, all N points, into LDS
    constexpr int NH = N / 2 + 1;
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
        long long ia = cc * d->in_s0 + c1 * d->in_s1 + (long long)k * d->in_sj;
        long long ib = ia + d->in_s0;
        if (LINES && d->npack) {  // receive-buffer order (fused unpack; x stages)
          ia = pack_index(d, k, cc + c1 * d->pack_l1);
          ib = pack_index(d, k, cc + 1 + c1 * d->pack_l1);
        }
        if (cc < d->C0r) A[it] = gload<T>(in, ia);
        if (cc + 1 < d->C0r) B[it] = gload<T>(in, ib);
      }
    }
    // conj(Z) of the merged half spectra into LDS, then the inputs of pass 0
    // (both halves in turn for the split exchange)
    constexpr int R = RL::r[0], NBR = N / R;
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
            T *ldr = reinterpret_cast<T *>(lds_raw);
            ldr[C::row(k) * LDJ + tt * LDT] = half ? zk.y : zk.x;
            if (mirror) ldr[C::row(N - k) * LDJ + tt * LDT] = half ? zm.y : zm.x;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < C::nb(0); ++q) {
        const int b = r0 + q * RPT;
        if (C::exact(0) || b < NBR) {
#pragma unroll
          for (int s = 0; s < R; ++s) {
            const int a = C::row(b + s * NBR) * LDJ + t * LDT;
            if constexpr (XM == 0) {
              v[q * R + s] = lds[a];
            } else {
              const T w = reinterpret_cast<const T *>(lds_raw)[a];
              if (half) v[q * R + s].y = w; else v[q * R + s].x = w;
            }
          }
        }
      }
    }
  } else {
    constexpr int R = RL::r[0], NBR = N / R;
#pragma unroll
    for (int q = 0; q < C::nb(0); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(0) || b < NBR) {
#pragma unroll
        for (int s = 0; s < R; ++s)
          v[q * R + s] = fload<T, IO>(d, in, in_base, cvalid, true, ca, b + s * NBR, conj, c1);
      }
    }
  }

  // ---- passes -------------------------------------------------------------
  static_for<0, P>([&](auto pc) __attribute__((always_inline)) {
    constexpr int p = decltype(pc)::value;
    constexpr int R = RL::r[p], NBR = N / R, L = RL::L(p);
    if constexpr (SHUF && p == P - 1) {
      block_transpose4(v);  // radix-16 outputs -> radix-4 inputs, in the wave
    } else if constexpr (p > 0) {
      // outputs of pass p-1 -> LDS -> inputs of pass p
      constexpr int Rq = RL::r[p - 1], NBq = N / Rq, Lq = RL::L(p - 1);
      if constexpr (p > 1 || IO == IO_C2R || IO == IO_DTT3) __syncthreads();
      if constexpr (XM == 0) {
#pragma unroll
        for (int q = 0; q < C::nb(p - 1); ++q) {
          const int b = r0 + q * RPT;
          if (C::exact(p - 1) || b < NBq) {
            const int k = b % Lq, g = b / Lq;
#pragma unroll
            for (int s = 0; s < Rq; ++s)
              lds[C::row(g * Lq * Rq + k + s * Lq) * LDJ + t * LDT] = v[q * Rq + s];
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < C::nb(p); ++q) {
          const int b = r0 + q * RPT;
          if (C::exact(p) || b < NBR) {
#pragma unroll
            for (int s = 0; s < R; ++s) v[q * R + s] = lds[C::row(b + s * NBR) * LDJ + t * LDT];
          }
        }
      } else {
        // real parts, then imaginary parts; .x of the new layout is filled
        // before .y of the old layout is written, and the two never alias
        T *ldr = reinterpret_cast<T *>(lds_raw);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          if (half) __syncthreads();
#pragma unroll
          for (int q = 0; q < C::nb(p - 1); ++q) {
            const int b = r0 + q * RPT;
            if (C::exact(p - 1) || b < NBq) {
              const int k = b % Lq, g = b / Lq;
#pragma unroll
              for (int s = 0; s < Rq; ++s)
                ldr[C::row(g * Lq * Rq + k + s * Lq) * LDJ + t * LDT] =
                    half ? v[q * Rq + s].y : v[q * Rq + s].x;
            }
          }
          __syncthreads();
#pragma unroll
          for (int q = 0; q < C::nb(p); ++q) {
            const int b = (SHUF && p == P - 2) ? kq + 16 * gq : r0 + q * RPT;
            if (C::exact(p) || b < NBR) {
#pragma unroll
              for (int s = 0; s < R; ++s) {
                const T w = ldr[C::row(b + s * NBR) * LDJ + t * LDT];
                if (half) v[q * R + s].y = w; else v[q * R + s].x = w;
              }
            }
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < C::nb(p); ++q) {
      const int b = !SHUF ? r0 + q * RPT : p == P - 2 ? kq + 16 * gq : p == P - 1 ? kq + 16 * (4 * gq + q) : r0 + q * RPT;
      if (C::exact(p) || b < NBR) {
        if constexpr (L > 1) {
          const int k = b % L;
          const int tp = RL::tw_off(p) + k * (R - 1);
#pragma unroll
          for (int s = 1; s < R; ++s) v[q * R + s] = cmul(v[q * R + s], gload<T>(tw, tp + s - 1));
        }
        dft<T, R>(&v[q * R]);
      }
    }
  });

  // ---- outputs of the last pass ---------------------------------------------
  constexpr int RZ = RL::r[P - 1], NBZ = N / RZ, LZ = RL::L(P - 1);
  if constexpr (IO == IO_DTT2) {
    // Z -> LDS; Y_k = 2 Re(w_q V_q) of each line of the pair, V split from Z
    // as in a real-to-complex transform, q = k (DCT-II) or N-1-k (DST-II)
    if constexpr (P > 1) __syncthreads();
#pragma unroll
    for (int q = 0; q < C::nb(P - 1); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(P - 1) || b < NBZ) {
        const int k = b % LZ, g = b / LZ;
#pragma unroll
        for (int s = 0; s < RZ; ++s) lds[C::row(g * LZ * RZ + k + s * LZ) * LDJ + t * LDT] = v[q * RZ + s];
      }
    }
    __syncthreads();
    const bool flip = d->dtt == 6;
    constexpr int ITER = (TX * N + NT - 1) / NT;
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int w = tid + it * NT;
      if (w < TX * N) {
        const int tt = LINES ? w / N : w % TX;
        const int k = LINES ? w % N : w / TX;
        const long long cc = 2 * (c0b + tt);
        if (cc < d->C0r) {
          const int q = flip ? N - 1 - k : k;
          const cpx<T> zk = lds[C::row(q) * LDJ + tt * LDT];
          const cpx<T> zm = lds[C::row(q == 0 ? 0 : N - q) * LDJ + tt * LDT];
          const cpx<T> cs = gload<T>(d->dtw, q);
          const long long o = cc * d->out_s0 + c1 * d->out_s1 + (f + k) * d->out_sj;
          gstore_r<T>(out, o, cs.x * (zk.x + zm.x) + cs.y * (zk.y - zm.y));
          if (cc + 1 < d->C0r)
            gstore_r<T>(out, o + d->out_s0, cs.x * (zk.y + zm.y) + cs.y * (zm.x - zk.x));
        }
      }
    }
  } else if constexpr (IO == IO_DTT4) {
    // u_k = V_k exp(-i pi k / (2N)); DCT-IV: Y_{2k} = 2 Re u_k, Y_{2N-1-2k} =
    // -2 Im u_k; DST-IV the same outputs read backwards
    const bool alt = d->dtt == 8;
#pragma unroll
    for (int q = 0; q < C::nb(P - 1); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(P - 1) || b < NBZ) {
        const int k = b % LZ, g = b / LZ;
#pragma unroll
        for (int s = 0; s < RZ; ++s) {
          const int pv = g * LZ * RZ + k + s * LZ;
          const cpx<T> u = cmul(v[q * RZ + s], gload<T>(d->dtw, N + pv));
          const int ie = alt ? 2 * N - 1 - 2 * pv : 2 * pv, io = alt ? 2 * pv : 2 * N - 1 - 2 * pv;
          if (cvalid) {
            gstore_r<T>(out, out_base + (long long)ie * d->out_sj, T(2) * u.x);
            gstore_r<T>(out, out_base + (long long)io * d->out_sj, T(-2) * u.y);
          }
        }
      }
    }
  } else if constexpr (IO == IO_DTTX) {
    // Y_k from E[p] (E_a + i E_b): DCT-I p = k, DST-I p = k + 1, DCT-IV and
    // DST-IV p = 2k + 1; cosines take (Re, Im), sines (-Im, Re)
    const int type = d->dtt, Lx = d->dtt_len;
#pragma unroll
    for (int q = 0; q < C::nb(P - 1); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(P - 1) || b < NBZ) {
        const int k = b % LZ, g = b / LZ;
#pragma unroll
        for (int s = 0; s < RZ; ++s) {
          const int pv = g * LZ * RZ + k + s * LZ;
          const int kk = type == 1 ? pv : type == 5 ? pv - 1 : ((pv & 1) ? pv >> 1 : -1);
          if (kk >= 0 && kk < Lx) {
            const cpx<T> z = v[q * RZ + s];
            const T a = type <= 4 ? z.x : -z.y, bb = type <= 4 ? z.y : z.x;
            const long long o = out_base + (long long)kk * d->out_sj;
            if (ca < d->C0r) gstore_r<T>(out, o, a);
            if (ca + 1 < d->C0r) gstore_r<T>(out, o + d->out_s0, bb);
          }
        }
      }
    }
  } else if constexpr (IO == IO_DTT3) {
    // z = conj(last pass) = v_a + i v_b; y_{2n} = v_n, y_{2n+1} = v_{N-1-n}
    // (DST-III: times (-1)^k)
    const bool alt = d->dtt == 7;
#pragma unroll
    for (int q = 0; q < C::nb(P - 1); ++q) {
      const int b = r0 + q * RPT;
      if (C::exact(P - 1) || b < NBZ) {
        const int k = b % LZ, g = b / LZ;
#pragma unroll
        for (int s = 0; s < RZ; ++s) {
          const int pv = g * LZ * RZ + k + s * LZ;
          const int kk = (2 * pv < N) ? 2 * pv : 2 * N - 1 - 2 * pv;
          const cpx<T> r = v[q * RZ + s];
          const bool neg = alt && (kk & 1);
          const long long o = out_base + (long long)kk * d->out_sj;
          if (ca < d->C0r) gstore_r<T>(out, o, neg ? -r.x : r.x);
          if (ca + 1 < d->C0r) gstore_r<T>(out, o + d->out_s0, neg ? r.y : -r.y);
        }
      }
    }
  } else if constexpr (IO == IO_R2C) {
    // the last pass's outputs Z -> LDS; A[k] = (Z[k] + conj Z[N-k]) / 2 and
    // B[k] = (Z[k] - conj Z[N-k]) / (2i) for k <= N/2.  Re A and Im B need
    // only real parts of Z, Im A and Re B only imaginary parts, so the split
    // exchange builds them in two rounds.
    constexpr int NH = N / 2 + 1;
    constexpr int ITER = (TX * NH + NT - 1) / NT;
    cpx<T> oa[ITER], ob[ITER];
    const T half_ = T(0.5);
#pragma unroll
    for (int half = 0; half < (XM ? 2 : 1); ++half) {
      if (P > 1 || half) __syncthreads();
#pragma unroll
      for (int q = 0; q < C::nb(P - 1); ++q) {
        const int b = r0 + q * RPT;
        if (C::exact(P - 1) || b < NBZ) {
          const int k = b % LZ, g = b / LZ;
#pragma unroll
          for (int s = 0; s < RZ; ++s) {
            const int a = C::row(g * LZ * RZ + k + s * LZ) * LDJ + t * LDT;
            if constexpr (XM == 0)
              lds[a] = v[q * RZ + s];
            else
              reinterpret_cast<T *>(lds_raw)[a] = half ? v[q * RZ + s].y : v[q * RZ + s].x;
          }
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
            const T *ldr = reinterpret_cast<const T *>(lds_raw);
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
    static_assert(P >= 2, "single-pass lengths go through the general kernel");
#pragma unroll
    for (int q = 0; q < C::nb(P - 1); ++q) {
      const int b = SHUF ? kq + 16 * (4 * gq + q) : r0 + q * RPT;
      if (C::exact(P - 1) || b < NBZ) {
        const int k = b % LZ, g = b / LZ;
#pragma unroll
        for (int s = 0; s < RZ; ++s)
          fstore<T, IO>(d, out, out_base, cvalid, ca, g * LZ * RZ + k + s * LZ, v[q * RZ + s], conj, c1);
      }
    }
  }
  if constexpr (DTT) {
    // the points outside [dtt_first, dtt_first + dtt_len) are copied through
    const int npt = d->dtt_n - d->dtt_len, fi = d->dtt_first, Lt = d->dtt_len;
    for (int w = tid; w < TX * npt; w += NT) {
      const int tt = LINES ? w / npt : w % TX;
      const int i = LINES ? w % npt : w / TX;
      const int k = i < fi ? i : i + Lt;
      const long long cc = PAIRED ? 2 * (c0b + tt) : c0b + tt;
      if (cc < d->C0r) {
        const long long ib = cc * d->in_s0 + c1 * d->in_s1 + (long long)k * d->in_sj;
        const long long ob = cc * d->out_s0 + c1 * d->out_s1 + (long long)k * d->out_sj;
        gstore_r<T>(out, ob, gload_r<T>(in, ib));
        if (PAIRED && cc + 1 < d->C0r)
          gstore_r<T>(out, ob + d->out_s0, gload_r<T>(in, ib + d->in_s0));
      }
    }
  }
}

template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
    fast_axis_kernel(const AxisDesc *__restrict__ d, const void *in, void *out, int conj) {
  const long long ntiles = (long long)d->tiles0 * d->C1;
  const long long tile = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, ntiles);
  if (tile >= ntiles) return;  // whole block, before any barrier
  fast_tile<T, N, TX, NT, IO, LINES, RL, XM>(desc_ptr(d), in, out, conj, tile);
}

// The same stage kernel over the tiles of several grids.
template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
    fast_multi_kernel(const MultiGroup *__restrict__ g, int ng, long long tiles) {
  const long long it = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, tiles);
  if (it >= tiles) return;  // whole block, before any barrier
  const int gi = find_multi_group(g, ng, it);
  fast_tile<T, N, TX, NT, IO, LINES, RL, XM>(desc_ptr(g[gi].d), g[gi].in, g[gi].out, g[gi].conj,
                                             it - g[gi].tile0);
}

// ---- registry -----------------------------------------------------------------

template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static int launch_fast(const AxisDesc *dd, const AxisDesc &h, const AxisIo &io, hipStream_t st) {
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  const long long ntiles = xcd_grid((long long)h.tiles0 * h.C1);
  if (ntiles <= 0) return D2DHIP_SUCCESS;
  const long long gx = ntiles < (1LL << 20) ? ntiles : (1LL << 20);
  const long long gy = (ntiles + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "axis stage has too many tiles");
  hipLaunchKernelGGL((fast_axis_kernel<T, N, TX, NT, IO, LINES, RL, XM, WPE>), dim3((unsigned)gx, (unsigned)gy),
                     dim3(NT), C::lds_bytes(), st, dd, io.in, io.out, io.conj);
  return hip_check(hipGetLastError(), "fast_axis_kernel launch");
}

template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static hipError_t prepare_fast() {
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  return hipFuncSetAttribute((const void *)fast_axis_kernel<T, N, TX, NT, IO, LINES, RL, XM, WPE>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::lds_bytes());
}

template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static int launch_fast_multi(const MultiGroup *g, int ng, long long tiles, hipStream_t st) {
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  const long long nb = xcd_grid(tiles);
  if (nb <= 0) return D2DHIP_SUCCESS;
  const long long gx = nb < (1LL << 20) ? nb : (1LL << 20);
  const long long gy = (nb + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "launch has too many tiles");
  hipLaunchKernelGGL((fast_multi_kernel<T, N, TX, NT, IO, LINES, RL, XM, WPE>),
                     dim3((unsigned)gx, (unsigned)gy), dim3(NT), C::lds_bytes(), st, g, ng, tiles);
  return hip_check(hipGetLastError(), "fast_multi_kernel launch");
}

template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM, int WPE>
static hipError_t prepare_fast_multi() {
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  return hipFuncSetAttribute((const void *)fast_multi_kernel<T, N, TX, NT, IO, LINES, RL, XM, WPE>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::lds_bytes());
}

template <int... Rs>
static std::vector<int> rvec(RList<Rs...>) {
  return {Rs...};
}

template <typename T, int N, int TX, int NT, int IO, bool LINES, class RL, int XM = 0, int WPE = 1>
static FastVariant make_variant(const char *name) {
  using C = FastCfg<T, N, TX, NT, LINES, RL, XM>;
  FastVariant v;
  v.rb = (int)sizeof(T);
  v.io = IO;
  v.lines = LINES ? 1 : 0;
  v.n = N;
  v.TX = TX;
  v.NT = NT;
  v.lds_bytes = C::lds_bytes();
  v.ldj = C::LDJ;
  v.ldt = C::LDT;
  v.radices = rvec(RL{});
  v.launch = &launch_fast<T, N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.prepare = &prepare_fast<T, N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.launch_multi = &launch_fast_multi<T, N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.prepare_multi = &prepare_fast_multi<T, N, TX, NT, IO, LINES, RL, XM, WPE>;
  v.name = name;
  return v;
}

#define D2D_VARIANTS(T, N, TX, NT, LINES, RL)                                             \
  make_variant<T, N, TX, NT, IO_C2C, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " c2c"), \
      make_variant<T, N, TX, NT, IO_R2C, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " r2c"), \
      make_variant<T, N, TX, NT, IO_C2R, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " c2r")

// real-to-real stages: DCT-II / DST-II and DCT-III / DST-III of N points,
// DCT-I / IV, DST-I / IV through an extension of N points, and DCT-IV /
// DST-IV of 2N points
#define D2D_DTT(T, N, TX, NT, LINES, RL)                                                    \
  make_variant<T, N, TX, NT, IO_DTT2, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " dct2"), \
      make_variant<T, N, TX, NT, IO_DTT3, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " dct3"), \
      make_variant<T, N, TX, NT, IO_DTTX, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " dctx"), \
      make_variant<T, N, TX, NT, IO_DTT4, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " dct4")

// the y- / z-pencil stage with its output in pack order (" ypk"), with its
// input in unpack order (" yun"), and with both (" yup"), twins of the c2c
// variant of the same name
#define D2D_YPACK(T, N, TX, NT, RL, XM, WPE, TAG)                                                     \
  make_variant<T, N, TX, NT, IO_YPK, false, RL, XM, WPE>(#T " n=" #N " tx=" #TX " nt=" #NT " ypk" TAG), \
      make_variant<T, N, TX, NT, IO_YUN, false, RL, XM, WPE>(#T " n=" #N " tx=" #TX " nt=" #NT " yun" TAG), \
      make_variant<T, N, TX, NT, IO_YUP, false, RL, XM, WPE>(#T " n=" #N " tx=" #TX " nt=" #NT " yup" TAG)

// c2c only, split exchange, at least WPE waves per SIMD
#define D2D_SPLIT_C2C(T, N, TX, NT, LINES, RL, WPE)                                      \
  make_variant<T, N, TX, NT, IO_C2C, LINES, RL, 1, WPE>(#T " n=" #N " tx=" #TX " nt=" #NT \
                                                        " c2c xs w" #WPE)

// c2c columns, split first exchange and in-wave last exchange, at least WPE
// waves per SIMD
#define D2D_SHUF_C2C(T, N, TX, NT, RL, WPE)                                             \
  make_variant<T, N, TX, NT, IO_C2C, false, RL, 2, WPE>(#T " n=" #N " tx=" #TX " nt=" #NT \
                                                        " c2c xsh w" #WPE)

// c2c, r2c and c2r, split exchange, at least WPE waves per SIMD
#define D2D_SPLIT_ALL(T, N, TX, NT, LINES, RL, WPE)                                           \
  D2D_SPLIT_C2C(T, N, TX, NT, LINES, RL, WPE),                                                \
      make_variant<T, N, TX, NT, IO_R2C, LINES, RL, 1, WPE>(#T " n=" #N " tx=" #TX " nt=" #NT \
                                                            " r2c xs w" #WPE),               \
      make_variant<T, N, TX, NT, IO_C2R, LINES, RL, 1, WPE>(#T " n=" #N " tx=" #TX " nt=" #NT \
                                                            " c2r xs w" #WPE)

using R16_16_4 = RList<16, 16, 4>;
using R16_16_8 = RList<16, 16, 8>;    // 2048
using R8_8_8 = RList<8, 8, 8>;
using R16_9_5 = RList<16, 9, 5>;     // 720
using R4_9_3_5 = RList<4, 9, 3, 5>;  // 540
using R16_9_3 = RList<16, 9, 3>;     // 432
using R16_8_3 = RList<16, 8, 3>;     // 384
using R16_8_5 = RList<16, 8, 5>;     // 640
using R16_16_3 = RList<16, 16, 3>;   // 768
using R8_8_5 = RList<8, 8, 5>;       // 320
using R16_9_4 = RList<16, 9, 4>;     // 576
using R16_10_5 = RList<16, 10, 5>;   // 800
using R16_12_5 = RList<16, 12, 5>;   // 960
using R10_10_10 = RList<10, 10, 10>; // 1000
using R16_9_8 = RList<16, 9, 8>;     // 1152
using R16_16_5 = RList<16, 16, 5>;   // 1280
using R16_16_6 = RList<16, 16, 6>;   // 1536
using R16_16 = RList<16, 16>;         // 256
using R16_8 = RList<16, 8>;           // 128
using R8_8 = RList<8, 8>;             // 64
using R8_4 = RList<8, 4>;             // 32
// Lengths 2^a 3^b 5^c of the uneven-grid batches: radices chosen so that
// every pass has a whole number of butterflies per row of the tile
// (N / r divisible by NT / TX), i.e. no idle lanes, at 12-24 values per lane.
using R12_4 = RList<12, 4>;           // 48, 4 rows per column
using R4_4_5 = RList<4, 4, 5>;        // 80, 4 rows
using R12_8 = RList<12, 8>;           // 96, 4 rows
using R4_4_10 = RList<4, 4, 10>;      // 160, 8 rows
using R4_4_12 = RList<4, 4, 12>;      // 192, 16 rows (or 8)

// c2c only, plain exchange (the one-wave tiles of the persistent batch kernel)
#define D2D_C2C(T, N, TX, NT, LINES, RL) \
  make_variant<T, N, TX, NT, IO_C2C, LINES, RL>(#T " n=" #N " tx=" #TX " nt=" #NT " c2c")

}  // namespace d2dhip
