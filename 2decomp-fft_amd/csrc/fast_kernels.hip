// fast_kernels.hip -- compile-time specialised axis stages for hot lengths.
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

// ---- persistent batch kernel on specialised tiles ---------------------------------
// c2c tiles for the lengths of uneven-grid batches, on strided columns and
// contiguous lines, all of one block size per precision.  Every kind has
// the radix list of the registered variants of its length, so its results
// equal the plans' own executions on any specialised kernel bit for bit.
//   FP64: tiles of up to 35 KB (4 blocks, 8 waves per CU), at least 8
//         columns (128-byte rows) on strided axes, 2-wave blocks;
//   FP32: the same tile bytes, twice the columns (at least 16: 128-byte
//         rows), 4-wave blocks (16 waves per CU; 2.95 -> 2.78 ms on the
//         64-grid batch against 2-wave blocks).  FP64 on 4-wave blocks gave
//         results that differ in the last bits from the standalone kernels
//         of the same tiles, so it keeps 2-wave blocks.
// Some kinds leave lanes idle in a pass (N / r not a multiple of the rows
// per column).
template <int N_, int TX_, bool LINES_, class RL_>
struct PV {
  static constexpr int N = N_, TX = TX_;
  static constexpr bool LINES = LINES_;
  using RL = RL_;
};
template <int NT_, int WPE_, class... Vs>
struct PSet {
  static constexpr int NT = NT_, WPE = WPE_;
  static constexpr int size = sizeof...(Vs);
};
using PersistSetD =
    PSet<128, 2, PV<256, 8, false, R16_16>, PV<256, 8, true, R16_16>, PV<192, 8, false, R4_4_12>,
         PV<192, 8, true, R4_4_12>, PV<160, 8, false, R4_4_10>, PV<160, 8, true, R4_4_10>,
         PV<128, 16, false, R16_8>, PV<128, 16, true, R16_8>, PV<96, 16, false, R12_8>,
         PV<96, 16, true, R12_8>, PV<80, 16, false, R4_4_5>, PV<80, 16, true, R4_4_5>,
         PV<64, 16, false, R8_8>, PV<64, 16, true, R8_8>, PV<48, 32, false, R12_4>,
         PV<48, 32, true, R12_4>, PV<32, 32, false, R8_4>, PV<32, 32, true, R8_4>>;
using PersistSetF =
    PSet<256, 4, PV<256, 16, false, R16_16>, PV<256, 16, true, R16_16>, PV<192, 16, false, R4_4_12>,
         PV<192, 16, true, R4_4_12>, PV<160, 16, false, R4_4_10>, PV<160, 16, true, R4_4_10>,
         PV<128, 32, false, R16_8>, PV<128, 32, true, R16_8>, PV<96, 32, false, R12_8>,
         PV<96, 32, true, R12_8>, PV<80, 32, false, R4_4_5>, PV<80, 32, true, R4_4_5>,
         PV<64, 32, false, R8_8>, PV<64, 32, true, R8_8>, PV<48, 64, false, R12_4>,
         PV<48, 64, true, R12_4>, PV<32, 64, false, R8_4>, PV<32, 64, true, R8_4>>;
template <typename T>
struct PersistSetOf {
  using type = PersistSetD;
};
template <>
struct PersistSetOf<float> {
  using type = PersistSetF;
};

template <int K, class V, class... Vs>
struct PAt {
  using type = typename PAt<K - 1, Vs...>::type;
};
template <class V, class... Vs>
struct PAt<0, V, Vs...> {
  using type = V;
};

// One kind's tile as a function of its own, so that its registers are
// allocated for its body alone (inlined side by side, the bodies of all
// kinds share one allocation and spill).  Function arguments arrive in
// vector registers; they are wave-uniform, and re-deriving them with
// readfirstlane lets the body keep the descriptor and its index arithmetic
// in scalar registers, as a kernel does with its arguments.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// When head is not null, thread 0 also takes the next position of that
// queue counter (after the call's entry wait, so that the atomic's latency
// overlaps the tile's first loads) and returns it.
template <typename T, int NT, class V>
__device__ __noinline__ unsigned long long persist_kind_tile(DescPtr d, const void *in, void *out,
                                                             int conj, long long tile,
                                                             unsigned long long *head) {
  unsigned long long *uhead = (unsigned long long *)uniform_u64((unsigned long long)head);
  unsigned long long r = 0;
  if (uhead && threadIdx.x == 0) r = atomicAdd(uhead, 1ull);
  const DescPtr ud = (DescPtr)uniform_u64((unsigned long long)d);
  const void *uin = (const void *)uniform_u64((unsigned long long)in);
  void *uout = (void *)uniform_u64((unsigned long long)out);
  const int uconj = __builtin_amdgcn_readfirstlane(conj);
  const long long utile = (long long)uniform_u64((unsigned long long)tile);
  fast_tile<T, V::N, V::TX, NT, IO_C2C, V::LINES, typename V::RL, 0>(ud, uin, uout, uconj, utile);
  return r;
}

// Tile of kind `kind` (wave-uniform): a chain of uniform branches.
template <typename T, int NT, int K, class... Vs>
__device__ __forceinline__ unsigned long long persist_tile(int kind, DescPtr d, const void *in,
                                                           void *out, int conj, long long tile,
                                                           unsigned long long *head) {
  if constexpr (K < (int)sizeof...(Vs)) {
    using V = typename PAt<K, Vs...>::type;
    if (kind == K) return persist_kind_tile<T, NT, V>(d, in, out, conj, tile, head);
    return persist_tile<T, NT, K + 1, Vs...>(kind, d, in, out, conj, tile, head);
  }
  return 0;
}

template <typename T, int NT, int WPE, class... Vs>
__device__ __forceinline__ unsigned long long persist_run(const BatchGroup &g, long long tile,
                                                          unsigned long long *head,
                                                          PSet<NT, WPE, Vs...>) {
  const int kind = __builtin_amdgcn_readfirstlane(g.kind);
  return persist_tile<T, NT, 0, Vs...>(kind, desc_ptr(g.d), g.in, g.out, g.conj, tile, head);
}

// The waves-per-SIMD bound also holds for the per-kind functions it calls,
// which would otherwise take up to 256 registers each.
template <typename T>
__global__ void __launch_bounds__(PersistSetOf<T>::type::NT)
    __attribute__((amdgpu_waves_per_eu(PersistSetOf<T>::type::WPE)))
    batch_fast_kernel(const BatchQueue q) {
  batch_loop(q, [&](const BatchGroup &g, long long tile, unsigned long long *head)
                    __attribute__((always_inline)) {
    return persist_run<T>(g, tile, head, typename PersistSetOf<T>::type());
  });
}

template <typename T, int NT, int WPE, class... Vs>
static std::vector<PersistKind> persist_table(PSet<NT, WPE, Vs...>) {
  const char *tname = sizeof(T) == 8 ? "double" : "float";
  std::vector<PersistKind> v;
  auto add = [&](int n, int tx, bool lines, size_t lds) {
    char b[96];
    std::snprintf(b, sizeof(b), "%s n=%d tx=%d nt=%d c2c", tname, n, tx, NT);
    PersistKind k;
    k.name = b;
    k.n = n;
    k.TX = tx;
    k.lines = lines ? 1 : 0;
    k.lds_bytes = lds;
    v.push_back(k);
  };
  (add(Vs::N, Vs::TX, Vs::LINES, FastCfg<T, Vs::N, Vs::TX, NT, Vs::LINES, typename Vs::RL>::lds_bytes()), ...);
  return v;
}

const std::vector<PersistKind> &persist_kinds(int rb) {
  static const std::vector<PersistKind> d = persist_table<double>(PersistSetD());
  static const std::vector<PersistKind> f = persist_table<float>(PersistSetF());
  return rb == 8 ? d : f;
}

int batch_fast_prepare(int rb, size_t lds, int *blocks_per_cu) {
  hipError_t e;
  if (rb == 8) {
    e = hipFuncSetAttribute((const void *)batch_fast_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_fast_kernel<double>,
                                                       PersistSetD::NT, lds);
  } else {
    e = hipFuncSetAttribute((const void *)batch_fast_kernel<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_fast_kernel<float>,
                                                       PersistSetF::NT, lds);
  }
  return hip_check(e, "persistent batch kernel set-up");
}

int launch_batch_fast(int rb, const BatchQueue &q, int blocks, size_t lds, hipStream_t st) {
  const int rst = launch_batch_reset(q, st);
  if (rst) return rst;
  if (rb == 8)
    hipLaunchKernelGGL(batch_fast_kernel<double>, dim3(blocks), dim3(PersistSetD::NT), lds, st, q);
  else
    hipLaunchKernelGGL(batch_fast_kernel<float>, dim3(blocks), dim3(PersistSetF::NT), lds, st, q);
  return hip_check(hipGetLastError(), "persistent batch kernel launch");
}

const std::vector<FastVariant> &fast_registry() {
  static const std::vector<FastVariant> reg = [] {
    std::vector<FastVariant> r = {
      // FP64, 1024: strided columns and contiguous lines
      D2D_VARIANTS(double, 1024, 4, 256, false, R16_16_4),
      D2D_VARIANTS(double, 1024, 2, 128, false, R16_16_4),
      D2D_VARIANTS(double, 1024, 8, 512, false, R16_16_4),
      D2D_VARIANTS(double, 1024, 8, 256, false, R16_16_4),
      D2D_VARIANTS(double, 1024, 1, 64, false, R16_16_4),
      D2D_VARIANTS(double, 1024, 4, 256, true, R16_16_4),
      D2D_VARIANTS(double, 1024, 2, 128, true, R16_16_4),
      // y-pencil stages fused with the pack / unpack of the y-pencil's exchange
      D2D_YPACK(double, 1024, 16, 1024, R16_16_4, 1, 4, " xs w4"),
      D2D_YPACK(double, 1024, 8, 512, R16_16_4, 1, 4, " xs w4"),
      D2D_YPACK(double, 1024, 8, 512, R16_16_4, 0, 1, ""),
      D2D_YPACK(float, 1024, 16, 1024, R16_16_4, 1, 8, " xs w8"),
      D2D_YPACK(float, 1024, 8, 512, R16_16_4, 0, 1, ""),
      D2D_YPACK(double, 1024, 16, 1024, R16_16_4, 2, 4, " xsh w4"),
      D2D_YPACK(double, 1024, 8, 512, R16_16_4, 2, 4, " xsh w4"),
      D2D_YPACK(float, 1024, 16, 1024, R16_16_4, 2, 8, " xsh w8"),
      D2D_YPACK(float, 1024, 8, 512, R16_16_4, 2, 6, " xsh w6"),
      D2D_YPACK(double, 512, 8, 512, R8_8_8, 0, 1, ""),
      D2D_YPACK(float, 512, 16, 512, R8_8_8, 0, 1, ""),
      D2D_SPLIT_C2C(double, 1024, 2, 128, false, R16_16_4, 4),
      D2D_SPLIT_C2C(double, 1024, 4, 256, false, R16_16_4, 4),
      D2D_SPLIT_C2C(double, 1024, 8, 512, false, R16_16_4, 4),
      D2D_SPLIT_C2C(double, 1024, 16, 1024, false, R16_16_4, 4),
      D2D_SHUF_C2C(double, 1024, 8, 512, R16_16_4, 4),
      D2D_SHUF_C2C(double, 1024, 16, 1024, R16_16_4, 4),
      D2D_SPLIT_C2C(double, 1024, 4, 256, false, R16_16_4, 3),
      D2D_SPLIT_C2C(double, 1024, 8, 512, false, R16_16_4, 3),
      D2D_SPLIT_ALL(double, 1024, 2, 128, true, R16_16_4, 3),
      D2D_SPLIT_ALL(double, 1024, 4, 256, true, R16_16_4, 3),
      D2D_SPLIT_ALL(double, 1024, 8, 512, true, R16_16_4, 3),
      // FP64, 512
      // FP64 2048 lines (the x stages of 2048-point pencils; the 2-line
      // tile first: 24-25 ms per 2048^3 stage against 26 ms for 1 line)
      D2D_VARIANTS(double, 2048, 2, 256, true, R16_16_8),
      D2D_SPLIT_ALL(double, 2048, 2, 256, true, R16_16_8, 3),
      D2D_VARIANTS(double, 2048, 1, 128, true, R16_16_8),
      D2D_VARIANTS(double, 512, 8, 512, false, R8_8_8),
      D2D_VARIANTS(double, 512, 4, 256, false, R8_8_8),
      D2D_VARIANTS(double, 512, 4, 256, true, R8_8_8),
      D2D_VARIANTS(double, 512, 8, 512, true, R8_8_8),
      // FP32, 1024 and 512
      D2D_VARIANTS(float, 1024, 8, 512, false, R16_16_4),
      D2D_VARIANTS(float, 1024, 4, 256, false, R16_16_4),
      D2D_VARIANTS(float, 1024, 16, 1024, false, R16_16_4),
      D2D_SPLIT_C2C(float, 1024, 8, 512, false, R16_16_4, 6),
      D2D_SPLIT_C2C(float, 1024, 4, 256, false, R16_16_4, 6),
      D2D_SPLIT_C2C(float, 1024, 16, 1024, false, R16_16_4, 8),
      D2D_SHUF_C2C(float, 1024, 16, 1024, R16_16_4, 8),
      D2D_SHUF_C2C(float, 1024, 8, 512, R16_16_4, 6),
      D2D_SPLIT_ALL(float, 1024, 4, 256, true, R16_16_4, 5),
      D2D_VARIANTS(float, 1024, 8, 512, true, R16_16_4),
      D2D_VARIANTS(float, 1024, 4, 256, true, R16_16_4),
      D2D_VARIANTS(float, 512, 16, 512, false, R8_8_8),
      D2D_VARIANTS(float, 512, 8, 256, true, R8_8_8),
      // FP64, the 720 x 540 x 432 grid's lengths (lines of 720, columns of 540 and 432)
      D2D_VARIANTS(double, 720, 4, 256, true, R16_9_5),
      D2D_VARIANTS(double, 720, 2, 128, true, R16_9_5),
      D2D_VARIANTS(double, 720, 8, 256, false, R16_9_5),
      D2D_VARIANTS(double, 540, 8, 512, false, R4_9_3_5),
      D2D_VARIANTS(double, 540, 4, 256, false, R4_9_3_5),
      D2D_VARIANTS(double, 540, 4, 256, true, R4_9_3_5),
      D2D_VARIANTS(double, 432, 8, 512, false, R16_9_3),
      D2D_VARIANTS(double, 432, 4, 256, false, R16_9_3),
      D2D_VARIANTS(double, 432, 4, 256, true, R16_9_3),
      // common mixed-radix lengths 384 = 16*8*3, 640 = 16*8*5, 768 = 16*16*3
      D2D_VARIANTS(double, 384, 8, 256, false, R16_8_3),
      D2D_VARIANTS(double, 384, 4, 128, true, R16_8_3),
      D2D_VARIANTS(double, 640, 8, 512, false, R16_8_5),
      D2D_VARIANTS(double, 640, 4, 256, false, R16_8_5),
      D2D_VARIANTS(double, 640, 4, 256, true, R16_8_5),
      D2D_VARIANTS(double, 768, 8, 512, false, R16_16_3),
      D2D_VARIANTS(double, 768, 4, 256, false, R16_16_3),
      D2D_VARIANTS(double, 768, 4, 256, true, R16_16_3),
      D2D_VARIANTS(float, 384, 16, 512, false, R16_8_3),
      D2D_VARIANTS(float, 384, 4, 128, true, R16_8_3),
      D2D_VARIANTS(float, 640, 16, 512, false, R16_8_5),
      D2D_VARIANTS(float, 640, 4, 256, true, R16_8_5),
      D2D_VARIANTS(float, 768, 16, 1024, false, R16_16_3),
      D2D_VARIANTS(float, 768, 8, 512, false, R16_16_3),
      D2D_VARIANTS(float, 768, 4, 256, true, R16_16_3),
      // and 320, 576, 800, 960, 1000, 1152
      D2D_VARIANTS(double, 320, 8, 256, false, R8_8_5),
      D2D_VARIANTS(double, 320, 4, 128, true, R8_8_5),
      D2D_VARIANTS(double, 576, 8, 512, false, R16_9_4),
      D2D_VARIANTS(double, 576, 4, 256, false, R16_9_4),
      D2D_VARIANTS(double, 576, 4, 256, true, R16_9_4),
      D2D_VARIANTS(double, 800, 8, 512, false, R16_10_5),
      D2D_VARIANTS(double, 800, 4, 256, false, R16_10_5),
      D2D_VARIANTS(double, 800, 4, 256, true, R16_10_5),
      D2D_VARIANTS(double, 960, 8, 512, false, R16_12_5),
      D2D_VARIANTS(double, 960, 4, 256, false, R16_12_5),
      D2D_VARIANTS(double, 960, 4, 256, true, R16_12_5),
      D2D_VARIANTS(double, 1000, 8, 512, false, R10_10_10),
      D2D_VARIANTS(double, 1000, 4, 256, false, R10_10_10),
      D2D_VARIANTS(double, 1000, 4, 256, true, R10_10_10),
      D2D_VARIANTS(double, 1152, 8, 512, false, R16_9_8),
      D2D_VARIANTS(double, 1152, 4, 256, false, R16_9_8),
      D2D_VARIANTS(double, 1152, 4, 256, true, R16_9_8),
      D2D_VARIANTS(float, 320, 16, 512, false, R8_8_5),
      D2D_VARIANTS(float, 320, 4, 128, true, R8_8_5),
      D2D_VARIANTS(float, 576, 16, 512, false, R16_9_4),
      D2D_VARIANTS(float, 576, 4, 256, true, R16_9_4),
      D2D_VARIANTS(float, 800, 16, 512, false, R16_10_5),
      D2D_VARIANTS(float, 800, 4, 256, true, R16_10_5),
      D2D_VARIANTS(float, 960, 16, 512, false, R16_12_5),
      D2D_VARIANTS(float, 960, 4, 256, true, R16_12_5),
      D2D_VARIANTS(float, 1000, 16, 512, false, R10_10_10),
      D2D_VARIANTS(float, 1000, 4, 256, true, R10_10_10),
      D2D_VARIANTS(float, 1152, 16, 512, false, R16_9_8),
      D2D_VARIANTS(float, 1152, 4, 256, true, R16_9_8),
      // 1280, 1536 and 2048-point columns (the y and z stages of those cubes)
      D2D_VARIANTS(double, 1280, 4, 256, false, R16_16_5),
      D2D_VARIANTS(double, 1280, 4, 256, true, R16_16_5),
      D2D_SPLIT_C2C(double, 1280, 8, 512, false, R16_16_5, 4),
      D2D_VARIANTS(double, 1536, 4, 256, false, R16_16_6),
      D2D_VARIANTS(double, 1536, 2, 256, true, R16_16_6),
      D2D_SPLIT_C2C(double, 1536, 8, 512, false, R16_16_6, 4),
      // (the split-exchange 8-column tile first: the default of plans too
      // large to tune, measured 30-31 ms against 39 ms per 2048^3 FP64 stage)
      D2D_SPLIT_C2C(double, 2048, 8, 1024, false, R16_16_8, 4),
      D2D_VARIANTS(double, 2048, 4, 512, false, R16_16_8),
      D2D_VARIANTS(float, 1280, 8, 512, false, R16_16_5),
      D2D_VARIANTS(float, 1280, 4, 256, true, R16_16_5),
      D2D_VARIANTS(float, 1536, 8, 512, false, R16_16_6),
      D2D_VARIANTS(float, 1536, 4, 256, true, R16_16_6),
      D2D_VARIANTS(float, 2048, 8, 512, false, R16_16_8),
      D2D_VARIANTS(float, 2048, 2, 256, true, R16_16_8),
      // short power-of-two lengths: 128-byte rows on strided axes
      D2D_VARIANTS(double, 256, 8, 128, false, R16_16),
      D2D_VARIANTS(double, 256, 4, 64, true, R16_16),
      D2D_VARIANTS(double, 256, 8, 128, true, R16_16),
      D2D_VARIANTS(double, 128, 8, 64, false, R16_8),
      D2D_VARIANTS(double, 128, 16, 128, false, R16_8),
      D2D_VARIANTS(double, 128, 8, 64, true, R16_8),
      D2D_VARIANTS(double, 64, 8, 64, false, R8_8),
      D2D_VARIANTS(double, 64, 16, 128, true, R8_8),
      D2D_VARIANTS(double, 32, 16, 64, false, R8_4),
      D2D_VARIANTS(double, 32, 32, 128, true, R8_4),
      D2D_VARIANTS(double, 48, 16, 64, false, R12_4),
      D2D_VARIANTS(double, 48, 16, 64, true, R12_4),
      D2D_VARIANTS(double, 80, 16, 64, false, R4_4_5),
      D2D_VARIANTS(double, 80, 16, 64, true, R4_4_5),
      D2D_VARIANTS(double, 96, 16, 64, false, R12_8),
      D2D_VARIANTS(double, 96, 16, 64, true, R12_8),
      D2D_VARIANTS(double, 160, 8, 64, false, R4_4_10),
      D2D_VARIANTS(double, 160, 16, 128, false, R4_4_10),
      D2D_VARIANTS(double, 160, 8, 64, true, R4_4_10),
      D2D_VARIANTS(double, 192, 8, 128, false, R4_4_12),
      D2D_VARIANTS(double, 192, 8, 64, false, R4_4_12),
      D2D_VARIANTS(double, 192, 4, 64, true, R4_4_12),
      D2D_VARIANTS(double, 192, 8, 64, true, R4_4_12),
      // tiles of the persistent batch kernel not registered above
      D2D_C2C(double, 192, 8, 128, true, R4_4_12),
      D2D_C2C(double, 160, 8, 128, false, R4_4_10),
      D2D_C2C(double, 160, 8, 128, true, R4_4_10),
      D2D_C2C(double, 128, 16, 128, true, R16_8),
      D2D_C2C(double, 96, 16, 128, false, R12_8),
      D2D_C2C(double, 96, 16, 128, true, R12_8),
      D2D_C2C(double, 80, 16, 128, false, R4_4_5),
      D2D_C2C(double, 80, 16, 128, true, R4_4_5),
      D2D_C2C(double, 64, 16, 128, false, R8_8),
      D2D_C2C(double, 48, 32, 128, false, R12_4),
      D2D_C2C(double, 48, 32, 128, true, R12_4),
      D2D_C2C(double, 32, 32, 128, false, R8_4),
      D2D_VARIANTS(float, 256, 16, 256, false, R16_16),
      D2D_VARIANTS(float, 256, 4, 64, true, R16_16),
      D2D_VARIANTS(float, 128, 16, 128, false, R16_8),
      D2D_VARIANTS(float, 128, 8, 64, true, R16_8),
      D2D_VARIANTS(float, 64, 16, 128, false, R8_8),
      D2D_VARIANTS(float, 64, 16, 128, true, R8_8),
      // real-to-real stages of (2^k + 1)-point grids (k = 6..10)
      D2D_DTT(double, 1024, 4, 256, false, R16_16_4),
      D2D_DTT(double, 1024, 4, 256, true, R16_16_4),
      D2D_DTT(double, 512, 8, 512, false, R8_8_8),
      D2D_DTT(double, 512, 4, 256, true, R8_8_8),
      D2D_DTT(double, 256, 8, 128, false, R16_16),
      D2D_DTT(double, 256, 4, 64, true, R16_16),
      D2D_DTT(double, 128, 8, 64, false, R16_8),
      D2D_DTT(double, 128, 8, 64, true, R16_8),
      D2D_DTT(double, 64, 8, 64, false, R8_8),
      D2D_DTT(double, 64, 16, 128, true, R8_8),
      D2D_DTT(float, 1024, 8, 512, false, R16_16_4),
      D2D_DTT(float, 1024, 4, 256, true, R16_16_4),
      D2D_DTT(float, 512, 16, 512, false, R8_8_8),
      D2D_DTT(float, 512, 8, 256, true, R8_8_8),
      D2D_DTT(float, 256, 16, 256, false, R16_16),
      D2D_DTT(float, 256, 4, 64, true, R16_16),
      D2D_DTT(float, 128, 16, 128, false, R16_8),
      D2D_DTT(float, 128, 8, 64, true, R16_8),
      D2D_DTT(float, 64, 16, 128, false, R8_8),
      D2D_DTT(float, 64, 16, 128, true, R8_8),
      D2D_VARIANTS(float, 32, 16, 64, false, R8_4),
      D2D_VARIANTS(float, 32, 32, 128, true, R8_4),
      D2D_VARIANTS(float, 48, 16, 64, false, R12_4),
      D2D_VARIANTS(float, 48, 16, 64, true, R12_4),
      D2D_VARIANTS(float, 80, 16, 64, false, R4_4_5),
      D2D_VARIANTS(float, 80, 16, 64, true, R4_4_5),
      D2D_VARIANTS(float, 96, 16, 64, false, R12_8),
      D2D_VARIANTS(float, 96, 16, 64, true, R12_8),
      D2D_VARIANTS(float, 160, 16, 128, false, R4_4_10),
      D2D_VARIANTS(float, 160, 8, 64, true, R4_4_10),
      D2D_VARIANTS(float, 192, 8, 128, false, R4_4_12),
      D2D_VARIANTS(float, 192, 4, 64, true, R4_4_12),
      D2D_C2C(float, 256, 16, 256, true, R16_16),
      D2D_C2C(float, 192, 16, 256, false, R4_4_12),
      D2D_C2C(float, 192, 16, 256, true, R4_4_12),
      D2D_C2C(float, 160, 16, 256, false, R4_4_10),
      D2D_C2C(float, 160, 16, 256, true, R4_4_10),
      D2D_C2C(float, 128, 32, 256, false, R16_8),
      D2D_C2C(float, 128, 32, 256, true, R16_8),
      D2D_C2C(float, 96, 32, 256, false, R12_8),
      D2D_C2C(float, 96, 32, 256, true, R12_8),
      D2D_C2C(float, 80, 32, 256, false, R4_4_5),
      D2D_C2C(float, 80, 32, 256, true, R4_4_5),
      D2D_C2C(float, 64, 32, 256, false, R8_8),
      D2D_C2C(float, 64, 32, 256, true, R8_8),
      D2D_C2C(float, 48, 64, 256, false, R12_4),
      D2D_C2C(float, 48, 64, 256, true, R12_4),
      D2D_C2C(float, 32, 64, 256, false, R8_4),
      D2D_C2C(float, 32, 64, 256, true, R8_4),
    };
    for (const FastVariant &v : mf_variants()) r.push_back(v);
    return r;
  }();
  return reg;
}

}  // namespace d2dhip
