// fast_common.h -- compile-time stage shapes and pencil IO shared by the
// specialised axis kernels (fast_impl.h: VALU passes; mfma_kernels.hip:
// radix-16 passes on the matrix cores).
#pragma once

#include <hip/hip_runtime.h>

#include "codelets.h"
#include "internal.h"

namespace d2dhip {

template <int... Rs>
struct RList {
  static constexpr int P = sizeof...(Rs);
  static constexpr int r[P] = {Rs...};
  static constexpr int L(int p) {
    int l = 1;
    for (int i = 0; i < p; ++i) l *= r[i];
    return l;
  }
  static constexpr int tw_off(int p) {
    int o = 0;
    for (int i = 0; i < p; ++i) o += L(i) * (r[i] - 1);
    return o;
  }
  static constexpr int prod() { return L(P); }
};

// XM = 0: the LDS exchange between passes moves complex values; XM = 1: it
// moves the real parts, then the imaginary parts, through a buffer of half
// the size (two more barriers per exchange), so that twice as many
// workgroups fit a CU's LDS.
template <typename T, int N, int TX, int NT, bool LINES, class RL, int XM = 0>
struct FastCfg {
  static constexpr int P = RL::P;
  static constexpr int RPT = NT / TX;  // butterfly rows per tile column
  static_assert(NT % TX == 0, "NT must be a multiple of TX");
  static_assert(RL::prod() == N, "radices must multiply to N");
  static constexpr int nb(int p) { return (N / RL::r[p] + RPT - 1) / RPT; }
  static constexpr bool exact(int p) { return (N / RL::r[p]) % RPT == 0; }
  static constexpr int emax() {
    int e = 0;
    for (int p = 0; p < P; ++p) e = e > nb(p) * RL::r[p] ? e : nb(p) * RL::r[p];
    return e;
  }
  // LDS row of point j: one pad row after every 16, so that the stride-R
  // writes of the first exchange spread over all banks.
  static constexpr int row(int j) { return j + (j >> 4); }
  static constexpr int ROWS = row(N - 1) + 1;
  static constexpr int LDT = LINES ? (ROWS | 1) : 1;
  static constexpr int LDJ = LINES ? 1 : TX;
  static constexpr size_t lds_bytes() {
    return (size_t)(LINES ? TX * LDT : TX * ROWS) * (XM ? sizeof(T) : sizeof(cpx<T>));
  }
};

// base: the column's plain element offset; c1: its outer column (used by
// the y-pencil buffer forms).
template <typename T, int IO>
__device__ __forceinline__ cpx<T> fload(DescPtr d, const void *in,
                                        long long base, bool cvalid, bool bvalid, long long ca,
                                        int j, int conj, long long c1 = 0) {
  if constexpr (IO == IO_YUN || IO == IO_YUP) {
    cpx<T> v = {T(0), T(0)};
    if (cvalid) v = gload<T>(in, base + (long long)j * d->in_sj + packj_offset(d, j, c1));
    if (conj) v.y = -v.y;
    return v;
  } else if constexpr (IO == IO_R2C) {
    const long long o = base + (long long)j * d->in_sj;
    cpx<T> v = {T(0), T(0)};
    if (ca < d->C0r) v.x = gload_r<T>(in, o);
    if (ca + 1 < d->C0r) v.y = gload_r<T>(in, o + d->in_s0);
    return v;
  } else {
    cpx<T> v = {T(0), T(0)};
    if (cvalid) v = gload<T>(in, base + (long long)j * d->in_sj);
    if (conj) v.y = -v.y;
    return v;
  }
  (void)bvalid;
}

template <typename T, int IO>
__device__ __forceinline__ void fstore(DescPtr d, void *out, long long base,
                                       bool cvalid, long long ca, int j, cpx<T> v, int conj,
                                       long long c1 = 0) {
  if constexpr (IO == IO_YPK || IO == IO_YUP) {
    if (cvalid) {
      if (conj) v.y = -v.y;
      const long long o = IO == IO_YPK ? packj_offset(d, j, c1) : packj_offset2(d, j, c1);
      gstore<T>(out, base + (long long)j * d->out_sj + o, v);
    }
  } else if constexpr (IO == IO_C2R) {
    const long long o = base + (long long)j * d->out_sj;
    if (ca < d->C0r) gstore_r<T>(out, o, v.x);
    if (ca + 1 < d->C0r) gstore_r<T>(out, o + d->out_s0, -v.y);
  } else {
    if (cvalid) {
      if (conj) v.y = -v.y;
      gstore<T>(out, base + (long long)j * d->out_sj, v);
    }
  }
}

}  // namespace d2dhip
