// codelets.h -- small-radix DFT butterflies for the gfx950 Stockham kernels.
//
// Every codelet computes the forward DFT y_s = sum_t v_t exp(-2 pi i s t / R)
// in place on a register array.  Inverse transforms are obtained by
// conjugating on load and on store, so one set of codelets serves both signs.
// Constants come from radix_consts.h (correctly rounded FP64 roots of unity)
// and are folded at compile time: indices are template parameters.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "radix_consts.h"

namespace d2dhip {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

// Device-memory (HBM) accesses through address-space-1 pointers: the
// compiler emits global_load / global_store for them.  A pointer that a
// kernel loads from memory (a descriptor's twiddle table, a grid's buffers
// in a multi-grid launch) is otherwise a generic pointer, and generic
// accesses become flat instructions, which also count against the LDS wait
// counter and so hold up the LDS exchanges between passes.
template <typename T>
using gvec2 = T __attribute__((ext_vector_type(2)));
#define D2D_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ cpx<T> gload(const void *base, long long i) {
  const gvec2<T> v = ((const D2D_GLOBAL gvec2<T> *)base)[i];
  return cpx<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void gstore(void *base, long long i, cpx<T> c) {
  gvec2<T> v;
  v.x = c.x;
  v.y = c.y;
  ((D2D_GLOBAL gvec2<T> *)base)[i] = v;
}
template <typename T>
__device__ __forceinline__ T gload_r(const void *base, long long i) {
  return ((const D2D_GLOBAL T *)base)[i];
}
template <typename T>
__device__ __forceinline__ void gstore_r(void *base, long long i, T v) {
  ((D2D_GLOBAL T *)base)[i] = v;
}

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) {
  return {a.x, -a.y};
}
// -i * a
template <typename T>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) {
  return {a.y, -a.x};
}
// +i * a
template <typename T>
__device__ __forceinline__ cpx<T> mul_pi(cpx<T> a) {
  return {-a.y, a.x};
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// v * exp(-2 pi i M / R), with the trivial angles special-cased.
template <typename T, int R, int M>
__device__ __forceinline__ cpx<T> twc(cpx<T> v) {
  constexpr int m = ((M % R) + R) % R;
  if constexpr (m == 0) {
    return v;
  } else if constexpr (2 * m == R) {
    return {-v.x, -v.y};
  } else if constexpr (4 * m == R) {
    return mul_mi(v);
  } else if constexpr (4 * m == 3 * R) {
    return mul_pi(v);
  } else if constexpr (8 * m == R) {  // (1 - i)/sqrt2
    constexpr T h = (T)kRootCos[kRootOff[8] + 1];
    return {(v.x + v.y) * h, (v.y - v.x) * h};
  } else if constexpr (8 * m == 3 * R) {  // (-1 - i)/sqrt2
    constexpr T h = (T)kRootCos[kRootOff[8] + 1];
    return {(v.y - v.x) * h, -(v.x + v.y) * h};
  } else if constexpr (8 * m == 5 * R) {  // (-1 + i)/sqrt2
    constexpr T h = (T)kRootCos[kRootOff[8] + 1];
    return {-(v.x + v.y) * h, (v.x - v.y) * h};
  } else if constexpr (8 * m == 7 * R) {  // (1 + i)/sqrt2
    constexpr T h = (T)kRootCos[kRootOff[8] + 1];
    return {(v.x - v.y) * h, (v.x + v.y) * h};
  } else {
    constexpr T c = (T)kRootCos[kRootOff[R] + m];
    constexpr T s = (T)(-kRootSin[kRootOff[R] + m]);
    return cmul(v, cpx<T>{c, s});
  }
}

template <typename T, int R>
__device__ __forceinline__ void dft(cpx<T> *v);

// Odd R by the symmetric pairing of t and R-t:
//   y_s     = v0 + sum_t a_t cos(2 pi s t/R) - i b_t sin(2 pi s t/R)
//   y_{R-s} = v0 + sum_t a_t cos(2 pi s t/R) + i b_t sin(2 pi s t/R)
// with a_t = v_t + v_{R-t}, b_t = v_t - v_{R-t}, t = 1..(R-1)/2.
template <typename T, int R>
__device__ __forceinline__ void dft_odd(cpx<T> *v) {
  constexpr int H = (R - 1) / 2;
  cpx<T> a[H], b[H];
  static_for<1, H + 1>([&](auto tc) __attribute__((always_inline)) {
    constexpr int t = decltype(tc)::value;
    a[t - 1] = cadd(v[t], v[R - t]);
    b[t - 1] = csub(v[t], v[R - t]);
  });
  const cpx<T> v0 = v[0];
  cpx<T> y0 = v0;
  static_for<0, H>([&](auto tc) __attribute__((always_inline)) { y0 = cadd(y0, a[decltype(tc)::value]); });
  static_for<1, H + 1>([&](auto sc) __attribute__((always_inline)) {
    constexpr int s = decltype(sc)::value;
    cpx<T> A = v0;
    cpx<T> B = {T(0), T(0)};
    static_for<1, H + 1>([&](auto tc) __attribute__((always_inline)) {
      constexpr int t = decltype(tc)::value;
      constexpr int m = (s * t) % R;
      constexpr T c = (T)kRootCos[kRootOff[R] + m];
      constexpr T sn = (T)kRootSin[kRootOff[R] + m];
      A.x += a[t - 1].x * c;
      A.y += a[t - 1].y * c;
      B.x += b[t - 1].y * sn;
      B.y -= b[t - 1].x * sn;
    });
    v[s] = cadd(A, B);
    v[R - s] = csub(A, B);
  });
  v[0] = y0;
}

// Composite R = R1*R2: input n = R2*n1 + n2, output k = k1 + R1*k2.
//   X[k1 + R1 k2] = sum_n2 W_R2^{n2 k2} W_R^{n2 k1} sum_n1 W_R1^{n1 k1} v[R2 n1 + n2]
template <typename T, int R1, int R2>
__device__ __forceinline__ void dft_comp(cpx<T> *v) {
  constexpr int R = R1 * R2;
  cpx<T> u[R];
  static_for<0, R2>([&](auto n2c) __attribute__((always_inline)) {
    constexpr int n2 = decltype(n2c)::value;
    cpx<T> w[R1];
    static_for<0, R1>([&](auto n1c) __attribute__((always_inline)) { w[decltype(n1c)::value] = v[R2 * decltype(n1c)::value + n2]; });
    dft<T, R1>(w);
    static_for<0, R1>([&](auto k1c) __attribute__((always_inline)) {
      constexpr int k1 = decltype(k1c)::value;
      u[k1 * R2 + n2] = twc<T, R, n2 * k1>(w[k1]);
    });
  });
  static_for<0, R1>([&](auto k1c) __attribute__((always_inline)) {
    constexpr int k1 = decltype(k1c)::value;
    cpx<T> w[R2];
    static_for<0, R2>([&](auto n2c) __attribute__((always_inline)) { w[decltype(n2c)::value] = u[k1 * R2 + decltype(n2c)::value]; });
    dft<T, R2>(w);
    static_for<0, R2>([&](auto k2c) __attribute__((always_inline)) {
      constexpr int k2 = decltype(k2c)::value;
      v[k1 + R1 * k2] = w[k2];
    });
  });
}

template <typename T, int R>
__device__ __forceinline__ void dft(cpx<T> *v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const cpx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const cpx<T> s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  } else if constexpr (R == 8) {
    dft_comp<T, 2, 4>(v);
  } else if constexpr (R == 16) {
    dft_comp<T, 4, 4>(v);
  } else if constexpr (R == 32) {
    dft_comp<T, 4, 8>(v);
  } else if constexpr (R == 6) {
    dft_comp<T, 2, 3>(v);
  } else if constexpr (R == 9) {
    dft_comp<T, 3, 3>(v);
  } else if constexpr (R == 10) {
    dft_comp<T, 2, 5>(v);
  } else if constexpr (R == 12) {
    dft_comp<T, 4, 3>(v);
  } else if constexpr (R == 15) {
    dft_comp<T, 3, 5>(v);
  } else {
    static_assert(R % 2 == 1, "odd radix expected");
    dft_odd<T, R>(v);
  }
}

}  // namespace d2dhip
