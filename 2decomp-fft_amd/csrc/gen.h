// gen.h -- counter-based synthetic inputs shared by host and device code.
//
// The stream is SplitMix64 (Steele, Lea & Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): the idx-th output of the
// generator seeded with `key` is mix(key + (idx + 1) * golden_gamma), so any
// element can be produced independently of the others.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/d2dhip.h"

namespace d2dhip {

__host__ __device__ inline unsigned long long gen_mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline unsigned long long gen_bits(unsigned long long key,
                                                       unsigned long long idx) {
  return gen_mix64(key + (idx + 1ull) * 0x9E3779B97F4A7C15ull);
}

// u in [0, 1) with 53 random bits.
__host__ __device__ inline double gen_u01(unsigned long long key, unsigned long long idx) {
  return (double)(gen_bits(key, idx) >> 11) * 0x1.0p-53;
}

// 2u - 1 in [-1, 1).
__host__ __device__ inline double gen_uniform_pm1(unsigned long long key,
                                                  unsigned long long idx) {
  return 2.0 * gen_u01(key, idx) - 1.0;
}

// Independent key for sub-stream `stream` of a seed.
__host__ __device__ inline unsigned long long gen_stream_key(unsigned long long seed,
                                                             unsigned long long stream) {
  return gen_mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
}

// Box-Muller: N(0,1) from two uniforms in [0,1).
__host__ __device__ inline double gen_box_muller(double u1, double u2) {
  return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}

// Value at 0-based (i, j, k) of the modal field
//   sum_m amp_m cos(2 pi (kx_m i/nx + ky_m j/ny + kz_m k/nz)) + sigma * xi(idx)
// The phase is reduced exactly in integers before the cosine.
__host__ __device__ inline double gen_modal_value(const d2dhip_modes *m, long long nx,
                                                  long long ny, long long nz, long long i,
                                                  long long j, long long k, long long idx) {
  const long long den = nx * ny * nz;
  double v = 0.0;
  for (int q = 0; q < m->nmodes; ++q) {
    long long a = ((long long)m->kx[q] * i) % nx;
    long long b = ((long long)m->ky[q] * j) % ny;
    long long c = ((long long)m->kz[q] * k) % nz;
    if (a < 0) a += nx;
    if (b < 0) b += ny;
    if (c < 0) c += nz;
    const long long num = (a * ny * nz + b * nx * nz + c * nx * ny) % den;
    v += m->amp[q] * cos(6.283185307179586 * ((double)num / (double)den));
  }
  if (m->sigma != 0.0) {
    const double u1 = gen_u01(m->noise_seed, 2ull * (unsigned long long)idx);
    const double u2 = gen_u01(m->noise_seed, 2ull * (unsigned long long)idx + 1ull);
    v += m->sigma * gen_box_muller(u1, u2);
  }
  return v;
}

}  // namespace d2dhip
