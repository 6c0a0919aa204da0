// halo_kernels.hip -- halo cells of one rank's pencil (update_halo,
// halo_exchange; src/halo.f90, src/halo_common.f90,
// src/halo_exchange_{x,y,z}_body.f90).
//
// A halo array has the pencil's local extents plus `level` cells on both
// sides of the two distributed axes (y, z for an x-pencil; x, z for a
// y-pencil; x, y for a z-pencil).  update_halo copies the pencil into its
// interior, then halo_exchange fills the halo slabs axis by axis: for an
// x-pencil y then z, for a y-pencil x then z, for a z-pencil x then y.  Each
// step moves, over the whole extent of the other two axes (their halo cells
// included), the `level` interior slabs next to each side into the
// neighbour's opposite halo.  At one rank the neighbour along an axis is
// the rank itself when that axis is periodic (MPI_CART_SHIFT of a periodic
// communicator of size 1) and none otherwise, so the low halo takes the
// last `level` interior slabs, the high halo the first ones, and a
// non-periodic axis leaves its halo cells as they were.  Doing the axes in
// the reference's order with the same slab extents gives the same corner
// cells.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace d2dhip {

struct alignas(16) H16 {
  unsigned long long a, b;
};

// out (n1, n2, n3) at offset (o1, o2, o3) <- in (s1, s2, s3), i fastest: a
// block per pencil row (j, k), its threads along i.
template <typename E>
__global__ void __launch_bounds__(256) halo_interior_kernel(const E *__restrict__ in,
                                                            E *__restrict__ out, int s1, int s2,
                                                            int s3, int o1, int o2, int o3, int n1,
                                                            int n2) {
  const long long rows = (long long)s2 * s3;
  for (long long row = blockIdx.x; row < rows; row += gridDim.x) {
    const int k = (int)(row / s2), j = (int)(row - (long long)k * s2);
    const E *src = in + row * s1;
    E *dst = out + o1 + (long long)(j + o2) * n1 + (long long)(k + o3) * n1 * n2;
    for (int i = threadIdx.x; i < s1; i += blockDim.x) dst[i] = src[i];
  }
}

// The 2 * level halo slabs of axis d of arr (n1, n2, n3) from the slabs on
// the other side of the interior.
template <typename E>
__global__ void __launch_bounds__(256) halo_wrap_kernel(E *arr, int n1, int n2, int n3, int d,
                                                        int level) {
  const int m1 = d == 0 ? 2 * level : n1, m2 = d == 1 ? 2 * level : n2;
  const int m3 = d == 2 ? 2 * level : n3;
  const int nd = d == 0 ? n1 : (d == 1 ? n2 : n3);
  const long long total = (long long)m1 * m2 * m3;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long r = t / m1;
    int c[3];
    c[0] = (int)(t - r * m1);
    c[2] = (int)(r / m2);
    c[1] = (int)(r - (long long)c[2] * m2);
    const int q = c[d];
    int s[3] = {c[0], c[1], c[2]};
    if (q < level) {  // low halo <- last interior slabs
      s[d] = nd - 2 * level + q;
    } else {  // high halo <- first interior slabs
      c[d] = nd - level + (q - level);
      s[d] = q;
    }
    arr[c[0] + (long long)c[1] * n1 + (long long)c[2] * n1 * n2] =
        arr[s[0] + (long long)s[1] * n1 + (long long)s[2] * n1 * n2];
  }
}

static long long halo_grid(long long total) {
  long long g = (total + 255) / 256;
  return g < 1 ? 1 : (g > 65536 ? 65536 : g);
}

template <typename E>
static int halo_run(const void *in, void *arr, const int s[3], const int n[3], const int o[3],
                    const int order[2], const int wrap[2], int level, hipStream_t st) {
  if (in) {
    const long long rows = (long long)s[1] * s[2];
    if (rows > 0 && s[0] > 0) {
      const long long g = rows < (1LL << 20) ? rows : (1LL << 20);
      hipLaunchKernelGGL(halo_interior_kernel<E>, dim3((unsigned)g), dim3(256), 0, st,
                         (const E *)in, (E *)arr, s[0], s[1], s[2], o[0], o[1], o[2], n[0], n[1]);
    }
  }
  for (int a = 0; a < 2; ++a) {
    if (!wrap[a] || level == 0) continue;
    const int d = order[a];
    long long total = 2LL * level;
    for (int e = 0; e < 3; ++e)
      if (e != d) total *= n[e];
    if (total > 0)
      hipLaunchKernelGGL(halo_wrap_kernel<E>, dim3((unsigned)halo_grid(total)), dim3(256), 0, st,
                         (E *)arr, n[0], n[1], n[2], d, level);
  }
  return hip_check(hipGetLastError(), "halo kernel launch");
}

int launch_halo(int esize, const void *in, void *arr, int s1, int s2, int s3, int level,
                int ipencil, const int periodic[3], hipStream_t st) {
  // halo axes and their order (src/halo_exchange_{x,y,z}_body.f90)
  int order[2];
  if (ipencil == 1) {
    order[0] = 1;
    order[1] = 2;
  } else if (ipencil == 2) {
    order[0] = 0;
    order[1] = 2;
  } else {
    order[0] = 0;
    order[1] = 1;
  }
  const int s[3] = {s1, s2, s3};
  int n[3] = {s1, s2, s3}, o[3] = {0, 0, 0};
  int wrap[2];
  for (int a = 0; a < 2; ++a) {
    n[order[a]] += 2 * level;
    o[order[a]] = level;
    wrap[a] = periodic[order[a]] ? 1 : 0;
  }
  switch (esize) {
    case 4: return halo_run<unsigned int>(in, arr, s, n, o, order, wrap, level, st);
    case 8: return halo_run<unsigned long long>(in, arr, s, n, o, order, wrap, level, st);
    default: return halo_run<H16>(in, arr, s, n, o, order, wrap, level, st);
  }
}

}  // namespace d2dhip
