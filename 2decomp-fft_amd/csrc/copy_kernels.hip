// copy_kernels.hip -- alltoallv pack/unpack maps and synthetic-input fills.
//
// The eight mem_split_* / mem_merge_* maps of the reference
// (src/transpose_x_to_y.f90:266-500, src/transpose_y_to_z.f90:370-606,
// src/transpose_z_to_y.f90:370-606, src/transpose_y_to_x.f90:266-500) are
// all block copies that keep natural (i fastest) order inside each
// destination block: element (i,j,k) of a pencil (n1,n2,n3) whose block axis
// value lies in block m = [b1, b1 + w) sits in the flat buffer at
//   init_m + (i-b1) + j*w + k*n2*w      (block on i)
//   init_m + i + (j-b1)*n1 + k*w*n1     (block on j)
//   init_m + i + j*n1 + (k-b1)*n1*n2    (block on k)
// with init_m the alltoallv displacement (or m*count for the padded EVEN
// layout).  One kernel covers every map: threads walk the pencil with i
// fastest, so both the pencil side and each block's run in the buffer are
// read and written in consecutive addresses.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "codelets.h"
#include "gen.h"
#include "internal.h"

namespace d2dhip {

struct PackArgs {
  const void *src;
  void *dst;
  int n1, n2, n3;
  int ax;     // block axis: 0 = i, 1 = j, 2 = k
  int split;  // 1: pencil -> buffer (mem_split), 0: buffer -> pencil (mem_merge)
  int np;
  int b1[D2DHIP_MAX_PROCS];
  int w[D2DHIP_MAX_PROCS];
  long long init[D2DHIP_MAX_PROCS];
};

struct alignas(16) E16 {
  unsigned long long a, b;
};

// One unit = a chunk of up to 1024 consecutive i of one pencil row (j, k);
// each thread moves 4 elements of it, 256 apart, loads first.  AX is the
// block axis: for j or k the destination block is fixed per row, for i it is
// found per element.
template <typename E, int AX>
__global__ void __launch_bounds__(256) pack_kernel(PackArgs a, long long units, int cpr) {
  constexpr int V = 4, CH = 256 * V;
  const E *src = reinterpret_cast<const E *>(a.src);
  E *dst = reinterpret_cast<E *>(a.dst);
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const long long row = u / cpr;
    const int chunk = (int)(u - row * cpr);
    const int j = (int)(row % a.n2);
    const int k = (int)(row / a.n2);
    long long rbase = 0;  // row part of the buffer position (blocks along j or k)
    if constexpr (AX != 0) {
      const int key = AX == 1 ? j : k;
      int m = 0;
      while (m + 1 < a.np && key >= a.b1[m + 1]) ++m;
      rbase = a.init[m] + (AX == 1 ? (long long)(j - a.b1[m]) * a.n1 + (long long)k * a.w[m] * a.n1
                                   : (long long)j * a.n1 + (long long)(k - a.b1[m]) * a.n1 * a.n2);
    }
    const long long lin0 = (long long)a.n1 * row;
    long long pos[V];
    E v[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
      const int i = chunk * CH + e * 256 + (int)threadIdx.x;
      pos[e] = -1;
      if (i < a.n1) {
        if constexpr (AX == 0) {
          int m = 0;
          while (m + 1 < a.np && i >= a.b1[m + 1]) ++m;
          pos[e] = a.init[m] + (long long)(i - a.b1[m]) + (long long)j * a.w[m] +
                   (long long)k * a.n2 * a.w[m];
        } else {
          pos[e] = rbase + i;
        }
        v[e] = a.split ? src[lin0 + i] : src[pos[e]];
      }
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
      const int i = chunk * CH + e * 256 + (int)threadIdx.x;
      if (pos[e] >= 0) {
        if (a.split)
          dst[pos[e]] = v[e];
        else
          dst[lin0 + i] = v[e];
      }
    }
  }
}

// Block of a position along the block axis: the m with b1[m] <= key <
// b1[m+1], by a binary search of fixed depth.
__device__ __forceinline__ int pack_block(const PackArgs &a, int key) {
  int lo = 0, hi = a.np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (key >= a.b1[mid]) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Vector form for maps whose runs all start on 16-byte boundaries (checked
// on the host: n1, every block start / width on the i axis and every
// displacement multiples of the vector): 16 bytes per lane.  A unit is rpu
// whole pencil rows (rows of fewer than 1024 vectors), a 1024-vector chunk
// of a row (rows of a multiple of 1024 vectors) or, for other long rows,
// 1024 consecutive vectors of the pencil taken as one flat run (a row of
// 1025 vectors would otherwise leave a unit with one), the row found by a
// division through the FP64 reciprocal of the row length, corrected to
// exact; each thread moves 4 vectors, loads first.  The block of a row
// (blocks on j or k) or of a vector (blocks on i) is looked up once, not per
// element.
template <int AX>
__global__ void __launch_bounds__(256) pack_vec_kernel(PackArgs a, int vpe, long long units,
                                                       int rpu, long long rows, double inv_nvr) {
  constexpr int NV = 4, CH = 256 * NV;
  const E16 *src = reinterpret_cast<const E16 *>(a.src);
  E16 *dst = reinterpret_cast<E16 *>(a.dst);
  const int nvr = a.n1 / vpe;  // vectors per pencil row
  const long long total = rows * nvr;
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    long long pv[NV], bv[NV];  // pencil and buffer vector index
    E16 v[NV];
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const int t = e * 256 + (int)threadIdx.x;
      long long row;
      int iv;
      if (rpu > 1) {
        row = u * rpu + t / nvr;
        iv = t % nvr;
        if (t >= rpu * nvr) row = rows;  // past the unit
      } else if (nvr % CH == 0) {  // whole chunks of a row
        const int cpr = nvr / CH;
        row = u / cpr;
        iv = (int)(u - row * cpr) * CH + t;
      } else {
        const long long g = u * CH + t;  // flat vector index
        row = (long long)((double)g * inv_nvr);
        long long r = g - row * nvr;
        if (r < 0) {  // the product rounded up past a row boundary
          --row;
          r += nvr;
        } else if (r >= nvr) {
          ++row;
          r -= nvr;
        }
        iv = (int)r;
        if (g >= total) row = rows;
      }
      pv[e] = -1;
      if (row < rows) {
        const int j = (int)(row % a.n2), k = (int)(row / a.n2);
        const int i = iv * vpe;
        long long pos;
        if constexpr (AX == 0) {
          const int m = pack_block(a, i);
          pos = a.init[m] + (long long)(i - a.b1[m]) + (long long)j * a.w[m] +
                (long long)k * a.n2 * a.w[m];
        } else {
          const int m = pack_block(a, AX == 1 ? j : k);
          pos = a.init[m] + i + (AX == 1 ? (long long)(j - a.b1[m]) * a.n1 + (long long)k * a.w[m] * a.n1
                                         : (long long)j * a.n1 + (long long)(k - a.b1[m]) * a.n1 * a.n2);
        }
        pv[e] = row * nvr + iv;
        bv[e] = pos / vpe;
        v[e] = a.split ? src[pv[e]] : src[bv[e]];
      }
    }
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      if (pv[e] >= 0) {
        if (a.split)
          dst[bv[e]] = v[e];
        else
          dst[pv[e]] = v[e];
      }
    }
  }
}

template <typename E>
static void launch_pack_e(const PackArgs &a, long long units, int cpr, long long grid, hipStream_t st) {
  if (a.ax == 0)
    hipLaunchKernelGGL((pack_kernel<E, 0>), dim3((unsigned)grid), dim3(256), 0, st, a, units, cpr);
  else if (a.ax == 1)
    hipLaunchKernelGGL((pack_kernel<E, 1>), dim3((unsigned)grid), dim3(256), 0, st, a, units, cpr);
  else
    hipLaunchKernelGGL((pack_kernel<E, 2>), dim3((unsigned)grid), dim3(256), 0, st, a, units, cpr);
}

int launch_pack(int dir, int split, int esize, const void *src, void *dst, const d2dhip_decomp *d,
                int even, hipStream_t st) {
  PackArgs a;
  a.src = src;
  a.dst = dst;
  // The pencil each map reads (split) or writes (merge), the block axis, and
  // the dist / displacement arrays, as in the reference's calls
  // (e.g. src/transpose_x_to_y.f90:94-95 and :132-133).
  const int *sz;
  const int *dist;
  const long long *disp;
  long long count;
  int np;
  switch (dir) {
    case D2DHIP_X_TO_Y:
      np = d->p_row;
      sz = split ? d->xsz : d->ysz;
      a.ax = split ? 0 : 1;
      dist = split ? d->x1dist : d->y1dist;
      disp = split ? d->x1disp : d->y1disp;
      count = split ? d->x1count : d->y1count;
      break;
    case D2DHIP_Y_TO_Z:
      np = d->p_col;
      sz = split ? d->ysz : d->zsz;
      a.ax = split ? 1 : 2;
      dist = split ? d->y2dist : d->z2dist;
      disp = split ? d->y2disp : d->z2disp;
      count = split ? d->y2count : d->z2count;
      break;
    case D2DHIP_Z_TO_Y:
      np = d->p_col;
      sz = split ? d->zsz : d->ysz;
      a.ax = split ? 2 : 1;
      dist = split ? d->z2dist : d->y2dist;
      disp = split ? d->z2disp : d->y2disp;
      count = split ? d->z2count : d->y2count;
      break;
    default:
      np = d->p_row;
      sz = split ? d->ysz : d->xsz;
      a.ax = split ? 1 : 0;
      dist = split ? d->y1dist : d->x1dist;
      disp = split ? d->y1disp : d->x1disp;
      count = split ? d->y1count : d->x1count;
      break;
  }
  if (np < 1 || np > D2DHIP_MAX_PROCS)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "process count out of range");
  a.n1 = sz[0];
  a.n2 = sz[1];
  a.n3 = sz[2];
  a.split = split;
  a.np = np;
  int b = 0;
  long long wsum = 0;
  for (int m = 0; m < np; ++m) {
    a.b1[m] = b;
    a.w[m] = dist[m];
    a.init[m] = even ? (long long)m * count : disp[m];
    b += dist[m];
    wsum += dist[m];
  }
  const int axlen = a.ax == 0 ? a.n1 : (a.ax == 1 ? a.n2 : a.n3);
  if (wsum != axlen)
    return set_error(D2DHIP_ERR_INVALID_VALUE, "decomposition does not match the pencil");
  if ((long long)a.n1 * a.n2 * a.n3 == 0) return D2DHIP_SUCCESS;
  // 16-byte vectors when every run starts on a vector boundary
  const int vpe = 16 / esize;
  bool vec = (a.n1 % vpe == 0) && ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 16 == 0);
  for (int m = 0; m < np && vec; ++m) {
    vec = a.init[m] % vpe == 0;
    if (a.ax == 0) vec = vec && a.b1[m] % vpe == 0 && a.w[m] % vpe == 0;
  }
  if (vec) {
    const long long rows = (long long)a.n2 * a.n3;
    const int nvr = a.n1 / vpe;
    const int rpu = nvr >= 1024 ? 1 : 1024 / nvr;
    const long long vunits = rpu > 1 ? (rows + rpu - 1) / rpu : (rows * nvr + 1023) / 1024;
    const long long vgrid = vunits < (1LL << 20) ? vunits : (1LL << 20);
    const double inv = 1.0 / (double)nvr;
    if (a.ax == 0)
      hipLaunchKernelGGL((pack_vec_kernel<0>), dim3((unsigned)vgrid), dim3(256), 0, st, a, vpe, vunits,
                         rpu, rows, inv);
    else if (a.ax == 1)
      hipLaunchKernelGGL((pack_vec_kernel<1>), dim3((unsigned)vgrid), dim3(256), 0, st, a, vpe, vunits,
                         rpu, rows, inv);
    else
      hipLaunchKernelGGL((pack_vec_kernel<2>), dim3((unsigned)vgrid), dim3(256), 0, st, a, vpe, vunits,
                         rpu, rows, inv);
    return hip_check(hipGetLastError(), "pack_vec_kernel launch");
  }
  const int cpr = (a.n1 + 1023) / 1024;
  const long long units = (long long)a.n2 * a.n3 * cpr;
  const long long grid = units < (1LL << 20) ? units : (1LL << 20);
  switch (esize) {
    case 4: launch_pack_e<unsigned int>(a, units, cpr, grid, st); break;
    case 8: launch_pack_e<unsigned long long>(a, units, cpr, grid, st); break;
    default: launch_pack_e<E16>(a, units, cpr, grid, st); break;
  }
  return hip_check(hipGetLastError(), "pack_kernel launch");
}

// ---- synthetic inputs ------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) fill_uniform_kernel(T *dst, long long nv,
                                                           unsigned long long seed) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < nv; i += (long long)gridDim.x * 256)
    dst[i] = (T)gen_uniform_pm1(seed, (unsigned long long)i);
}

// Copy of complex columns with the access pattern of an axis stage and no
// arithmetic: tiles of 16 columns (c0, c1) of n points sj apart, 1024-thread
// blocks in XCD-aware order; each thread loads its n/64 points, then stores
// them.  It measures the bandwidth that pattern allows.
template <typename T>
__global__ void __launch_bounds__(1024)
    column_copy_kernel(const cpx<T> *__restrict__ src, cpx<T> *__restrict__ dst, long long C0,
                       long long C1, long long s1, long long sj, int n) {
  constexpr int TX = 16, RPT = 64, NP = 16;
  const long long tiles0 = (C0 + TX - 1) / TX, ntiles = tiles0 * C1;
  const long long tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const long long c1 = tile / tiles0;
  const long long c0 = (tile - c1 * tiles0) * TX + threadIdx.x % TX;
  if (c0 >= C0) return;
  const long long base = c0 + c1 * s1;
  for (int j0 = (int)(threadIdx.x / TX); j0 < n; j0 += RPT * NP) {
    cpx<T> v[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m) {  // all loads in flight before the stores
      const int j = j0 + m * RPT;
      v[m] = (j < n) ? src[base + (long long)j * sj] : cpx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      const int j = j0 + m * RPT;
      if (j < n) dst[base + (long long)j * sj] = v[m];
    }
  }
}

int launch_column_copy(int elem, const void *src, void *dst, long long C0, long long C1,
                       long long s1, long long sj, int n, hipStream_t st) {
  const long long ntiles = (C0 + 15) / 16 * C1;
  const long long nb = xcd_grid(ntiles);
  if (nb <= 0) return D2DHIP_SUCCESS;
  if (nb > 0x7fffffffLL) return set_error(D2DHIP_ERR_UNSUPPORTED, "too many columns");
  if (elem == D2DHIP_C16)
    hipLaunchKernelGGL(column_copy_kernel<double>, dim3((unsigned)nb), dim3(1024), 0, st,
                       (const cpx<double> *)src, (cpx<double> *)dst, C0, C1, s1, sj, n);
  else
    hipLaunchKernelGGL(column_copy_kernel<float>, dim3((unsigned)nb), dim3(1024), 0, st,
                       (const cpx<float> *)src, (cpx<float> *)dst, C0, C1, s1, sj, n);
  return hip_check(hipGetLastError(), "column_copy_kernel launch");
}

int launch_fill_uniform(void *dst, long long n, int elem, unsigned long long seed,
                        hipStream_t st) {
  const bool cplx = (elem == D2DHIP_C16 || elem == D2DHIP_C8);
  const long long nv = cplx ? 2 * n : n;
  if (nv == 0) return D2DHIP_SUCCESS;
  long long grid = (nv + 255) / 256;
  if (grid > 65536) grid = 65536;
  if (elem == D2DHIP_R8 || elem == D2DHIP_C16)
    hipLaunchKernelGGL(fill_uniform_kernel<double>, dim3((unsigned)grid), dim3(256), 0, st,
                       (double *)dst, nv, seed);
  else
    hipLaunchKernelGGL(fill_uniform_kernel<float>, dim3((unsigned)grid), dim3(256), 0, st,
                       (float *)dst, nv, seed);
  return hip_check(hipGetLastError(), "fill_uniform launch");
}

template <typename T>
__global__ void __launch_bounds__(256) fill_modal_kernel(T *dst, int nx, int ny, int nz,
                                                         d2dhip_modes m) {
  const long long n = (long long)nx * ny * nz;
  for (long long idx = blockIdx.x * 256LL + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * 256) {
    const long long i = idx % nx;
    const long long r = idx / nx;
    const long long j = r % ny;
    const long long k = r / ny;
    dst[idx] = (T)gen_modal_value(&m, nx, ny, nz, i, j, k, idx);
  }
}

int launch_fill_modal(void *dst, int nx, int ny, int nz, int elem, const d2dhip_modes *m,
                      hipStream_t st) {
  const long long n = (long long)nx * ny * nz;
  long long grid = (n + 255) / 256;
  if (grid > 65536) grid = 65536;
  if (elem == D2DHIP_R8)
    hipLaunchKernelGGL(fill_modal_kernel<double>, dim3((unsigned)grid), dim3(256), 0, st,
                       (double *)dst, nx, ny, nz, *m);
  else
    hipLaunchKernelGGL(fill_modal_kernel<float>, dim3((unsigned)grid), dim3(256), 0, st,
                       (float *)dst, nx, ny, nz, *m);
  return hip_check(hipGetLastError(), "fill_modal launch");
}

}  // namespace d2dhip
