// fft_kernels.hip -- batched 1D FFT stages along one pencil axis (gfx950).
//
// One workgroup transforms a tile of TX columns of length n.  The tile lives
// in LDS between Stockham passes; within a pass each thread holds up to
// kRegCap complex values in registers (E/R butterflies of radix R).  The
// first pass reads its butterfly inputs straight from HBM and the last pass
// writes straight to HBM, so a stage costs one read and one write of the
// pencil.  Consecutive threads take consecutive tile columns (the fastest
// tile index), which makes HBM accesses of strided axes 16*TX-byte runs and
// those of contiguous lines TX runs of consecutive points.
//
// Stockham autosort pass (decimation in time) with radix R over
// sub-transforms of length L (product of the earlier radices): butterfly b
// has k = b mod L, g = b div L, inputs x[b + s*n/R] (s < R) scaled by
// exp(-2 pi i s k / (L R)), and outputs y[g*L*R + k + s*L].
//
// Real transforms pair two real columns into one complex column
// (z = a + i b) and separate / merge them with the Hermitian identities
//   A[k] = (Z[k] + conj Z[n-k]) / 2,   B[k] = (Z[k] - conj Z[n-k]) / (2i),
// which holds for every n, odd or even.
//
// The stage descriptor sits in device memory and is read with scalar loads;
// each radix's pass is a separate (non-inlined) function so that its
// register arrays are allocated on their own.
#include <hip/hip_runtime.h>

#include "batch_loop.h"
#include "codelets.h"
#include "internal.h"

namespace d2dhip {

// Input value of tile column c0 (global), outer index c1, point j.
template <typename T, int IO>
__device__ __forceinline__ cpx<T> pload(DescPtr d, const AxisIo &io, long long c0, long long c1,
                                        int j) {
  if constexpr (IO == IO_R2C) {
    const long long base = c1 * d->in_s1 + (long long)j * d->in_sj;
    const long long ca = 2 * c0;
    cpx<T> v = {T(0), T(0)};
    if (ca < d->C0r) v.x = gload_r<T>(io.in, base + ca * d->in_s0);
    if (ca + 1 < d->C0r) v.y = gload_r<T>(io.in, base + (ca + 1) * d->in_s0);
    return v;
  } else {
    cpx<T> v = {T(0), T(0)};
    if (c0 < d->C0)  // (a c2c stage may read its pencil in unpack order, packj = 2 or 3)
      v = gload<T>(io.in, c0 * d->in_s0 + (c0 >> d->bsh) * d->in_b0 + c1 * d->in_s1 +
                              (long long)j * d->in_sj + (d->packj >= 2 ? packj_offset(d, j, c1) : 0));
    if (io.conj) v.y = -v.y;
    return v;
  }
}

// Store of an output point of the (conjugated, for inverses) transform.
template <typename T, int IO>
__device__ __forceinline__ void pstore(DescPtr d, const AxisIo &io, long long c0, long long c1,
                                       int j, cpx<T> v) {
  if constexpr (IO == IO_C2R) {
    // the pre-phase stored conj(Z): z = conj(result) = a + i b
    const long long base = c1 * d->out_s1 + (long long)j * d->out_sj;
    const long long ca = 2 * c0;
    if (ca < d->C0r) gstore_r<T>(io.out, base + ca * d->out_s0, v.x);
    if (ca + 1 < d->C0r) gstore_r<T>(io.out, base + (ca + 1) * d->out_s0, -v.y);
  } else {
    if (c0 < d->C0) {
      if (io.conj) v.y = -v.y;
      gstore<T>(io.out, c0 * d->out_s0 + (c0 >> d->bsh) * d->out_b0 + c1 * d->out_s1 + (long long)j * d->out_sj +
                            (d->packj == 1 ? packj_offset(d, j, c1) : d->packj == 3 ? packj_offset2(d, j, c1) : 0),
                v);
    }
  }
}

// The stage's LDS tile: the dynamic LDS of the launch.  Taken from the
// array itself rather than passed as a pointer, so that the non-inlined pass
// function addresses it with ds_read / ds_write.
template <typename T>
__device__ __forceinline__ cpx<T> *lds_tile() {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  return reinterpret_cast<cpx<T> *>(lds_raw);
}

// One Stockham pass of radix R.  `sg` reads the pass inputs from HBM (first
// pass), `dg` writes the outputs to HBM (last pass); otherwise LDS.  Ends
// with a barrier whenever it wrote LDS.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Function arguments arrive in vector registers; they are wave-uniform, and
// re-deriving them with readfirstlane keeps the descriptor reads and the
// index arithmetic of the pass on the scalar unit.
template <typename T, int E, int IO, int R>
__device__ __noinline__ void run_pass(DescPtr d_, AxisIo io_, int p_, bool sg_, bool dg_,
                                      long long c0b_, long long c1_) {
  const DescPtr d = (DescPtr)uniform_u64((unsigned long long)d_);
  const AxisIo io = {(const void *)uniform_u64((unsigned long long)io_.in),
                     (void *)uniform_u64((unsigned long long)io_.out),
                     __builtin_amdgcn_readfirstlane(io_.conj)};
  const int p = __builtin_amdgcn_readfirstlane(p_);
  const bool sg = __builtin_amdgcn_readfirstlane((int)sg_) != 0;
  const bool dg = __builtin_amdgcn_readfirstlane((int)dg_) != 0;
  const long long c0b = (long long)uniform_u64((unsigned long long)c0b_);
  const long long c1 = (long long)uniform_u64((unsigned long long)c1_);
  cpx<T> *lds = lds_tile<T>();
  constexpr int NB = E / R;
  static_assert(NB >= 1, "radix larger than the register capacity");
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int n = d->n, nb = n / R, L = d->pass[p].L;
  const float invL = d->pass[p].invL;
  const int TX = d->TX, lg = d->log2TX, tmask = TX - 1;
  const int ldj = d->ldj, ldt = d->ldt;
  const int total = TX * nb;
  cpx<T> v[NB][R];

#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int w = tid + q * nthr;
    if (w < total) {
      const int t = w & tmask, b = w >> lg;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const int j = b + s * nb;
        v[q][s] = sg ? pload<T, IO>(d, io, c0b + t, c1, j) : lds[j * ldj + t * ldt];
      }
    }
  }
  if (!sg && !dg) __syncthreads();  // all LDS reads before LDS writes

  const void *tw = d->tw;
  const int tw_off = d->pass[p].tw_off;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int w = tid + q * nthr;
    if (w < total) {
      const int t = w & tmask, b = w >> lg;
      int g = (int)((float)b * invL);
      int k = b - g * L;
      if (k < 0) {
        --g;
        k += L;
      } else if (k >= L) {
        ++g;
        k -= L;
      }
      if (L > 1) {
#pragma unroll
        for (int s = 1; s < R; ++s) v[q][s] = cmul(v[q][s], gload<T>(tw, tw_off + k * (R - 1) + (s - 1)));
      }
      dft<T, R>(v[q]);
      const int jb = g * L * R + k;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const int j = jb + s * L;
        if (dg)
          pstore<T, IO>(d, io, c0b + t, c1, j, v[q][s]);
        else
          lds[j * ldj + t * ldt] = v[q][s];
      }
    }
  }
  if (!dg) __syncthreads();
}

template <typename T, int E, int IO>
__device__ __forceinline__ void dispatch_pass(DescPtr d, const AxisIo &io, int p, bool sg, bool dg,
                                              long long c0b, long long c1) {
  switch (d->pass[p].r) {
    case 1: run_pass<T, E, IO, 1>(d, io, p, sg, dg, c0b, c1); break;
    case 2: run_pass<T, E, IO, 2>(d, io, p, sg, dg, c0b, c1); break;
    case 3: run_pass<T, E, IO, 3>(d, io, p, sg, dg, c0b, c1); break;
    case 4: run_pass<T, E, IO, 4>(d, io, p, sg, dg, c0b, c1); break;
    case 5: run_pass<T, E, IO, 5>(d, io, p, sg, dg, c0b, c1); break;
    case 6: run_pass<T, E, IO, 6>(d, io, p, sg, dg, c0b, c1); break;
    case 7: run_pass<T, E, IO, 7>(d, io, p, sg, dg, c0b, c1); break;
    case 8: run_pass<T, E, IO, 8>(d, io, p, sg, dg, c0b, c1); break;
    case 9: run_pass<T, E, IO, 9>(d, io, p, sg, dg, c0b, c1); break;
    case 10: run_pass<T, E, IO, 10>(d, io, p, sg, dg, c0b, c1); break;
    case 11: run_pass<T, E, IO, 11>(d, io, p, sg, dg, c0b, c1); break;
    case 12: run_pass<T, E, IO, 12>(d, io, p, sg, dg, c0b, c1); break;
    case 13: run_pass<T, E, IO, 13>(d, io, p, sg, dg, c0b, c1); break;
    case 15: run_pass<T, E, IO, 15>(d, io, p, sg, dg, c0b, c1); break;
    case 16: run_pass<T, E, IO, 16>(d, io, p, sg, dg, c0b, c1); break;
    case 17: run_pass<T, E, IO, 17>(d, io, p, sg, dg, c0b, c1); break;
    default: break;  // the planner only emits the radices above
  }
}

// Waves per SIMD the general kernels' register allocation must leave room
// for; the bound also holds for the non-inlined pass functions.  FP32 passes
// fit 4 waves without spilling; FP64 passes would spill, so they are left
// unbounded.
template <typename T>
constexpr int general_wpe() {
  return sizeof(T) == 8 ? 1 : 4;
}

// ---- real-to-real (DCT / DST) stages ---------------------------------------------
// FFTW's REDFT / RODFT kinds, unnormalised, of the L points [dtt_first,
// dtt_first + L) of each line; the other points of the line are copied
// through, as FFTW's transform of the sub-array leaves them.  Two real lines
// a, b always go through one complex transform, z = a' + i b'.
//
// Types 2, 3, 6, 7 (DCT-II / III, DST-II / III) run an L-point DFT
// (J. Makhoul, "A fast cosine transform in one and two dimensions", IEEE
// Trans. ASSP 28(1), 1980), with w_k = exp(-i pi k / (2L)):
//   DCT-II   v[j/2] = x_j (j even), v[L-1-j/2] = x_j (j odd), V = DFT_L(v),
//            Y_k = 2 Re(w_k V_k); the pair separates as in a real-to-complex
//            transform, V_a = (Z_k + conj Z_{L-k}) / 2, V_b = (Z_k - conj Z_{L-k}) / 2i.
//   DST-II   Y_k = DCT-II(x')_{L-1-k} with x'_j = (-1)^j x_j.
//   DCT-III  V_j = conj(w_j) (X_j - i X_{L-j}) (X_L = 0) is Hermitian, so
//            v = IDFT_L(V) is real and y_{2n} = v_n, y_{2n+1} = v_{L-1-n};
//            the pair goes in as Z_j = V_a,j + i V_b,j, the inverse DFT as
//            conj(DFT(conj Z)).
//   DST-III  (-1)^k DCT-III of X read backwards.
// Types 1, 4, 5, 8 run the DFT of an even / odd extension e of length m:
//   type 1 DCT-I    m = 2(L-1)  e[j] = e[m-j] = X_j                 Y_k =  Re E[k]
//   type 4 DCT-IV   m = 8L      e[2j+1] = e[m-2j-1] = X_j           Y_k =  Re E[2k+1]
//   type 5 DST-I    m = 2(L+1)  e[j+1] = -e[m-j-1] = X_j            Y_k = -Im E[k+1]
//   type 8 DST-IV   m = 8L      e[2j+1] = -e[m-2j-1] = X_j          Y_k = -Im E[2k+1]
// (other e[i] zero).  An even real extension has a real DFT and an odd one
// an imaginary DFT, so Z = E_a + i E_b separates as (Re Z, Im Z) for the
// DCTs and as (Im Z, -Re Z) for the DSTs.
__device__ __forceinline__ bool dtt_half(int type) {
  return type == 2 || type == 3 || type == 6 || type == 7;
}

// The tile's line pairs as the complex sequence the DFT takes, in LDS
// points [0, m) (m = the stage's DFT length).
template <typename T>
__device__ __forceinline__ void dtt_fill(DescPtr d, const void *in, cpx<T> *lds, long long c0b,
                                         long long c1, int m) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int TX = d->TX, lg = d->log2TX, tmask = TX - 1;
  const int ldj = d->ldj, ldt = d->ldt;
  const int type = d->dtt, f = d->dtt_first, L = d->dtt_len;
  const bool half = dtt_half(type);
  if (!half && type != 1) {  // the zeros of the extension (DCT-I writes every point)
    for (int w = tid; w < TX * m; w += nthr) {
      const int t = w & tmask, j = w >> lg;
      lds[j * ldj + t * ldt] = cpx<T>{T(0), T(0)};
    }
    __syncthreads();
  }
  for (int w = tid; w < TX * L; w += nthr) {
    const int t = w & tmask, j = w >> lg;
    const long long ca = 2 * (c0b + t);
    const long long base = c1 * d->in_s1 + (long long)(f + j) * d->in_sj;
    T a = T(0), b = T(0);
    if (ca < d->C0r) a = gload_r<T>(in, base + ca * d->in_s0);
    if (ca + 1 < d->C0r) b = gload_r<T>(in, base + (ca + 1) * d->in_s0);
    int p1, p2 = -1;
    switch (type) {
      case 2: p1 = (j & 1) ? L - 1 - (j >> 1) : (j >> 1); break;
      case 6:  // (-1)^j x_j in the DCT-II order
        if (j & 1) {
          a = -a;
          b = -b;
        }
        p1 = (j & 1) ? L - 1 - (j >> 1) : (j >> 1);
        break;
      case 3: p1 = j; break;
      case 7: p1 = L - 1 - j; break;
      case 1: p1 = j; p2 = (j > 0 && j < L - 1) ? m - j : -1; break;
      case 5: p1 = j + 1; p2 = m - j - 1; break;
      default: p1 = 2 * j + 1; p2 = m - 2 * j - 1; break;  // types 4, 8
    }
    lds[p1 * ldj + t * ldt] = cpx<T>{a, b};
    if (p2 >= 0) lds[p2 * ldj + t * ldt] = type >= 5 ? cpx<T>{-a, -b} : cpx<T>{a, b};
  }
  __syncthreads();
  if (type == 3 || type == 7) {
    // conj(Z_j) = w_j conj(u_j), u_j = (Xa_j + Xb_{L-j}, Xb_j - Xa_{L-j}):
    // each thread takes the points j and L-j together
    const int nj = L / 2 + 1;
    for (int w = tid; w < TX * nj; w += nthr) {
      const int t = w & tmask, j = w >> lg, jm = L - j;
      if (j > jm) continue;
      const cpx<T> A = lds[j * ldj + t * ldt];
      const cpx<T> B = (j > 0) ? lds[jm * ldj + t * ldt] : cpx<T>{T(0), T(0)};
      const cpx<T> cs = gload<T>(d->dtw, j);
      const T ux = A.x + B.y, uy = A.y - B.x;
      lds[j * ldj + t * ldt] = cpx<T>{cs.x * ux - cs.y * uy, -(cs.x * uy + cs.y * ux)};
      if (j > 0 && j < jm) {
        const cpx<T> cm = gload<T>(d->dtw, jm);
        const T vx = B.x + A.y, vy = B.y - A.x;
        lds[jm * ldj + t * ldt] = cpx<T>{cm.x * vx - cm.y * vy, -(cm.x * vy + cm.y * vx)};
      }
    }
    __syncthreads();
  }
}

// The transformed lines from the DFT in LDS, and the copied-through points.
template <typename T>
__device__ __forceinline__ void dtt_store(DescPtr d, const void *in, void *out, const cpx<T> *lds,
                                          long long c0b, long long c1) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int TX = d->TX, lg = d->log2TX, tmask = TX - 1;
  const int ldj = d->ldj, ldt = d->ldt;
  const int type = d->dtt, f = d->dtt_first, L = d->dtt_len, np = d->dtt_n;
  for (int w = tid; w < TX * np; w += nthr) {
    const int t = w & tmask, k = w >> lg;
    const long long ca = 2 * (c0b + t);
    if (ca >= d->C0r) continue;
    const bool two = ca + 1 < d->C0r;
    const long long ib = c1 * d->in_s1 + (long long)k * d->in_sj;
    const long long ob = c1 * d->out_s1 + (long long)k * d->out_sj;
    T a, b = T(0);
    const int kk = k - f;
    if (kk >= 0 && kk < L) {
      if (type == 2 || type == 6) {
        const int q = type == 2 ? kk : L - 1 - kk;
        const cpx<T> zk = lds[q * ldj + t * ldt];
        const cpx<T> zm = lds[(q == 0 ? 0 : L - q) * ldj + t * ldt];
        const cpx<T> cs = gload<T>(d->dtw, q);
        a = cs.x * (zk.x + zm.x) + cs.y * (zk.y - zm.y);
        b = cs.x * (zk.y + zm.y) + cs.y * (zm.x - zk.x);
      } else if (type == 3 || type == 7) {
        const int q = (kk & 1) ? L - 1 - (kk >> 1) : (kk >> 1);
        const cpx<T> r = lds[q * ldj + t * ldt];
        const bool neg = type == 7 && (kk & 1);
        a = neg ? -r.x : r.x;
        b = neg ? r.y : -r.y;
      } else {
        const int q = (type == 4 || type == 8) ? 2 * kk + 1 : type == 5 ? kk + 1 : kk;
        const cpx<T> z = lds[q * ldj + t * ldt];
        if (type <= 4) {
          a = z.x;
          b = z.y;
        } else {
          a = -z.y;
          b = z.x;
        }
      }
    } else {
      a = gload_r<T>(in, ib + ca * d->in_s0);
      if (two) b = gload_r<T>(in, ib + (ca + 1) * d->in_s0);
    }
    gstore_r<T>(out, ob + ca * d->out_s0, a);
    if (two) gstore_r<T>(out, ob + (ca + 1) * d->out_s0, b);
  }
}

// One tile of a stage, by the whole workgroup (any blockDim at least the
// stage's planned block size).
template <typename T, int E, int IO>
__device__ __forceinline__ void axis_tile(DescPtr d, const AxisIo &io, long long tile) {
  cpx<T> *lds = lds_tile<T>();
  const void *in = io.in;
  void *out = io.out;
  const long long tiles0 = d->tiles0;
  const long long c1 = tile / tiles0;
  const long long c0b = (tile - c1 * tiles0) * d->TX;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int n = d->n, TX = d->TX, lg = d->log2TX, tmask = TX - 1;
  const int ldj = d->ldj, ldt = d->ldt, nh = d->nh;

  if constexpr (IO == IO_C2R) {
    // Merge the two half-spectra of columns (2c0, 2c0+1) into conj(Z) over
    // all n points; the imaginary parts of the zero and Nyquist modes are
    // ignored, as a complex-to-real transform does.
    for (int w = tid; w < TX * nh; w += nthr) {
      const int t = w & tmask, k = w >> lg;
      const long long ca = 2 * (c0b + t);
      const long long base = c1 * d->in_s1 + (long long)k * d->in_sj;
      long long ia = base + ca * d->in_s0, ib = base + (ca + 1) * d->in_s0;
      if (d->npack) {  // receive-buffer order (fused unpack)
        ia = pack_index(d, k, ca + c1 * d->pack_l1);
        ib = pack_index(d, k, ca + 1 + c1 * d->pack_l1);
      }
      cpx<T> A = {T(0), T(0)}, B = {T(0), T(0)};
      if (ca < d->C0r) A = gload<T>(in, ia);
      if (ca + 1 < d->C0r) B = gload<T>(in, ib);
      if (k == 0 || 2 * k == n) {
        A.y = T(0);
        B.y = T(0);
      }
      lds[k * ldj + t * ldt] = cpx<T>{A.x - B.y, -(A.y + B.x)};
      if (k > 0 && 2 * k < n) lds[(n - k) * ldj + t * ldt] = cpx<T>{A.x + B.y, A.y - B.x};
    }
    __syncthreads();
  }
  if constexpr (IO == IO_R2R) dtt_fill<T>(d, in, lds, c0b, c1, n);

  // real-to-real stages keep the line in LDS from the fill to the store:
  // their passes are those of a complex transform
  constexpr int PIO = IO == IO_R2R ? IO_C2C : IO;
  const int np = d->npass;
  for (int p = 0; p < np; ++p) {
    const bool sg = (p == 0) && (IO != IO_C2R) && (IO != IO_R2R);
    const bool dg = (p == np - 1) && (IO != IO_R2C) && (IO != IO_R2R) && (np >= 2);
    dispatch_pass<T, E, PIO>(d, io, p, sg, dg, c0b, c1);
  }

  if constexpr (IO == IO_R2R) {
    dtt_store<T>(d, in, out, lds, c0b, c1);
  } else if constexpr (IO == IO_R2C) {
    const T half = T(0.5);
    for (int w = tid; w < TX * nh; w += nthr) {
      const int t = w & tmask, k = w >> lg;
      const long long ca = 2 * (c0b + t);
      if (ca >= d->C0r) continue;
      const cpx<T> zk = lds[k * ldj + t * ldt];
      const int km = (k == 0) ? 0 : n - k;
      const cpx<T> zm = lds[km * ldj + t * ldt];
      const long long base = c1 * d->out_s1 + (long long)k * d->out_sj;
      long long oa_i = base + ca * d->out_s0, ob_i = base + (ca + 1) * d->out_s0;
      if (d->npack) {  // send-buffer order (fused pack)
        oa_i = pack_index(d, k, ca + c1 * d->pack_l1);
        ob_i = pack_index(d, k, ca + 1 + c1 * d->pack_l1);
      }
      gstore<T>(out, oa_i, cpx<T>{(zk.x + zm.x) * half, (zk.y - zm.y) * half});
      if (ca + 1 < d->C0r) gstore<T>(out, ob_i, cpx<T>{(zk.y + zm.y) * half, (zm.x - zk.x) * half});
    }
  } else {
    if (np < 2) {
      for (int w = tid; w < TX * n; w += nthr) {
        const int t = w & tmask, j = w >> lg;
        pstore<T, IO>(d, io, c0b + t, c1, j, lds[j * ldj + t * ldt]);
      }
    }
  }
}

template <typename T, int E, int IO>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(general_wpe<T>())))
    axis_fft_kernel(const AxisDesc *__restrict__ d, const void *in, void *out, int conj) {
  const AxisIo io = {in, out, conj};
  const long long ntiles = (long long)d->tiles0 * d->C1;
  const long long tile = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, ntiles);
  if (tile >= ntiles) return;  // whole block, before any barrier
  axis_tile<T, E, IO>(desc_ptr(d), io, tile);
}

// ---- many grids in one launch ---------------------------------------------------
// The persistent batch kernel on the general kernel's tile function (the
// queue and phase protocol: batch_loop.h).
template <typename T, int E>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(general_wpe<T>())))
    batch_kernel(const BatchQueue q) {
  batch_loop(q, [&](const BatchGroup &g, long long tile, unsigned long long *head)
                    __attribute__((always_inline)) {
    unsigned long long r = 0;
    if (head && threadIdx.x == 0) r = atomicAdd(head, 1ull);
    const AxisIo io = {g.in, g.out, g.conj};
    axis_tile<T, E, IO_C2C>(desc_ptr(g.d), io, tile);
    return r;
  });
}

__global__ void batch_reset_kernel(const BatchQueue q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < q.nphase * kBatchQueues) q.done[(long long)i * kHeadStride] = 0ull;
  if (i < kBatchQueues) q.heads[i * kHeadStride] = 0ull;
  if (i == 0) *q.err = 0u;
}

int batch_prepare(int prec_bytes, int nthreads, size_t lds, int *blocks_per_cu) {
  hipError_t e;
  if (prec_bytes == 8) {
    e = hipFuncSetAttribute((const void *)batch_kernel<double, kRegCap>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_kernel<double, kRegCap>,
                                                       nthreads, lds);
  } else {
    e = hipFuncSetAttribute((const void *)batch_kernel<float, kRegCap>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_kernel<float, kRegCap>,
                                                       nthreads, lds);
  }
  return hip_check(e, "batch kernel set-up");
}

int launch_batch_reset(const BatchQueue &q, hipStream_t st) {
  const int n = q.nphase * kBatchQueues > kBatchQueues ? q.nphase * kBatchQueues : kBatchQueues;
  hipLaunchKernelGGL(batch_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, st, q);
  return hip_check(hipGetLastError(), "batch reset launch");
}

int launch_batch(int prec_bytes, const BatchQueue &q, int blocks, int nthreads, size_t lds,
                 hipStream_t st) {
  const int rst = launch_batch_reset(q, st);
  if (rst) return rst;
  if (prec_bytes == 8)
    hipLaunchKernelGGL((batch_kernel<double, kRegCap>), dim3(blocks), dim3(nthreads), lds, st, q);
  else
    hipLaunchKernelGGL((batch_kernel<float, kRegCap>), dim3(blocks), dim3(nthreads), lds, st, q);
  return hip_check(hipGetLastError(), "batch kernel launch");
}

// ---- many grids, one launch per phase and kernel -------------------------------
// The alternative executor of a batch: per phase, the grids that run a stage
// on the same kernel share one launch whose workgroups cover all their tiles
// (XCD-aware order); a tile's grid is found by binary search over tile0, and
// the kernel boundary between phases orders the stages.
// The general kernel over the tiles of several grids, each with its own
// descriptor (block size and LDS at least every grid's planned shape).
template <typename T, int E>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(general_wpe<T>())))
    general_multi_kernel(const MultiGroup *__restrict__ g, int ng, long long tiles) {
  const long long it = xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, tiles);
  if (it >= tiles) return;  // whole block, before any barrier
  const int gi = find_multi_group(g, ng, it);
  const AxisIo io = {g[gi].in, g[gi].out, g[gi].conj};
  axis_tile<T, E, IO_C2C>(desc_ptr(g[gi].d), io, it - g[gi].tile0);
}

int launch_general_multi(int prec_bytes, const MultiGroup *g, int ng, long long tiles,
                         int nthreads, size_t lds, hipStream_t st) {
  const long long nb = xcd_grid(tiles);
  if (nb <= 0) return D2DHIP_SUCCESS;
  const long long gx = nb < (1LL << 20) ? nb : (1LL << 20);
  const long long gy = (nb + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "launch has too many tiles");
  if (prec_bytes == 8)
    hipLaunchKernelGGL((general_multi_kernel<double, kRegCap>), dim3((unsigned)gx, (unsigned)gy),
                       dim3(nthreads), lds, st, g, ng, tiles);
  else
    hipLaunchKernelGGL((general_multi_kernel<float, kRegCap>), dim3((unsigned)gx, (unsigned)gy),
                       dim3(nthreads), lds, st, g, ng, tiles);
  return hip_check(hipGetLastError(), "general multi kernel launch");
}

int general_multi_prepare(int prec_bytes, size_t lds) {
  const void *f = prec_bytes == 8 ? (const void *)general_multi_kernel<double, kRegCap>
                                  : (const void *)general_multi_kernel<float, kRegCap>;
  return hip_check(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                   "hipFuncSetAttribute(general multi kernel)");
}

// Bluestein's chirp-z algorithm (L. Bluestein, 1970) for a line length nb
// with a prime factor the Stockham codelets do not cover:
//   y_k = c_k sum_j (x_j c_j) w_{k-j},  c_j = exp(-i pi j^2/nb), w_m = exp(i pi m^2/nb),
// the convolution done as two power-of-two FFTs of length M >= 2 nb - 1 in
// LDS with the DFT of w precomputed (bhat).  The IO pairing of real
// transforms wraps around the complex line transform as in axis_fft_kernel.
template <typename T, int E, int IO>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(general_wpe<T>())))
    bluestein_kernel(const AxisDesc *__restrict__ d, const void *in, void *out, int conj) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(lds_raw);
  const AxisIo io = {in, out, conj};
  const long long tiles0 = d->tiles0;
  const long long tile =
      xcd_tile((long long)blockIdx.x + (long long)blockIdx.y * gridDim.x, tiles0 * d->C1);
  if (tile >= tiles0 * d->C1) return;  // whole block, before any barrier
  const long long c1 = tile / tiles0;
  const long long c0b = (tile - c1 * tiles0) * d->TX;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int M = d->n, n = d->nblue, TX = d->TX, lg = d->log2TX, tmask = TX - 1;
  const int ldj = d->ldj, ldt = d->ldt, nh = d->nh;
  const void *chirp = d->chirp;
  const void *bhat = d->bhat;

  // line values (conjugated for inverse transforms) into LDS[0, n)
  if constexpr (IO == IO_R2R) {
    dtt_fill<T>(desc_ptr(d), in, lds, c0b, c1, n);
  } else if constexpr (IO == IO_C2R) {
    for (int w = tid; w < TX * nh; w += nthr) {
      const int t = w & tmask, k = w >> lg;
      const long long ca = 2 * (c0b + t);
      const long long base = c1 * d->in_s1 + (long long)k * d->in_sj;
      long long ia = base + ca * d->in_s0, ib = base + (ca + 1) * d->in_s0;
      if (d->npack) {  // receive-buffer order (fused unpack)
        ia = pack_index(desc_ptr(d), k, ca + c1 * d->pack_l1);
        ib = pack_index(desc_ptr(d), k, ca + 1 + c1 * d->pack_l1);
      }
      cpx<T> A = {T(0), T(0)}, B = {T(0), T(0)};
      if (ca < d->C0r) A = gload<T>(in, ia);
      if (ca + 1 < d->C0r) B = gload<T>(in, ib);
      if (k == 0 || 2 * k == n) {
        A.y = T(0);
        B.y = T(0);
      }
      lds[k * ldj + t * ldt] = cpx<T>{A.x - B.y, -(A.y + B.x)};
      if (k > 0 && 2 * k < n) lds[(n - k) * ldj + t * ldt] = cpx<T>{A.x + B.y, A.y - B.x};
    }
  } else {
    for (int w = tid; w < TX * n; w += nthr) {
      const int t = w & tmask, j = w >> lg;
      lds[j * ldj + t * ldt] = pload<T, IO>(desc_ptr(d), io, c0b + t, c1, j);
    }
  }
  __syncthreads();
  // a_j = x_j c_j, zero padded to M
  for (int w = tid; w < TX * M; w += nthr) {
    const int t = w & tmask, j = w >> lg;
    cpx<T> &e = lds[j * ldj + t * ldt];
    e = (j < n) ? cmul(e, gload<T>(chirp, j)) : cpx<T>{T(0), T(0)};
  }
  __syncthreads();
  for (int p = 0; p < d->npass; ++p) dispatch_pass<T, E, IO_C2C>(desc_ptr(d), io, p, false, false, c0b, c1);
  // conj(A * W): the inverse DFT as conj(DFT(conj(.)))
  for (int w = tid; w < TX * M; w += nthr) {
    const int t = w & tmask, j = w >> lg;
    cpx<T> &e = lds[j * ldj + t * ldt];
    e = cconj(cmul(e, gload<T>(bhat, j)));
  }
  __syncthreads();
  for (int p = 0; p < d->npass; ++p) dispatch_pass<T, E, IO_C2C>(desc_ptr(d), io, p, false, false, c0b, c1);
  // y_k = c_k conj(r_k) / M
  const T invM = T(1) / T(M);
  for (int w = tid; w < TX * n; w += nthr) {
    const int t = w & tmask, k = w >> lg;
    cpx<T> &e = lds[k * ldj + t * ldt];
    const cpx<T> y = cmul(cconj(e), gload<T>(chirp, k));
    e = cpx<T>{y.x * invM, y.y * invM};
  }
  __syncthreads();
  if constexpr (IO == IO_R2R) {
    dtt_store<T>(desc_ptr(d), in, out, lds, c0b, c1);
  } else if constexpr (IO == IO_R2C) {
    const T half = T(0.5);
    for (int w = tid; w < TX * nh; w += nthr) {
      const int t = w & tmask, k = w >> lg;
      const long long ca = 2 * (c0b + t);
      if (ca >= d->C0r) continue;
      const cpx<T> zk = lds[k * ldj + t * ldt];
      const int km = (k == 0) ? 0 : n - k;
      const cpx<T> zm = lds[km * ldj + t * ldt];
      const long long base = c1 * d->out_s1 + (long long)k * d->out_sj;
      long long oa_i = base + ca * d->out_s0, ob_i = base + (ca + 1) * d->out_s0;
      if (d->npack) {  // send-buffer order (fused pack)
        oa_i = pack_index(desc_ptr(d), k, ca + c1 * d->pack_l1);
        ob_i = pack_index(desc_ptr(d), k, ca + 1 + c1 * d->pack_l1);
      }
      gstore<T>(out, oa_i, cpx<T>{(zk.x + zm.x) * half, (zk.y - zm.y) * half});
      if (ca + 1 < d->C0r) gstore<T>(out, ob_i, cpx<T>{(zk.y + zm.y) * half, (zm.x - zk.x) * half});
    }
  } else {
    for (int w = tid; w < TX * n; w += nthr) {
      const int t = w & tmask, j = w >> lg;
      pstore<T, IO>(desc_ptr(d), io, c0b + t, c1, j, lds[j * ldj + t * ldt]);
    }
  }
}

template <typename T, int IO>
static hipError_t set_lds_limit() {
  hipError_t e = hipFuncSetAttribute((const void *)axis_fft_kernel<T, kRegCap, IO>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)bluestein_kernel<T, kRegCap, IO>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

int axis_kernel_prepare(int prec_bytes) {
  hipError_t e;
  if (prec_bytes == 8) {
    e = set_lds_limit<double, IO_C2C>();
    if (e == hipSuccess) e = set_lds_limit<double, IO_R2C>();
    if (e == hipSuccess) e = set_lds_limit<double, IO_C2R>();
    if (e == hipSuccess) e = set_lds_limit<double, IO_R2R>();
  } else {
    e = set_lds_limit<float, IO_C2C>();
    if (e == hipSuccess) e = set_lds_limit<float, IO_R2C>();
    if (e == hipSuccess) e = set_lds_limit<float, IO_C2R>();
    if (e == hipSuccess) e = set_lds_limit<float, IO_R2R>();
  }
  return hip_check(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

// Launches one axis stage.  The tiles0 * C1 tiles are spread over gridDim.x
// (at most 2^20 blocks) and gridDim.y; surplus blocks exit at once.
template <typename T, int IO>
static int launch_t(const AxisDesc *dd, const AxisDesc &h, const AxisIo &io, const AxisShape &s,
                    hipStream_t st) {
  const long long ntiles = xcd_grid((long long)h.tiles0 * h.C1);
  if (ntiles <= 0) return D2DHIP_SUCCESS;
  const long long gx = ntiles < (1LL << 20) ? ntiles : (1LL << 20);
  const long long gy = (ntiles + gx - 1) / gx;
  if (gy > 65535) return set_error(D2DHIP_ERR_UNSUPPORTED, "axis stage has too many tiles");
  if (h.blue)
    hipLaunchKernelGGL((bluestein_kernel<T, kRegCap, IO>), dim3((unsigned)gx, (unsigned)gy),
                       dim3(s.nthreads), s.lds_bytes, st, dd, io.in, io.out, io.conj);
  else
    hipLaunchKernelGGL((axis_fft_kernel<T, kRegCap, IO>), dim3((unsigned)gx, (unsigned)gy),
                       dim3(s.nthreads), s.lds_bytes, st, dd, io.in, io.out, io.conj);
  return hip_check(hipGetLastError(), "axis_fft_kernel launch");
}

int launch_axis(int prec_bytes, int io, const AxisDesc *d_desc, const AxisDesc &h_desc,
                const AxisIo &io_ptrs, const AxisShape &s, hipStream_t st) {
  if (prec_bytes == 8) {
    if (io == IO_C2C) return launch_t<double, IO_C2C>(d_desc, h_desc, io_ptrs, s, st);
    if (io == IO_R2C) return launch_t<double, IO_R2C>(d_desc, h_desc, io_ptrs, s, st);
    if (io == IO_R2R) return launch_t<double, IO_R2R>(d_desc, h_desc, io_ptrs, s, st);
    return launch_t<double, IO_C2R>(d_desc, h_desc, io_ptrs, s, st);
  }
  if (io == IO_C2C) return launch_t<float, IO_C2C>(d_desc, h_desc, io_ptrs, s, st);
  if (io == IO_R2C) return launch_t<float, IO_R2C>(d_desc, h_desc, io_ptrs, s, st);
  if (io == IO_R2R) return launch_t<float, IO_R2R>(d_desc, h_desc, io_ptrs, s, st);
  return launch_t<float, IO_C2R>(d_desc, h_desc, io_ptrs, s, st);
}

}  // namespace d2dhip
