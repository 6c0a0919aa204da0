/*
 * glassman.c -- CPU restatement of the reference's "generic" FFT backend.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline); never part of the
 * shipped library.
 *
 * Follows, operation for operation:
 *   - SPCFFT / SPCPFT, Glassman's general-N FFT: src/glassman.f90:29-108
 *     (smallest-prime-first factoring :45-58, ping-pong U/WORK with the
 *     copy-back when the pass count is odd :52-64, Horner evaluation of the
 *     length-C DFTs :93-102, OMEGA accumulated by repeated multiplication
 *     with DELTA :85-91 and :103, and the truncated 2*pi constant :84).
 *   - the batched 1D drivers c2c_1m_x/y/z, r2c_1m_x/z, c2r_1m_x/z:
 *     src/fft_generic.f90:112-384 (gather line -> spcfft -> scatter).
 *   - the one-rank 3D drivers fft_3d_c2c / fft_3d_r2c / fft_3d_c2r:
 *     src/fft_common_3d.f90:9-296, with dims(1) = dims(2) = 1 so that every
 *     transpose is the copy dst = src (src/transpose_x_to_y.f90:41-50 and
 *     the matching branches of the other three transpose files).
 *
 * Pinned by (see tests/test_oracle.py): numpy pocketfft golden fixtures
 * (tests/golden), the long-double direct DFT below on tiny grids, round-trip
 * identity, Parseval, Hermitian symmetry of r2c output, and the closed-form
 * spectrum of single Fourier modes.
 *
 * Built with -ffp-contract=off so each complex product is the textbook
 * (ac - bd, ad + bc) without fused multiply-adds.
 */
#include "d2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
  cplx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

/* SPCPFT(A, B, C, UIN(B,C,A), UOUT(B,A,C), ISIGN): src/glassman.f90:69-108 */
static void spcpft(int A, int B, int C, const cplx *uin, cplx *uout,
                   int isign) {
  /* ANGLE = 6.28318530717958_mytype / REAL(A * C, kind=mytype)  (:84) */
  const double angle = 6.28318530717958 / (double)(A * C);
  cplx omega = {1.0, 0.0};
  cplx delta;
  delta.re = cos(angle);
  delta.im = (isign == 1) ? sin(angle) : -sin(angle);
  for (int ic = 1; ic <= C; ++ic) {
    for (int ia = 1; ia <= A; ++ia) {
      for (int ib = 1; ib <= B; ++ib) {
        /* SUM = UIN(IB, C, IA) */
        cplx sum = uin[(ib - 1) + (size_t)B * (C - 1) + (size_t)B * C * (ia - 1)];
        for (int jcr = 2; jcr <= C; ++jcr) {
          const int jc = C + 1 - jcr;
          const cplx u =
              uin[(ib - 1) + (size_t)B * (jc - 1) + (size_t)B * C * (ia - 1)];
          const cplx p = cmul(omega, sum);
          sum.re = u.re + p.re;
          sum.im = u.im + p.im;
        }
        uout[(ib - 1) + (size_t)B * (ia - 1) + (size_t)B * A * (ic - 1)] = sum;
      }
      omega = cmul(delta, omega);
    }
  }
}

/* SPCFFT(U, N, ISIGN, WORK): src/glassman.f90:29-67 */
void oracle_spcfft(double *u_, int n, int isign, double *work_) {
  cplx *u = (cplx *)u_;
  cplx *work = (cplx *)work_;
  int a = 1, b = n, c = 1;
  int inu = 1;
  while (b > 1) {
    a = c * a;
    c = 2;
    while (b % c != 0) c = c + 1;
    b = b / c;
    if (inu)
      spcpft(a, b, c, u, work, isign);
    else
      spcpft(a, b, c, work, u, isign);
    inu = !inu;
  }
  if (!inu) {
    for (int i = 0; i < n; ++i) u[i] = work[i];
  }
}

/* ---- batched 1D drivers: src/fft_generic.f90 ---------------------------- */

#define IDX3(i, j, k, n1, n2) \
  ((size_t)(i) + (size_t)(n1) * ((size_t)(j) + (size_t)(n2) * (size_t)(k)))

/* c2c_1m_x: src/fft_generic.f90:112-142 */
void oracle_c2c_1m_x(double *inout_, const int sz[3], int isign) {
  cplx *inout = (cplx *)inout_;
  const int n1 = sz[0], n2 = sz[1], n3 = sz[2];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n1);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)n1);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < n3; ++k)
      for (int j = 0; j < n2; ++j) {
        for (int i = 0; i < n1; ++i) buf[i] = inout[IDX3(i, j, k, n1, n2)];
        oracle_spcfft((double *)buf, n1, isign, (double *)scr);
        for (int i = 0; i < n1; ++i) inout[IDX3(i, j, k, n1, n2)] = buf[i];
      }
    free(buf);
    free(scr);
  }
}

/* c2c_1m_y: src/fft_generic.f90:145-175 */
void oracle_c2c_1m_y(double *inout_, const int sz[3], int isign) {
  cplx *inout = (cplx *)inout_;
  const int n1 = sz[0], n2 = sz[1], n3 = sz[2];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n2);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)n2);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < n3; ++k)
      for (int i = 0; i < n1; ++i) {
        for (int j = 0; j < n2; ++j) buf[j] = inout[IDX3(i, j, k, n1, n2)];
        oracle_spcfft((double *)buf, n2, isign, (double *)scr);
        for (int j = 0; j < n2; ++j) inout[IDX3(i, j, k, n1, n2)] = buf[j];
      }
    free(buf);
    free(scr);
  }
}

/* c2c_1m_z: src/fft_generic.f90:178-208 */
void oracle_c2c_1m_z(double *inout_, const int sz[3], int isign) {
  cplx *inout = (cplx *)inout_;
  const int n1 = sz[0], n2 = sz[1], n3 = sz[2];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n3);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)n3);
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        for (int k = 0; k < n3; ++k) buf[k] = inout[IDX3(i, j, k, n1, n2)];
        oracle_spcfft((double *)buf, n3, isign, (double *)scr);
        for (int k = 0; k < n3; ++k) inout[IDX3(i, j, k, n1, n2)] = buf[k];
      }
    free(buf);
    free(scr);
  }
}

/* r2c_1m_x: src/fft_generic.f90:211-251.  in (s1,s2,s3) real, out (d1,s2,s3) */
void oracle_r2c_1m_x(const double *in, const int isz[3], double *out_, int d1) {
  cplx *out = (cplx *)out_;
  const int s1 = isz[0], s2 = isz[1], s3 = isz[2];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)s1);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)s1);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < s3; ++k)
      for (int j = 0; j < s2; ++j) {
        for (int i = 0; i < s1; ++i) {
          buf[i].re = in[IDX3(i, j, k, s1, s2)];
          buf[i].im = 0.0;
        }
        oracle_spcfft((double *)buf, s1, -1, (double *)scr);
        for (int i = 0; i < d1; ++i) out[IDX3(i, j, k, d1, s2)] = buf[i];
      }
    free(buf);
    free(scr);
  }
}

/* r2c_1m_z: src/fft_generic.f90:254-294.  in (s1,s2,s3) real, out (s1,s2,d3) */
void oracle_r2c_1m_z(const double *in, const int isz[3], double *out_, int d3) {
  cplx *out = (cplx *)out_;
  const int s1 = isz[0], s2 = isz[1], s3 = isz[2];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)s3);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)s3);
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < s2; ++j)
      for (int i = 0; i < s1; ++i) {
        for (int k = 0; k < s3; ++k) {
          buf[k].re = in[IDX3(i, j, k, s1, s2)];
          buf[k].im = 0.0;
        }
        oracle_spcfft((double *)buf, s3, -1, (double *)scr);
        for (int k = 0; k < d3; ++k) out[IDX3(i, j, k, s1, s2)] = buf[k];
      }
    free(buf);
    free(scr);
  }
}

/* c2r_1m_x: src/fft_generic.f90:297-344.  in (in_d1,d2,d3) complex with
 * in_d1 >= d1/2+1, out (d1,d2,d3) real.  Hermitian extension
 * buf(i) = conjg(buf(d1+2-i)) for i = d1/2+2..d1 (1-based), then the real
 * part of the inverse transform. */
void oracle_c2r_1m_x(const double *in_, int in_d1, double *out,
                     const int osz[3]) {
  const cplx *in = (const cplx *)in_;
  const int d1 = osz[0], d2 = osz[1], d3 = osz[2];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)d1);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)d1);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < d3; ++k)
      for (int j = 0; j < d2; ++j) {
        for (int i = 1; i <= d1 / 2 + 1; ++i)
          buf[i - 1] = in[IDX3(i - 1, j, k, in_d1, d2)];
        for (int i = d1 / 2 + 2; i <= d1; ++i) {
          buf[i - 1].re = buf[d1 + 2 - i - 1].re;
          buf[i - 1].im = -buf[d1 + 2 - i - 1].im;
        }
        oracle_spcfft((double *)buf, d1, 1, (double *)scr);
        for (int i = 0; i < d1; ++i) out[IDX3(i, j, k, d1, d2)] = buf[i].re;
      }
    free(buf);
    free(scr);
  }
}

/* c2r_1m_z: src/fft_generic.f90:347-384.  in (d1,d2,in_d3), out (d1,d2,d3). */
void oracle_c2r_1m_z(const double *in_, int in_d3, double *out,
                     const int osz[3]) {
  const cplx *in = (const cplx *)in_;
  const int d1 = osz[0], d2 = osz[1], d3 = osz[2];
  (void)in_d3;
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)d3);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)d3);
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < d2; ++j)
      for (int i = 0; i < d1; ++i) {
        for (int k = 1; k <= d3 / 2 + 1; ++k)
          buf[k - 1] = in[IDX3(i, j, k - 1, d1, d2)];
        for (int k = d3 / 2 + 2; k <= d3; ++k) {
          buf[k - 1].re = buf[d3 + 2 - k - 1].re;
          buf[k - 1].im = -buf[d3 + 2 - k - 1].im;
        }
        oracle_spcfft((double *)buf, d3, 1, (double *)scr);
        for (int k = 0; k < d3; ++k) out[IDX3(i, j, k, d1, d2)] = buf[k].re;
      }
    free(buf);
    free(scr);
  }
}

/* ---- one-rank 3D drivers: src/fft_common_3d.f90 ------------------------ */

static int skip_of(const int *skip, int axis) { return skip ? skip[axis] : 0; }

/* fft_3d_c2c: src/fft_common_3d.f90:9-116 (non-inplace; dims = (1,1)). */
int oracle_fft3d_c2c(int format, int nx, int ny, int nz, const double *in,
                     double *out, int isign, const int *skip) {
  if (format != 1 && format != 3) return -1;
  if (isign != -1 && isign != 1) return -2;
  const int sz[3] = {nx, ny, nz};
  const size_t n = (size_t)nx * ny * nz;
  /* wk1 = in (:35-38 / :78-81); the one-rank transposes make wk1 -> out a
   * copy, so the stages can run on `out` directly with the same result. */
  memcpy(out, in, n * 2 * sizeof(double));
  const int x_first = (format == 1 && isign == -1) || (format == 3 && isign == 1);
  if (x_first) {
    if (!skip_of(skip, 0)) oracle_c2c_1m_x(out, sz, isign);
    if (!skip_of(skip, 1)) oracle_c2c_1m_y(out, sz, isign);
    if (!skip_of(skip, 2)) oracle_c2c_1m_z(out, sz, isign);
  } else {
    if (!skip_of(skip, 2)) oracle_c2c_1m_z(out, sz, isign);
    if (!skip_of(skip, 1)) oracle_c2c_1m_y(out, sz, isign);
    if (!skip_of(skip, 0)) oracle_c2c_1m_x(out, sz, isign);
  }
  return 0;
}

/* fft_3d_r2c: src/fft_common_3d.f90:121-189 (dims = (1,1)). */
int oracle_fft3d_r2c(int format, int nx, int ny, int nz, const double *in,
                     double *out, const int *skip) {
  const int isz[3] = {nx, ny, nz};
  if (format == 1) {
    const int sp[3] = {nx / 2 + 1, ny, nz};
    oracle_r2c_1m_x(in, isz, out, sp[0]);          /* r2c_1m_x(in_r, wk13) */
    if (!skip_of(skip, 1)) oracle_c2c_1m_y(out, sp, -1);
    if (!skip_of(skip, 2)) oracle_c2c_1m_z(out, sp, -1);
  } else if (format == 3) {
    const int sp[3] = {nx, ny, nz / 2 + 1};
    oracle_r2c_1m_z(in, isz, out, sp[2]);          /* r2c_1m_z(in_r, wk13) */
    if (!skip_of(skip, 1)) oracle_c2c_1m_y(out, sp, -1);
    if (!skip_of(skip, 0)) oracle_c2c_1m_x(out, sp, -1);
  } else {
    return -1;
  }
  return 0;
}

/* fft_3d_c2r: src/fft_common_3d.f90:194-296 (non-inplace; dims = (1,1)). */
int oracle_fft3d_c2r(int format, int nx, int ny, int nz, const double *in,
                     double *out, const int *skip) {
  const int osz[3] = {nx, ny, nz};
  if (format == 1) {
    const int sp[3] = {nx / 2 + 1, ny, nz};
    const size_t ns = (size_t)sp[0] * sp[1] * sp[2];
    double *wk = (double *)malloc(ns * 2 * sizeof(double));
    if (!wk) return -3;
    memcpy(wk, in, ns * 2 * sizeof(double));      /* wk1 = in_c (:223-226) */
    if (!skip_of(skip, 2)) oracle_c2c_1m_z(wk, sp, 1);
    if (!skip_of(skip, 1)) oracle_c2c_1m_y(wk, sp, 1);
    oracle_c2r_1m_x(wk, sp[0], out, osz);
    free(wk);
  } else if (format == 3) {
    const int sp[3] = {nx, ny, nz / 2 + 1};
    const size_t ns = (size_t)sp[0] * sp[1] * sp[2];
    double *wk = (double *)malloc(ns * 2 * sizeof(double));
    if (!wk) return -3;
    memcpy(wk, in, ns * 2 * sizeof(double));      /* wk1 = in_c (:252-255) */
    if (!skip_of(skip, 0)) oracle_c2c_1m_x(wk, sp, 1);
    if (!skip_of(skip, 1)) oracle_c2c_1m_y(wk, sp, 1);
    oracle_c2r_1m_z(wk, sp[2], out, osz);
    free(wk);
  } else {
    return -1;
  }
  return 0;
}

/* ---- CPU-baseline sampler ------------------------------------------------
 * Times the generic backend's six per-line stages of an r2c+c2r round trip
 * (PHYSICAL_IN_X) on `lines` lines each, gathered and scattered at the real
 * strides of the full grid.  The full-size arrays are reserved with calloc
 * so that only the pages the sampled lines touch are ever committed; each
 * stage runs twice and the second (warm-page) run is timed. */
#include <time.h>
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int oracle_sample_stage_times(int nx, int ny, int nz, int lines, int nthreads,
                              double t[6]) {
  const int nxc = nx / 2 + 1;
  const size_t nreal = (size_t)nx * ny * nz;
  const size_t nspec = (size_t)nxc * ny * nz;
  double *rbuf = (double *)calloc(nreal, sizeof(double));
  cplx *cbuf = (cplx *)calloc(nspec, sizeof(cplx));
  if (!rbuf || !cbuf) {
    free(rbuf);
    free(cbuf);
    return -1;
  }
  int used = 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
  }
#endif
  /* Lines are taken as consecutive runs in the fastest batch index so the
   * touched pages stay few; each stage's lines match the reference's loop
   * order (src/fft_generic.f90 loops). */
  for (int s = 0; s < 6; ++s)
  for (int rep = 0; rep < 2; ++rep) { /* rep 0 touches the pages, rep 1 is timed */
    const int nl = lines;
    double t0 = now_s();
#pragma omp parallel
    {
      const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
      cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
      cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
#pragma omp for schedule(static)
      for (int l = 0; l < nl; ++l) {
        if (s == 0) { /* x r2c: line (j=l%ny, k=l/ny) */
          const int j = l % ny, k = (l / ny) % nz;
          for (int i = 0; i < nx; ++i) {
            buf[i].re = rbuf[IDX3(i, j, k, nx, ny)];
            buf[i].im = 0.0;
          }
          oracle_spcfft((double *)buf, nx, -1, (double *)scr);
          for (int i = 0; i < nxc; ++i) cbuf[IDX3(i, j, k, nxc, ny)] = buf[i];
        } else if (s == 1 || s == 4) { /* y c2c: line (i=l%nxc, k=l/nxc) */
          const int i = l % nxc, k = (l / nxc) % nz;
          for (int j = 0; j < ny; ++j) buf[j] = cbuf[IDX3(i, j, k, nxc, ny)];
          oracle_spcfft((double *)buf, ny, s == 1 ? -1 : 1, (double *)scr);
          for (int j = 0; j < ny; ++j) cbuf[IDX3(i, j, k, nxc, ny)] = buf[j];
        } else if (s == 2 || s == 3) { /* z c2c: line (i=l%nxc, j=l/nxc) */
          const int i = l % nxc, j = (l / nxc) % ny;
          for (int k = 0; k < nz; ++k) buf[k] = cbuf[IDX3(i, j, k, nxc, ny)];
          oracle_spcfft((double *)buf, nz, s == 2 ? -1 : 1, (double *)scr);
          for (int k = 0; k < nz; ++k) cbuf[IDX3(i, j, k, nxc, ny)] = buf[k];
        } else { /* x c2r */
          const int j = l % ny, k = (l / ny) % nz;
          for (int i = 1; i <= nx / 2 + 1; ++i)
            buf[i - 1] = cbuf[IDX3(i - 1, j, k, nxc, ny)];
          for (int i = nx / 2 + 2; i <= nx; ++i) {
            buf[i - 1].re = buf[nx + 2 - i - 1].re;
            buf[i - 1].im = -buf[nx + 2 - i - 1].im;
          }
          oracle_spcfft((double *)buf, nx, 1, (double *)scr);
          for (int i = 0; i < nx; ++i) rbuf[IDX3(i, j, k, nx, ny)] = buf[i].re;
        }
      }
      free(buf);
      free(scr);
    }
    t[s] = (now_s() - t0) / (double)nl;
  }
  free(rbuf);
  free(cbuf);
  return used;
}

/* ---- direct DFT in long double (pinning aid) ---------------------------
 * X(kx,ky,kz) = sum x(i,j,k) exp(isign*2*pi*i*(kx*i/nx + ky*j/ny + kz*k/nz)),
 * with the phase reduced exactly in integer arithmetic.  O(N^2): tiny grids
 * only. */
void oracle_dft3d_ld(int nx, int ny, int nz, const double *in, double *out,
                     int isign) {
  const long double twopi = 6.283185307179586476925286766559L;
  const size_t n = (size_t)nx * ny * nz;
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx) {
        long double sr = 0.0L, si = 0.0L;
        for (int k = 0; k < nz; ++k)
          for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
              /* phase = (kx*i/nx + ky*j/ny + kz*k/nz) mod 1, as an exact
               * fraction over L = nx*ny*nz */
              const long long L = (long long)nx * ny * nz;
              long long num = (long long)((long long)kx * i % nx) * ny * nz +
                              (long long)((long long)ky * j % ny) * nx * nz +
                              (long long)((long long)kz * k % nz) * nx * ny;
              num %= L;
              const long double ang = isign * twopi * (long double)num / (long double)L;
              const long double c = cosl(ang), s = sinl(ang);
              const size_t p = IDX3(i, j, k, nx, ny);
              const long double xr = in[2 * p], xi = in[2 * p + 1];
              sr += xr * c - xi * s;
              si += xr * s + xi * c;
            }
        const size_t q = IDX3(kx, ky, kz, nx, ny);
        out[2 * q] = (double)sr;
        out[2 * q + 1] = (double)si;
      }
  (void)n;
}
