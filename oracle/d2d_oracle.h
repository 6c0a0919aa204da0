/*
 * d2d_oracle.h -- CPU restatement of 2DECOMP&FFT's single-rank FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Layout conventions follow the reference: Fortran column-major arrays,
 * (i, j, k) with i fastest, complex stored as interleaved (re, im) doubles.
 *
 * What each function follows in the reference (paths relative to the
 * reference checkout) and what pins it is listed in the .c files.
 */
#ifndef D2D_ORACLE_H
#define D2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Glassman general-N FFT (src/glassman.f90:29-108) ---------------- */
/* u: n complex (interleaved), work: n complex scratch. isign = -1 / +1. */
void oracle_spcfft(double *u, int n, int isign, double *work);

/* ---- generic-backend 1D batched drivers (src/fft_generic.f90) -------- */
void oracle_c2c_1m_x(double *inout, const int sz[3], int isign);
void oracle_c2c_1m_y(double *inout, const int sz[3], int isign);
void oracle_c2c_1m_z(double *inout, const int sz[3], int isign);
void oracle_r2c_1m_x(const double *in, const int isz[3], double *out, int d1);
void oracle_r2c_1m_z(const double *in, const int isz[3], double *out, int d3);
void oracle_c2r_1m_x(const double *in, int in_d1, double *out, const int osz[3]);
void oracle_c2r_1m_z(const double *in, int in_d3, double *out, const int osz[3]);

/* ---- 3D drivers at one rank (src/fft_common_3d.f90) ------------------ */
/* format: 1 = PHYSICAL_IN_X, 3 = PHYSICAL_IN_Z.  skip[3] may be NULL.     */
/* c2c: in/out are nx*ny*nz complex.  in is preserved (non-inplace mode). */
int oracle_fft3d_c2c(int format, int nx, int ny, int nz, const double *in,
                     double *out, int isign, const int *skip);
/* r2c: in nx*ny*nz real; out (nx/2+1)*ny*nz (X) or nx*ny*(nz/2+1) (Z).   */
int oracle_fft3d_r2c(int format, int nx, int ny, int nz, const double *in,
                     double *out, const int *skip);
/* c2r: in spectral complex (preserved), out nx*ny*nz real.               */
int oracle_fft3d_c2r(int format, int nx, int ny, int nz, const double *in,
                     double *out, const int *skip);

/* Wall-clock seconds per line of each generic stage of an r2c+c2r round
 * trip at the given grid, measured on `lines` lines per stage at the real
 * strides (PHYSICAL_IN_X).  t[0..5] = x-r2c, y-fwd, z-fwd, z-bwd, y-bwd,
 * x-c2r.  nthreads <= 0 uses OpenMP's default.  Returns threads used.    */
int oracle_sample_stage_times(int nx, int ny, int nz, int lines, int nthreads,
                              double t[6]);

/* ---- direct DFT in long double (pinning aid, tiny sizes only) -------- */
void oracle_dft3d_ld(int nx, int ny, int nz, const double *in, double *out,
                     int isign);

/* ---- decomposition metadata (src/decomp_2d.f90:1016-1206) ------------ */
typedef struct {
  int xst[3], xen[3], xsz[3];
  int yst[3], yen[3], ysz[3];
  int zst[3], zen[3], zsz[3];
  int p_row, p_col;            /* dims(1), dims(2) */
  int x1dist[64], y1dist[64];  /* length p_row */
  int y2dist[64], z2dist[64];  /* length p_col */
  int64_t x1cnts[64], y1cnts[64], x1disp[64], y1disp[64];
  int64_t y2cnts[64], z2cnts[64], y2disp[64], z2disp[64];
  int64_t x1count, y1count, y2count, z2count;
  int even;
} oracle_decomp;

int oracle_distribute(int data1, int proc, int *st, int *en, int *sz);
/* coord_row = coord(1) in [0,p_row), coord_col = coord(2) in [0,p_col). */
int oracle_decomp_info_init(int nx, int ny, int nz, int p_row, int p_col,
                            int coord_row, int coord_col, oracle_decomp *d);

/* ---- pack / unpack maps (src/transpose_*.f90 mem_split_* / mem_merge_*)
 * dir: 0 = x->y, 1 = y->z, 2 = z->y, 3 = y->x.  esize = bytes per element
 * (8 real, 16 complex double).  even selects the padded MPI_ALLTOALL
 * offsets (m*count) instead of the alltoallv displacements.              */
void oracle_mem_split(int dir, const void *in, int n1, int n2, int n3,
                      void *out, const oracle_decomp *d, int esize, int even);
void oracle_mem_merge(int dir, const void *in, int n1, int n2, int n3,
                      void *out, const oracle_decomp *d, int esize, int even);

#ifdef __cplusplus
}
#endif
#endif
