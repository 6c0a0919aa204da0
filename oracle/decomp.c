/*
 * decomp.c -- CPU restatement of 2DECOMP&FFT's decomposition metadata and
 * of the eight pack/unpack maps that feed MPI_ALLTOALL(V).
 *
 * TEST INFRASTRUCTURE ONLY (checker); never part of the shipped library.
 *
 * Follows:
 *   distribute     src/decomp_2d.f90:1070-1101 (floor blocks first, the
 *                  nu = data1 mod proc larger blocks last)
 *   partition      src/decomp_2d.f90:1016-1064 (pdim 1 local, 2 over
 *                  dims(1) at coord(1), 3 over dims(2) at coord(2))
 *   get_dist       src/decomp_2d.f90:1108-1130
 *   prepare_buffer src/decomp_2d.f90:1135-1206 (alltoallv counts and
 *                  displacements; EVEN padded counts from the last blocks)
 *   decomp_info_init src/decomp_2d.f90:382-490 (grid check :399-405,
 *                  x/y/z partitions :413-418, even flag :443-454)
 *   mem_split_xy / mem_merge_xy  src/transpose_x_to_y.f90:266-500
 *   mem_split_yz / mem_merge_yz  src/transpose_y_to_z.f90:370-606
 *   mem_split_zy / mem_merge_zy  src/transpose_z_to_y.f90:370-606
 *   mem_split_yx / mem_merge_yx  src/transpose_y_to_x.f90:266-500
 *
 * Pinned by (tests/test_oracle.py): the init_test partition-sum invariant
 * (examples/init_test/init_test.f90:84-112), hand-checked distributions
 * from the reference's own comments (17 over 4 = 4,4,4,5 at
 * src/decomp_2d.f90:1104-1106), and a full multi-rank transpose simulated
 * with every rank's split -> alltoallv -> merge on the integer index field
 * u = i + (j-1)nx + (k-1)nx*ny of examples/test2d/test2d.f90:77, which must
 * land exactly on each rank's target pencil.
 */
#include <stdint.h>
#include <string.h>

#include "d2d_oracle.h"

int oracle_distribute(int data1, int proc, int *st, int *en, int *sz) {
  if (proc <= 0) return -1;
  int size1 = data1 / proc;
  const int nu = data1 - size1 * proc;
  const int nl = proc - nu;
  st[0] = 1;
  sz[0] = size1;
  en[0] = size1;
  for (int i = 1; i <= nl - 1; ++i) {
    st[i] = st[i - 1] + size1;
    sz[i] = size1;
    en[i] = en[i - 1] + size1;
  }
  size1 = size1 + 1;
  for (int i = nl; i <= proc - 1; ++i) {
    st[i] = en[i - 1] + 1;
    sz[i] = size1;
    en[i] = en[i - 1] + size1;
  }
  if (en[proc - 1] != data1) return -2;
  if (sz[proc - 1] != data1 - st[proc - 1] + 1) return -3;
  return 0;
}

static int partition(int nx, int ny, int nz, const int pdim[3], int p_row,
                     int p_col, int crow, int ccol, int lst[3], int len[3],
                     int lsz[3]) {
  int st[64], en[64], sz[64];
  const int g[3] = {nx, ny, nz};
  for (int i = 0; i < 3; ++i) {
    if (pdim[i] == 1) {
      lst[i] = 1;
      len[i] = g[i];
      lsz[i] = g[i];
    } else if (pdim[i] == 2) {
      if (oracle_distribute(g[i], p_row, st, en, sz)) return -1;
      lst[i] = st[crow];
      len[i] = en[crow];
      lsz[i] = sz[crow];
    } else {
      if (oracle_distribute(g[i], p_col, st, en, sz)) return -1;
      lst[i] = st[ccol];
      len[i] = en[ccol];
      lsz[i] = sz[ccol];
    }
  }
  return 0;
}

int oracle_decomp_info_init(int nx, int ny, int nz, int p_row, int p_col,
                            int crow, int ccol, oracle_decomp *d) {
  int st[64], en[64];
  memset(d, 0, sizeof(*d));
  if (p_row < 1 || p_col < 1 || p_row > 64 || p_col > 64) return -1;
  if (crow < 0 || crow >= p_row || ccol < 0 || ccol >= p_col) return -1;
  /* src/decomp_2d.f90:399-405 (errorcode 6) */
  if (nx < p_row || ny < p_row || ny < p_col || nz < p_col) return 6;
  d->p_row = p_row;
  d->p_col = p_col;
  /* get_dist */
  if (oracle_distribute(nx, p_row, st, en, d->x1dist)) return -2;
  if (oracle_distribute(ny, p_row, st, en, d->y1dist)) return -2;
  if (oracle_distribute(ny, p_col, st, en, d->y2dist)) return -2;
  if (oracle_distribute(nz, p_col, st, en, d->z2dist)) return -2;
  /* partition ×3 */
  const int px[3] = {1, 2, 3}, py[3] = {2, 1, 3}, pz[3] = {2, 3, 1};
  if (partition(nx, ny, nz, px, p_row, p_col, crow, ccol, d->xst, d->xen, d->xsz)) return -3;
  if (partition(nx, ny, nz, py, p_row, p_col, crow, ccol, d->yst, d->yen, d->ysz)) return -3;
  if (partition(nx, ny, nz, pz, p_row, p_col, crow, ccol, d->zst, d->zen, d->zsz)) return -3;
  /* prepare_buffer */
  for (int i = 0; i < p_row; ++i) {
    d->x1cnts[i] = (int64_t)d->x1dist[i] * d->xsz[1] * d->xsz[2];
    d->y1cnts[i] = (int64_t)d->ysz[0] * d->y1dist[i] * d->ysz[2];
    d->x1disp[i] = i == 0 ? 0 : d->x1disp[i - 1] + d->x1cnts[i - 1];
    d->y1disp[i] = i == 0 ? 0 : d->y1disp[i - 1] + d->y1cnts[i - 1];
  }
  for (int i = 0; i < p_col; ++i) {
    d->y2cnts[i] = (int64_t)d->ysz[0] * d->y2dist[i] * d->ysz[2];
    d->z2cnts[i] = (int64_t)d->zsz[0] * d->zsz[1] * d->z2dist[i];
    d->y2disp[i] = i == 0 ? 0 : d->y2disp[i - 1] + d->y2cnts[i - 1];
    d->z2disp[i] = i == 0 ? 0 : d->z2disp[i - 1] + d->z2cnts[i - 1];
  }
  d->x1count = (int64_t)d->x1dist[p_row - 1] * d->y1dist[p_row - 1] * d->xsz[2];
  d->y1count = d->x1count;
  d->y2count = (int64_t)d->y2dist[p_col - 1] * d->z2dist[p_col - 1] * d->zsz[0];
  d->z2count = d->y2count;
  d->even = (nx % p_row == 0 && ny % p_row == 0 && ny % p_col == 0 && nz % p_col == 0);
  return 0;
}

/* Element copy helper. */
static inline void cp(char *dst, int64_t di, const char *src, int64_t si, int es) {
  memcpy(dst + di * es, src + si * es, (size_t)es);
}

/* Which dist array, which displacements/count drive each side of a map. */
static void map_params(int dir, int split, const oracle_decomp *d, int *np,
                       const int **dist, const int64_t **disp,
                       int64_t *count) {
  switch (dir) {
    case 0: /* x -> y */
      *np = d->p_row;
      *dist = split ? d->x1dist : d->y1dist;
      *disp = split ? d->x1disp : d->y1disp;
      *count = split ? d->x1count : d->y1count;
      break;
    case 1: /* y -> z */
      *np = d->p_col;
      *dist = split ? d->y2dist : d->z2dist;
      *disp = split ? d->y2disp : d->z2disp;
      *count = split ? d->y2count : d->z2count;
      break;
    case 2: /* z -> y */
      *np = d->p_col;
      *dist = split ? d->z2dist : d->y2dist;
      *disp = split ? d->z2disp : d->y2disp;
      *count = split ? d->z2count : d->y2count;
      break;
    default: /* y -> x */
      *np = d->p_row;
      *dist = split ? d->y1dist : d->x1dist;
      *disp = split ? d->y1disp : d->x1disp;
      *count = split ? d->y1count : d->x1count;
      break;
  }
}

/* The block axis of each map: 0 = i, 1 = j, 2 = k.
 * split_xy i, merge_xy j; split_yz j, merge_yz k; split_zy k, merge_zy j;
 * split_yx j, merge_yx i. */
static int block_axis(int dir, int split) {
  static const int ax[4][2] = {{1, 0}, {2, 1}, {1, 2}, {0, 1}}; /* [dir][split] */
  return ax[dir][split];
}

/* Position in the flat buffer of element (i,j,k) (0-based) of block m whose
 * range along the block axis starts at b1 with width w:
 *   block on i: (i-b1) + j*w + k*n2*w
 *   block on j: i + (j-b1)*n1 + k*w*n1
 *   block on k: i + j*n1 + (k-b1)*n1*n2
 * which restates the pos formulas of every mem_split_* / mem_merge_* loop. */
static inline int64_t flat_pos(int ax, int i, int j, int k, int b1, int w,
                               int n1, int n2) {
  if (ax == 0) return (int64_t)(i - b1) + (int64_t)j * w + (int64_t)k * n2 * w;
  if (ax == 1) return (int64_t)i + (int64_t)(j - b1) * n1 + (int64_t)k * w * n1;
  return (int64_t)i + (int64_t)j * n1 + (int64_t)(k - b1) * n1 * n2;
}

static void mem_map(int dir, int split, const void *in, int n1, int n2, int n3,
                    void *out, const oracle_decomp *d, int es, int even) {
  int np;
  const int *dist;
  const int64_t *disp;
  int64_t count;
  map_params(dir, split, d, &np, &dist, &disp, &count);
  const int ax = block_axis(dir, split);
  int b1 = 0;
  for (int m = 0; m < np; ++m) {
    const int w = dist[m];
    const int64_t init = even ? (int64_t)m * count : disp[m];
    const int lo[3] = {ax == 0 ? b1 : 0, ax == 1 ? b1 : 0, ax == 2 ? b1 : 0};
    const int hi[3] = {ax == 0 ? b1 + w : n1, ax == 1 ? b1 + w : n2,
                       ax == 2 ? b1 + w : n3};
    for (int k = lo[2]; k < hi[2]; ++k)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int i = lo[0]; i < hi[0]; ++i) {
          const int64_t a = (int64_t)i + (int64_t)n1 * ((int64_t)j + (int64_t)n2 * k);
          const int64_t p = init + flat_pos(ax, i, j, k, b1, w, n1, n2);
          if (split)
            cp((char *)out, p, (const char *)in, a, es);
          else
            cp((char *)out, a, (const char *)in, p, es);
        }
    b1 += w;
  }
}

void oracle_mem_split(int dir, const void *in, int n1, int n2, int n3,
                      void *out, const oracle_decomp *d, int esize, int even) {
  mem_map(dir, 1, in, n1, n2, n3, out, d, esize, even);
}

void oracle_mem_merge(int dir, const void *in, int n1, int n2, int n3,
                      void *out, const oracle_decomp *d, int esize, int even) {
  mem_map(dir, 0, in, n1, n2, n3, out, d, esize, even);
}
