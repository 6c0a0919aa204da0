// internal.h -- shared declarations of libd2dhip (host side + kernel args).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/d2dhip.h"

namespace d2dhip {

constexpr int kMaxPass = 16;
// Register capacity (complex values per thread per Stockham pass) of the
// general axis kernel.  18 holds one radix-16/17/18 butterfly, two radix-8/9,
// four radix-4 and so on.
constexpr int kRegCap = 18;

// IO_R2R: a real-to-real trigonometric transform (DCT / DST) of each line,
// computed through a complex FFT of two real lines at once (of the line's
// length for DCT / DST II and III, of a symmetric extension for the others).
// IO_DTT2 / IO_DTT3 / IO_DTTX are the I/O forms of the specialised
// kernels' real-to-real stages: DCT-II / DST-II (a permuted first pass, a
// twiddled Hermitian split after the last), DCT-III / DST-III (a twiddled
// Hermitian merge before the first pass, a permuted last pass) and
// DCT-I / IV, DST-I / IV (the symmetric extension gathered by the first
// pass, the outputs picked from the last) and DCT-IV / DST-IV of even
// length L as an L/2-point DFT of one line (IO_DTT4: tile columns are single
// real lines, not pairs).
// IO_YPK / IO_YUN: a c2c stage on the columns of a y- or z-pencil whose
// output (YPK) is written in, or whose input (YUN) is read from, the
// alltoallv buffer order of the pencil's exchange (mem_split_yz /
// mem_split_yx / mem_split_zy, mem_merge_zy / mem_merge_xy / mem_merge_yz):
// the transpose's pack or unpack fused with the stage.  IO_YUP: both, the
// input in one exchange's receive order and the output in the next one's
// send order.
enum IoKind : int {
  IO_C2C = 0, IO_R2C = 1, IO_C2R = 2, IO_R2R = 3, IO_DTT2 = 4, IO_DTT3 = 5, IO_DTTX = 6,
  IO_DTT4 = 7, IO_YPK = 8, IO_YUN = 9, IO_YUP = 10
};

struct PassInfo {
  int r;       // radix
  int L;       // product of the radices of the earlier passes
  int tw_off;  // first twiddle (complex) of this pass
  float invL;  // 1/L for the quotient estimate
};

// One batched 1D FFT stage along one axis of a pencil.
//
// The stage sees the pencil as columns: column (c0, c1) starts at
// c0*s0 + c1*s1 and its n points are sj apart (all in elements of the
// array's own type).  A workgroup owns a tile of TX consecutive c0 of one
// c1 and all n points of each, staged in LDS.  For IO_R2C / IO_C2R the tile
// columns are pairs of real columns (2*c0, 2*c0+1) transformed together as
// one complex column and split / merged by Hermitian symmetry.
struct AxisDesc {
  const void *tw;  // per-pass twiddles, forward sign, complex of the plan type
  int n;
  int npass;
  PassInfo pass[kMaxPass];
  int TX, log2TX;
  int tiles0;         // ceil(C0 / TX)
  long long C0;       // tile-level columns along c0
  long long C0r;      // real columns along c0 (paired IO), else C0
  long long C1;       // columns along c1
  long long in_s0, in_s1, in_sj;
  long long out_s0, out_s1, out_sj;
  // Blocked columns (c2c stages): column c0 sits at c0 * s0 + (c0 >> bsh) * b0,
  // i.e. blocks of 2^bsh columns b0 + 2^bsh * s0 elements apart; 0 = plain.
  long long in_b0, out_b0;
  int bsh;
  int ldj, ldt;       // LDS strides (complex units) of point j and column t
  int nh;             // n/2 + 1 of the line length
  // Bluestein (chirp-z) stages: the line length is nblue, `n` and the passes
  // are those of the power-of-two convolution length M = n >= 2*nblue - 1.
  int blue;           // 1: Bluestein stage
  int nblue;          // line length of a Bluestein stage
  const void *chirp;  // exp(-i pi j^2 / nblue), j < nblue
  const void *bhat;   // DFT_M of the chirp sequence exp(+i pi m^2 / nblue), wrapped
  // Complex side of a line stage (r2c output, c2r input) in the x-pencil
  // send / receive-buffer order of a process grid (a transform fused with
  // the pack or unpack), or in a plan's blocked spectral layout: block m
  // of a line's points holds [b1, b1 + w) and puts point i of line l at
  // base - b1 + i + l * w; pack[2i .. 2i+1] = {base - b1, w} of the block
  // holding point i.  npack = 0: plain layout, > 0 the number of blocks.
  const long long *pack;
  int npack;
  long long pack_l1;  // line index l = c0 + c1 * pack_l1 of the packed order (npack != 0)
  long long pack_bs;  // npack < 0: the blocked layout's block stride (no table)
  // Columns of a y- or z-pencil in buffer order (fused pack / unpack):
  // packj = 1 on the output side, 2 on the input side, 3 on both (input
  // through pack, output through pack2), 0 plain.  The table holds one
  // entry per point j of the column axis, pack[2j .. 2j+1] = {dj, dc1}:
  // point j of column (c0, c1) sits at the plain index plus dj + c1 * dc1
  // (dj, dc1 those of the exchange block that holds j; one load per point,
  // no search, so that the loads of an unpacking stage wait on a single
  // cached read).  npack / npack2: the number of entries.
  int packj;
  const long long *pack2;
  int npack2;
  // IO_R2R stages: transform type (1-4 DCT-I..IV, 5-8 DST-I..IV), the
  // transformed points [dtt_first, dtt_first + dtt_len) of each line of
  // dtt_n points (the others pass through); `n` is the DFT length (dtt_len,
  // or the length of the line's symmetric extension).
  int dtt, dtt_first, dtt_len, dtt_n;
  // (cos, sin)(pi k / (2 dtt_len)), k < dtt_len: the twiddles of the
  // half-length DCT-II / III and DST-II / III stages (n = dtt_len); for
  // IO_DTT4 stages exp(-i pi (4j+1) / (4 dtt_len)), j < n, then
  // exp(-i pi k / dtt_len), k < n (n = dtt_len / 2)
  const void *dtw;
};

#ifdef __HIPCC__  // (a host compiler, for the host-only sanitizer build, skips these)
// Index of complex point i of line l in the packed order of d (npack > 0):
// one 16-byte read of the point's entry {base - b1, w} (those of the block
// holding i), so that a fused unpack's loads wait on a single cached read.
// npack < 0: the blocked spectral layout, blocks of 2^bsh points pack_bs apart.
__device__ __forceinline__ long long pack_index(const __attribute__((address_space(4))) AxisDesc *d,
                                                int i, long long l) {
  if (d->npack < 0)
    return (long long)(i >> d->bsh) * d->pack_bs + (i & ((1 << d->bsh) - 1)) + (l << d->bsh);
  typedef long long ll2 __attribute__((ext_vector_type(2)));
  const __attribute__((address_space(1))) ll2 *pk =
      (const __attribute__((address_space(1))) ll2 *)d->pack;
  const ll2 e = pk[i];
  return e.x + i + l * e.y;
}
#endif

#ifdef __HIPCC__  // (a host compiler, for the host-only sanitizer build, skips these)
// Offset of point j of outer column c1 from its plain position in the
// block buffer order of d (packj != 0): one 16-byte read of the point's entry.
__device__ __forceinline__ long long packj_offset_in(const long long *table, int j, long long c1) {
  typedef long long ll2 __attribute__((ext_vector_type(2)));
  const __attribute__((address_space(1))) ll2 *pk =
      (const __attribute__((address_space(1))) ll2 *)table;
  const ll2 e = pk[j];
  return e.x + c1 * e.y;
}
__device__ __forceinline__ long long packj_offset(const __attribute__((address_space(4))) AxisDesc *d,
                                                  int j, long long c1) {
  return packj_offset_in(d->pack, j, c1);
}
// the output side of a packj = 3 stage
__device__ __forceinline__ long long packj_offset2(const __attribute__((address_space(4))) AxisDesc *d,
                                                   int j, long long c1) {
  return packj_offset_in(d->pack2, j, c1);
}
#endif

// A stage descriptor as the kernels read it: constant address space (it is
// not written while a kernel runs), so its fields are scalar loads even when
// the pointer itself was loaded from memory (batch and multi-grid launches)
// or passed to a non-inlined function.
using DescPtr = const __attribute__((address_space(4))) AxisDesc *;
__device__ __forceinline__ DescPtr desc_ptr(const AxisDesc *d) { return (DescPtr)d; }

// Per-launch arguments: the descriptor lives in device memory (built once per
// plan), the data pointers and the direction come with each launch.
struct AxisIo {
  const void *in;
  void *out;
  int conj;  // 1: inverse transform (conjugate on load and store)
};

// One group of work items of a persistent batch launch: all tiles of one
// stage of one grid, ceil(ntiles / per_item) queue items.
struct BatchGroup {
  const AxisDesc *d;  // device descriptor of the stage for the tile function
  const void *in;
  void *out;
  int conj;
  int phase;          // stage phase: every item of phase p-1 completes before phase p starts
  int per_item;       // consecutive tiles one queue item covers
  long long ntiles;
  int kind;           // specialised tile function (persistent kernel on specialised tiles)
};

// The device queues of a persistent batch launch.  Each phase's items are
// dealt to kBatchQueues queues in chunks of a few consecutive items; block b
// takes items from queue b % kBatchQueues (blocks are placed on the 8 XCDs
// round robin, so each XCD mostly works its own queue and its own counter).
constexpr int kBatchQueues = 8;
constexpr int kHeadStride = 32;  // u64 between queue counters (256 bytes)
struct QueueEntry {
  int group;  // BatchGroup index
  int local;  // item within the group: tiles [local * per_item, ...)
  int phase;
  int pad;
};
struct BatchQueue {
  const BatchGroup *groups;
  const QueueEntry *entries;  // queue q: entries[off[q] .. off[q] + len[q])
  long long off[kBatchQueues];
  long long len[kBatchQueues];
  // items of phase p in queue q: phase_items[p * kBatchQueues + q]
  const unsigned long long *phase_items;
  int nphase;
  unsigned long long *heads;  // kBatchQueues counters, kHeadStride apart
  // completed items of phase p in queue q: done[(p * kBatchQueues + q) * kHeadStride]
  unsigned long long *done;
  unsigned int *err;
  long long wait_ticks;
};

// Launch shape chosen by the planner for one axis stage.
struct AxisShape {
  int TX = 1;
  int nthreads = 64;
  size_t lds_bytes = 0;
  int ldj = 1, ldt = 1;
};

// Per-axis plan: the radix schedule and twiddles of one FFT length.
struct AxisPlan {
  int n = 0;
  std::vector<int> radices;
  std::vector<PassInfo> passes;
  void *d_tw = nullptr;  // device twiddles (cpx<T>)
  size_t tw_bytes = 0;
};

// Workgroups of a launch are dealt to the 8 XCDs of an MI355X round-robin.
// Tile t of a stage goes to block xcd_block(t): each XCD then works through
// a contiguous range of tiles, so tiles that share 128-byte lines (TX * 16 B
// < 128) meet in one L2.  Launches use xcd_grid(ntiles) blocks; blocks past
// the last tile exit at once.
constexpr int kXcds = 8;
__host__ __device__ inline long long xcd_grid(long long ntiles) {
  return (ntiles + kXcds - 1) / kXcds * kXcds;
}
__device__ inline long long xcd_tile(long long block, long long ntiles) {
  const long long per = (ntiles + kXcds - 1) / kXcds;
  return (block % kXcds) * per + block / kXcds;
}

// Profiler regions: the reference's decomp_profiler_start / _end around its
// entry points (src/profiler_caliper.f90:156-177), here as rocTX ranges that
// `rocprofv3 --marker-trace` records, with the reference's region names and
// its four categories (src/profiler_caliper.f90:24-27).  A Region opens one
// for its scope when its category is on (api.cpp).
enum { PROF_TRANSPOSE = 1, PROF_IO = 2, PROF_FFT = 4, PROF_D2D = 8 };
struct Region {
  bool on;
  Region(int category, const char *name);
  ~Region();
  Region(const Region &) = delete;
  Region &operator=(const Region &) = delete;
};

// status helpers (api.cpp)
int set_error(int status, const std::string &msg);
int hip_check(hipError_t e, const char *what);

// axis kernels (fft_kernels.hip)
int launch_axis(int prec_bytes, int io, const AxisDesc *d_desc, const AxisDesc &h_desc,
                const AxisIo &io_ptrs, const AxisShape &s, hipStream_t st);
int axis_kernel_prepare(int prec_bytes);  // raise LDS limits once
int batch_prepare(int prec_bytes, int nthreads, size_t lds, int *blocks_per_cu);
// Clears the queue counters, the per-phase counters and the error flag (a
// kernel, not a memset, so that a captured graph replays it).
int launch_batch_reset(const BatchQueue &q, hipStream_t st);
int launch_batch(int prec_bytes, const BatchQueue &q, int blocks, int nthreads, size_t lds,
                 hipStream_t st);

// One grid's stage in a launch that covers several grids: its descriptor
// and buffers, and the index of its first tile among the launch's tiles.
struct MultiGroup {
  const AxisDesc *d;
  const void *in;
  void *out;
  int conj;
  long long tile0;
};

#ifdef __HIPCC__  // (skipped by a host compiler, as above)
// Group of launch tile `it` (groups sorted by tile0), wave-uniform.
__device__ inline int find_multi_group(const MultiGroup *__restrict__ g, int n, long long it) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].tile0 <= it) lo = mid; else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}
#endif

// The general kernel over the tiles of several grids (their own descriptors).
int launch_general_multi(int prec_bytes, const MultiGroup *g, int ng, long long tiles,
                         int nthreads, size_t lds, hipStream_t st);
int general_multi_prepare(int prec_bytes, size_t lds);

// copy / pack kernels (copy_kernels.hip)
// halo cells (halo_kernels.hip): in (s1, s2, s3) into the interior of arr
// (in may be null: exchange only), then the periodic axes' halo slabs
int launch_halo(int esize, const void *in, void *arr, int s1, int s2, int s3, int level,
                int ipencil, const int periodic[3], hipStream_t st);
int launch_pack(int dir, int split, int esize, const void *src, void *dst,
                const d2dhip_decomp *d, int even, hipStream_t st);
int launch_fill_uniform(void *dst, long long n, int elem, unsigned long long seed,
                        hipStream_t st);
int launch_column_copy(int elem, const void *src, void *dst, long long C0, long long C1,
                       long long s1, long long sj, int n, hipStream_t st);
int launch_fill_modal(void *dst, int nx, int ny, int nz, int elem,
                      const d2dhip_modes *m, hipStream_t st);

}  // namespace d2dhip
