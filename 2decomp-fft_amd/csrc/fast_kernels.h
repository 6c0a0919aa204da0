// fast_kernels.h -- registry of the compile-time specialised axis kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "internal.h"

namespace d2dhip {

struct FastVariant {
  int rb = 8;      // bytes per real (8 FP64, 4 FP32)
  int mf = 0;      // 1: split-BF16 radix-16 passes on the matrix cores
  int io = 0;      // IO_C2C / IO_R2C / IO_C2R
  int lines = 0;   // 1: contiguous lines (x axis), 0: strided columns
  int n = 0;       // FFT length
  int TX = 1, NT = 64;
  size_t lds_bytes = 0;
  int ldj = 1, ldt = 1;
  std::vector<int> radices;
  int (*launch)(const AxisDesc *dd, const AxisDesc &h, const AxisIo &io, hipStream_t st) = nullptr;
  hipError_t (*prepare)() = nullptr;
  // the same kernel over the tiles of several grids (MultiGroup list)
  int (*launch_multi)(const MultiGroup *g, int ng, long long tiles, hipStream_t st) = nullptr;
  hipError_t (*prepare_multi)() = nullptr;
  const char *name = "";
};

// All specialised variants compiled into the library.
const std::vector<FastVariant> &fast_registry();

// The persistent batch kernel on specialised tiles: one-wave c2c tile
// functions, one per (length, contiguous lines or strided columns); a queue
// group's kind is its index here.  name: the registered FastVariant with
// the same kernel body (its plan descriptor fits the persistent kernel).
struct PersistKind {
  int n = 0, TX = 1, lines = 0;
  size_t lds_bytes = 0;
  std::string name;
};
const std::vector<PersistKind> &persist_kinds(int rb);
int batch_fast_prepare(int rb, size_t lds, int *blocks_per_cu);
int launch_batch_fast(int rb, const BatchQueue &q, int blocks, size_t lds, hipStream_t st);
// The split-BF16 matrix-core variants (mfma_kernels.hip).
std::vector<FastVariant> mf_variants();

}  // namespace d2dhip
