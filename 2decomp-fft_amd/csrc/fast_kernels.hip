// fast_kernels.hip -- the persistent batch kernel on specialised tiles and the
// registry of the specialised axis-stage kernels (fast_impl.h holds the
// kernels; fast_variants_*.hip instantiate the registered ones).
#include "fast_impl.h"

namespace d2dhip {

// ---- persistent batch kernel on specialised tiles ---------------------------------
// c2c tiles for the lengths of uneven-grid batches, on strided columns and
// contiguous lines, all of one block size per precision.  Every kind has
// the radix list of the registered variants of its length, so its results
// equal the plans' own executions on any specialised kernel bit for bit.
//   FP64: tiles of up to 35 KB (4 blocks, 8 waves per CU), at least 8
//         columns (128-byte rows) on strided axes, 2-wave blocks;
//   FP32: the same tile bytes, twice the columns (at least 16: 128-byte
//         rows), 4-wave blocks (16 waves per CU; 2.95 -> 2.78 ms on the
//         64-grid batch against 2-wave blocks).  FP64 on 4-wave blocks gave
//         results that differ in the last bits from the standalone kernels
//         of the same tiles, so it keeps 2-wave blocks.
// Some kinds leave lanes idle in a pass (N / r not a multiple of the rows
// per column).
template <int N_, int TX_, bool LINES_, class RL_>
struct PV {
  static constexpr int N = N_, TX = TX_;
  static constexpr bool LINES = LINES_;
  using RL = RL_;
};
template <int NT_, int WPE_, class... Vs>
struct PSet {
  static constexpr int NT = NT_, WPE = WPE_;
  static constexpr int size = sizeof...(Vs);
};
using PersistSetD =
    PSet<128, 2, PV<256, 8, false, R16_16>, PV<256, 8, true, R16_16>, PV<192, 8, false, R4_4_12>,
         PV<192, 8, true, R4_4_12>, PV<160, 8, false, R4_4_10>, PV<160, 8, true, R4_4_10>,
         PV<128, 16, false, R16_8>, PV<128, 16, true, R16_8>, PV<96, 16, false, R12_8>,
         PV<96, 16, true, R12_8>, PV<80, 16, false, R4_4_5>, PV<80, 16, true, R4_4_5>,
         PV<64, 16, false, R8_8>, PV<64, 16, true, R8_8>, PV<48, 32, false, R12_4>,
         PV<48, 32, true, R12_4>, PV<32, 32, false, R8_4>, PV<32, 32, true, R8_4>>;
using PersistSetF =
    PSet<256, 4, PV<256, 16, false, R16_16>, PV<256, 16, true, R16_16>, PV<192, 16, false, R4_4_12>,
         PV<192, 16, true, R4_4_12>, PV<160, 16, false, R4_4_10>, PV<160, 16, true, R4_4_10>,
         PV<128, 32, false, R16_8>, PV<128, 32, true, R16_8>, PV<96, 32, false, R12_8>,
         PV<96, 32, true, R12_8>, PV<80, 32, false, R4_4_5>, PV<80, 32, true, R4_4_5>,
         PV<64, 32, false, R8_8>, PV<64, 32, true, R8_8>, PV<48, 64, false, R12_4>,
         PV<48, 64, true, R12_4>, PV<32, 64, false, R8_4>, PV<32, 64, true, R8_4>>;
template <typename T>
struct PersistSetOf {
  using type = PersistSetD;
};
template <>
struct PersistSetOf<float> {
  using type = PersistSetF;
};

template <int K, class V, class... Vs>
struct PAt {
  using type = typename PAt<K - 1, Vs...>::type;
};
template <class V, class... Vs>
struct PAt<0, V, Vs...> {
  using type = V;
};

// One kind's tile as a function of its own, so that its registers are
// allocated for its body alone (inlined side by side, the bodies of all
// kinds share one allocation and spill).  Function arguments arrive in
// vector registers; they are wave-uniform, and re-deriving them with
// readfirstlane lets the body keep the descriptor and its index arithmetic
// in scalar registers, as a kernel does with its arguments.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// When head is not null, thread 0 also takes the next position of that
// queue counter (after the call's entry wait, so that the atomic's latency
// overlaps the tile's first loads) and returns it.
template <typename T, int NT, class V>
__device__ __noinline__ unsigned long long persist_kind_tile(DescPtr d, const void *in, void *out,
                                                             int conj, long long tile,
                                                             unsigned long long *head) {
  unsigned long long *uhead = (unsigned long long *)uniform_u64((unsigned long long)head);
  unsigned long long r = 0;
  if (uhead && threadIdx.x == 0) r = atomicAdd(uhead, 1ull);
  const DescPtr ud = (DescPtr)uniform_u64((unsigned long long)d);
  const void *uin = (const void *)uniform_u64((unsigned long long)in);
  void *uout = (void *)uniform_u64((unsigned long long)out);
  const int uconj = __builtin_amdgcn_readfirstlane(conj);
  const long long utile = (long long)uniform_u64((unsigned long long)tile);
  fast_tile<T, V::N, V::TX, NT, IO_C2C, V::LINES, typename V::RL, 0>(ud, uin, uout, uconj, utile);
  return r;
}

// Tile of kind `kind` (wave-uniform): a chain of uniform branches.
template <typename T, int NT, int K, class... Vs>
__device__ __forceinline__ unsigned long long persist_tile(int kind, DescPtr d, const void *in,
                                                           void *out, int conj, long long tile,
                                                           unsigned long long *head) {
  if constexpr (K < (int)sizeof...(Vs)) {
    using V = typename PAt<K, Vs...>::type;
    if (kind == K) return persist_kind_tile<T, NT, V>(d, in, out, conj, tile, head);
    return persist_tile<T, NT, K + 1, Vs...>(kind, d, in, out, conj, tile, head);
  }
  return 0;
}

template <typename T, int NT, int WPE, class... Vs>
__device__ __forceinline__ unsigned long long persist_run(const BatchGroup &g, long long tile,
                                                          unsigned long long *head,
                                                          PSet<NT, WPE, Vs...>) {
  const int kind = __builtin_amdgcn_readfirstlane(g.kind);
  return persist_tile<T, NT, 0, Vs...>(kind, desc_ptr(g.d), g.in, g.out, g.conj, tile, head);
}

// The waves-per-SIMD bound also holds for the per-kind functions it calls,
// which would otherwise take up to 256 registers each.
template <typename T>
__global__ void __launch_bounds__(PersistSetOf<T>::type::NT)
    __attribute__((amdgpu_waves_per_eu(PersistSetOf<T>::type::WPE)))
    batch_fast_kernel(const BatchQueue q) {
  batch_loop(q, [&](const BatchGroup &g, long long tile, unsigned long long *head)
                    __attribute__((always_inline)) {
    return persist_run<T>(g, tile, head, typename PersistSetOf<T>::type());
  });
}

template <typename T, int NT, int WPE, class... Vs>
static std::vector<PersistKind> persist_table(PSet<NT, WPE, Vs...>) {
  const char *tname = sizeof(T) == 8 ? "double" : "float";
  std::vector<PersistKind> v;
  auto add = [&](int n, int tx, bool lines, size_t lds) {
    char b[96];
    std::snprintf(b, sizeof(b), "%s n=%d tx=%d nt=%d c2c", tname, n, tx, NT);
    PersistKind k;
    k.name = b;
    k.n = n;
    k.TX = tx;
    k.lines = lines ? 1 : 0;
    k.lds_bytes = lds;
    v.push_back(k);
  };
  (add(Vs::N, Vs::TX, Vs::LINES, FastCfg<T, Vs::N, Vs::TX, NT, Vs::LINES, typename Vs::RL>::lds_bytes()), ...);
  return v;
}

const std::vector<PersistKind> &persist_kinds(int rb) {
  static const std::vector<PersistKind> d = persist_table<double>(PersistSetD());
  static const std::vector<PersistKind> f = persist_table<float>(PersistSetF());
  return rb == 8 ? d : f;
}

int batch_fast_prepare(int rb, size_t lds, int *blocks_per_cu) {
  hipError_t e;
  if (rb == 8) {
    e = hipFuncSetAttribute((const void *)batch_fast_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_fast_kernel<double>,
                                                       PersistSetD::NT, lds);
  } else {
    e = hipFuncSetAttribute((const void *)batch_fast_kernel<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_fast_kernel<float>,
                                                       PersistSetF::NT, lds);
  }
  return hip_check(e, "persistent batch kernel set-up");
}

int launch_batch_fast(int rb, const BatchQueue &q, int blocks, size_t lds, hipStream_t st) {
  const int rst = launch_batch_reset(q, st);
  if (rst) return rst;
  if (rb == 8)
    hipLaunchKernelGGL(batch_fast_kernel<double>, dim3(blocks), dim3(PersistSetD::NT), lds, st, q);
  else
    hipLaunchKernelGGL(batch_fast_kernel<float>, dim3(blocks), dim3(PersistSetF::NT), lds, st, q);
  return hip_check(hipGetLastError(), "persistent batch kernel launch");
}

// The registered variants, in candidate order (the first able to run a stage
// is its default): the parts compiled in fast_variants_*.hip, then the
// matrix-core kernels.
std::vector<FastVariant> fast_variants_a();
std::vector<FastVariant> fast_variants_b();
std::vector<FastVariant> fast_variants_c();
std::vector<FastVariant> fast_variants_d();
std::vector<FastVariant> fast_variants_e();

const std::vector<FastVariant> &fast_registry() {
  static const std::vector<FastVariant> reg = [] {
    std::vector<FastVariant> r;
    for (auto part : {fast_variants_a, fast_variants_b, fast_variants_c, fast_variants_d,
                      fast_variants_e}) {
      const std::vector<FastVariant> v = part();
      r.insert(r.end(), v.begin(), v.end());
    }
    for (const FastVariant &v : mf_variants()) r.push_back(v);
    return r;
  }();
  return reg;
}

}  // namespace d2dhip
