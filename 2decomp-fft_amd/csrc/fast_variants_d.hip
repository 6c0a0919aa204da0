// fast_variants_d.hip -- registered specialised kernels, part 4 of 5:
// the real-to-real (DCT / DST) tiles of (2^k + 1)-point grids.
#include "fast_impl.h"

namespace d2dhip {

std::vector<FastVariant> fast_variants_d() {
  return {
    // real-to-real stages of (2^k + 1)-point grids (k = 6..10)
    D2D_DTT(double, 1024, 4, 256, false, R16_16_4),
    D2D_DTT(double, 1024, 4, 256, true, R16_16_4),
    D2D_DTT(double, 512, 8, 512, false, R8_8_8),
    D2D_DTT(double, 512, 4, 256, true, R8_8_8),
    D2D_DTT(double, 256, 8, 128, false, R16_16),
    D2D_DTT(double, 256, 4, 64, true, R16_16),
    D2D_DTT(double, 128, 8, 64, false, R16_8),
    D2D_DTT(double, 128, 8, 64, true, R16_8),
    D2D_DTT(double, 64, 8, 64, false, R8_8),
    D2D_DTT(double, 64, 16, 128, true, R8_8),
    D2D_DTT(float, 1024, 8, 512, false, R16_16_4),
    D2D_DTT(float, 1024, 4, 256, true, R16_16_4),
    D2D_DTT(float, 512, 16, 512, false, R8_8_8),
    D2D_DTT(float, 512, 8, 256, true, R8_8_8),
    D2D_DTT(float, 256, 16, 256, false, R16_16),
    D2D_DTT(float, 256, 4, 64, true, R16_16),
    D2D_DTT(float, 128, 16, 128, false, R16_8),
    D2D_DTT(float, 128, 8, 64, true, R16_8),
    D2D_DTT(float, 64, 16, 128, false, R8_8),
    D2D_DTT(float, 64, 16, 128, true, R8_8),
    D2D_VARIANTS(float, 32, 16, 64, false, R8_4),
    D2D_VARIANTS(float, 32, 32, 128, true, R8_4),
    D2D_VARIANTS(float, 48, 16, 64, false, R12_4),
    D2D_VARIANTS(float, 48, 16, 64, true, R12_4),
    D2D_VARIANTS(float, 80, 16, 64, false, R4_4_5),
    D2D_VARIANTS(float, 80, 16, 64, true, R4_4_5),
    D2D_VARIANTS(float, 96, 16, 64, false, R12_8),
    D2D_VARIANTS(float, 96, 16, 64, true, R12_8),
    D2D_VARIANTS(float, 160, 16, 128, false, R4_4_10),
    D2D_VARIANTS(float, 160, 8, 64, true, R4_4_10),
    D2D_VARIANTS(float, 192, 8, 128, false, R4_4_12),
    D2D_VARIANTS(float, 192, 4, 64, true, R4_4_12),
    D2D_C2C(float, 256, 16, 256, true, R16_16),
    D2D_C2C(float, 192, 16, 256, false, R4_4_12),
    D2D_C2C(float, 192, 16, 256, true, R4_4_12),
    D2D_C2C(float, 160, 16, 256, false, R4_4_10),
    D2D_C2C(float, 160, 16, 256, true, R4_4_10),
    D2D_C2C(float, 128, 32, 256, false, R16_8),
    D2D_C2C(float, 128, 32, 256, true, R16_8),
    D2D_C2C(float, 96, 32, 256, false, R12_8),
    D2D_C2C(float, 96, 32, 256, true, R12_8),
    D2D_C2C(float, 80, 32, 256, false, R4_4_5),
    D2D_C2C(float, 80, 32, 256, true, R4_4_5),
    D2D_C2C(float, 64, 32, 256, false, R8_8),
    D2D_C2C(float, 64, 32, 256, true, R8_8),
    D2D_C2C(float, 48, 64, 256, false, R12_4),
    D2D_C2C(float, 48, 64, 256, true, R12_4),
    D2D_C2C(float, 32, 64, 256, false, R8_4),
    D2D_C2C(float, 32, 64, 256, true, R8_4),
  };
}

}  // namespace d2dhip
