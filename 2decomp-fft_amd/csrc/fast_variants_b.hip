// fast_variants_b.hip -- registered specialised kernels, part 2 of 5:
// FP32 1024 / 512, the 720 / 540 / 432 lengths and 384 / 640 / 768.
#include "fast_impl.h"

namespace d2dhip {

std::vector<FastVariant> fast_variants_b() {
  return {
    // FP32, 1024 and 512
    D2D_VARIANTS(float, 1024, 8, 512, false, R16_16_4),
    D2D_VARIANTS(float, 1024, 4, 256, false, R16_16_4),
    D2D_VARIANTS(float, 1024, 16, 1024, false, R16_16_4),
    D2D_SPLIT_C2C(float, 1024, 8, 512, false, R16_16_4, 6),
    D2D_SPLIT_C2C(float, 1024, 4, 256, false, R16_16_4, 6),
    D2D_SPLIT_C2C(float, 1024, 16, 1024, false, R16_16_4, 8),
    D2D_SHUF_C2C(float, 1024, 16, 1024, R16_16_4, 8),
    D2D_SHUF_C2C(float, 1024, 8, 512, R16_16_4, 6),
    D2D_SPLIT_ALL(float, 1024, 4, 256, true, R16_16_4, 5),
    D2D_VARIANTS(float, 1024, 8, 512, true, R16_16_4),
    D2D_VARIANTS(float, 1024, 4, 256, true, R16_16_4),
    D2D_VARIANTS(float, 512, 16, 512, false, R8_8_8),
    D2D_VARIANTS(float, 512, 8, 256, true, R8_8_8),
    // FP64, the 720 x 540 x 432 grid's lengths (lines of 720, columns of 540 and 432)
    D2D_VARIANTS(double, 720, 4, 256, true, R16_9_5),
    D2D_VARIANTS(double, 720, 2, 128, true, R16_9_5),
    D2D_VARIANTS(double, 720, 8, 256, false, R16_9_5),
    D2D_VARIANTS(double, 540, 8, 512, false, R4_9_3_5),
    D2D_VARIANTS(double, 540, 4, 256, false, R4_9_3_5),
    D2D_VARIANTS(double, 540, 4, 256, true, R4_9_3_5),
    D2D_VARIANTS(double, 432, 8, 512, false, R16_9_3),
    D2D_VARIANTS(double, 432, 4, 256, false, R16_9_3),
    D2D_VARIANTS(double, 432, 4, 256, true, R16_9_3),
    // common mixed-radix lengths 384 = 16*8*3, 640 = 16*8*5, 768 = 16*16*3
    D2D_VARIANTS(double, 384, 8, 256, false, R16_8_3),
    D2D_VARIANTS(double, 384, 4, 128, true, R16_8_3),
    D2D_VARIANTS(double, 640, 8, 512, false, R16_8_5),
    D2D_VARIANTS(double, 640, 4, 256, false, R16_8_5),
    D2D_VARIANTS(double, 640, 4, 256, true, R16_8_5),
    D2D_SPLIT_C2C(double, 640, 8, 512, false, R16_8_5, 4),
    D2D_VARIANTS(double, 768, 8, 512, false, R16_16_3),
    D2D_VARIANTS(double, 768, 4, 256, false, R16_16_3),
    D2D_VARIANTS(double, 768, 4, 256, true, R16_16_3),
    D2D_SPLIT_C2C(double, 768, 8, 512, false, R16_16_3, 4),
    D2D_VARIANTS(float, 384, 16, 512, false, R16_8_3),
    D2D_VARIANTS(float, 384, 4, 128, true, R16_8_3),
    D2D_VARIANTS(float, 640, 16, 512, false, R16_8_5),
    D2D_VARIANTS(float, 640, 4, 256, true, R16_8_5),
    D2D_VARIANTS(float, 768, 16, 1024, false, R16_16_3),
    D2D_VARIANTS(float, 768, 8, 512, false, R16_16_3),
    D2D_VARIANTS(float, 768, 4, 256, true, R16_16_3),
  };
}

}  // namespace d2dhip
