// fast_variants_a.hip -- registered specialised kernels, part 1 of 5:
// 1024-point FP64 kernels (every exchange form), the fused pack / unpack twins, 512 and the 2048-point lines.
#include "fast_impl.h"

namespace d2dhip {

std::vector<FastVariant> fast_variants_a() {
  return {
    // FP64, 1024: strided columns and contiguous lines
    D2D_VARIANTS(double, 1024, 4, 256, false, R16_16_4),
    D2D_VARIANTS(double, 1024, 2, 128, false, R16_16_4),
    D2D_VARIANTS(double, 1024, 8, 512, false, R16_16_4),
    D2D_VARIANTS(double, 1024, 8, 256, false, R16_16_4),
    D2D_VARIANTS(double, 1024, 1, 64, false, R16_16_4),
    D2D_VARIANTS(double, 1024, 4, 256, true, R16_16_4),
    D2D_VARIANTS(double, 1024, 2, 128, true, R16_16_4),
    // y-pencil stages fused with the pack / unpack of the y-pencil's exchange
    D2D_YPACK(double, 1024, 16, 1024, R16_16_4, 1, 4, " xs w4"),
    D2D_YPACK(double, 1024, 8, 512, R16_16_4, 1, 4, " xs w4"),
    D2D_YPACK(double, 1024, 8, 512, R16_16_4, 0, 1, ""),
    D2D_YPACK(float, 1024, 16, 1024, R16_16_4, 1, 8, " xs w8"),
    D2D_YPACK(float, 1024, 8, 512, R16_16_4, 0, 1, ""),
    D2D_YPACK(double, 1024, 16, 1024, R16_16_4, 2, 4, " xsh w4"),
    D2D_YPACK(double, 1024, 8, 512, R16_16_4, 2, 4, " xsh w4"),
    D2D_YPACK(float, 1024, 16, 1024, R16_16_4, 2, 8, " xsh w8"),
    D2D_YPACK(float, 1024, 8, 512, R16_16_4, 2, 6, " xsh w6"),
    D2D_YPACK(double, 512, 8, 512, R8_8_8, 0, 1, ""),
    D2D_YPACK(float, 512, 16, 512, R8_8_8, 0, 1, ""),
    D2D_SPLIT_C2C(double, 1024, 2, 128, false, R16_16_4, 4),
    D2D_SPLIT_C2C(double, 1024, 4, 256, false, R16_16_4, 4),
    D2D_SPLIT_C2C(double, 1024, 8, 512, false, R16_16_4, 4),
    D2D_SPLIT_C2C(double, 1024, 16, 1024, false, R16_16_4, 4),
    D2D_SHUF_C2C(double, 1024, 8, 512, R16_16_4, 4),
    D2D_SHUF_C2C(double, 1024, 16, 1024, R16_16_4, 4),
    D2D_SPLIT_C2C(double, 1024, 4, 256, false, R16_16_4, 3),
    D2D_SPLIT_C2C(double, 1024, 8, 512, false, R16_16_4, 3),
    D2D_SPLIT_ALL(double, 1024, 2, 128, true, R16_16_4, 3),
    D2D_SPLIT_ALL(double, 1024, 4, 256, true, R16_16_4, 3),
    D2D_SPLIT_ALL(double, 1024, 8, 512, true, R16_16_4, 3),
    // FP64, 512
    // FP64 2048 lines (the x stages of 2048-point pencils; the 2-line
    // tile first: 24-25 ms per 2048^3 stage against 26 ms for 1 line)
    D2D_VARIANTS(double, 2048, 2, 256, true, R16_16_8),
    D2D_SPLIT_ALL(double, 2048, 2, 256, true, R16_16_8, 3),
    D2D_VARIANTS(double, 2048, 1, 128, true, R16_16_8),
    D2D_VARIANTS(double, 512, 8, 512, false, R8_8_8),
    D2D_VARIANTS(double, 512, 4, 256, false, R8_8_8),
    D2D_VARIANTS(double, 512, 4, 256, true, R8_8_8),
    D2D_VARIANTS(double, 512, 8, 512, true, R8_8_8),
  };
}

}  // namespace d2dhip
