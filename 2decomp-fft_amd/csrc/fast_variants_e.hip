// fast_variants_e.hip -- registered specialised kernels, part 5 of 5: the
// lengths 100, 120, 200, 240, 300, 360, 400, 480, 500, 600, 896, 1200 and 1600
// (FP64 and FP32; strided columns in 64-256-byte rows, contiguous lines).
#include "fast_impl.h"

namespace d2dhip {

using R4_5_5 = RList<4, 5, 5>;       // 100
using R8_15 = RList<8, 15>;          // 120
using R8_5_5 = RList<8, 5, 5>;       // 200
using R16_15 = RList<16, 15>;        // 240
using R12_5_5 = RList<12, 5, 5>;     // 300
using R8_9_5 = RList<8, 9, 5>;       // 360
using R16_5_5 = RList<16, 5, 5>;     // 400
using R16_6_5 = RList<16, 6, 5>;     // 480
using R4_5_5_5 = RList<4, 5, 5, 5>;  // 500
using R8_15_5 = RList<8, 15, 5>;     // 600
using R16_8_7 = RList<16, 8, 7>;     // 896
using R16_15_5 = RList<16, 15, 5>;   // 1200
using R16_10_10 = RList<16, 10, 10>; // 1600
using R8_8_5_5 = RList<8, 8, 5, 5>;   // 1600, 128 rows: 2 or 3 butterflies a lane

std::vector<FastVariant> fast_variants_e() {
  return {
    D2D_VARIANTS(double, 100, 16, 128, false, R4_5_5),
    D2D_VARIANTS(double, 100, 16, 128, true, R4_5_5),
    D2D_VARIANTS(double, 120, 16, 128, false, R8_15),
    D2D_VARIANTS(double, 120, 16, 128, true, R8_15),
    D2D_VARIANTS(double, 200, 8, 128, false, R8_5_5),
    D2D_VARIANTS(double, 200, 8, 128, true, R8_5_5),
    D2D_VARIANTS(double, 240, 8, 128, false, R16_15),
    D2D_VARIANTS(double, 240, 8, 128, true, R16_15),
    D2D_VARIANTS(double, 300, 8, 256, false, R12_5_5),
    D2D_VARIANTS(double, 300, 4, 128, true, R12_5_5),
    D2D_VARIANTS(double, 360, 8, 256, false, R8_9_5),
    D2D_VARIANTS(double, 360, 4, 128, true, R8_9_5),
    D2D_VARIANTS(double, 400, 8, 256, false, R16_5_5),
    D2D_VARIANTS(double, 400, 4, 128, true, R16_5_5),
    D2D_VARIANTS(double, 480, 8, 256, false, R16_6_5),
    D2D_VARIANTS(double, 480, 4, 128, true, R16_6_5),
    D2D_VARIANTS(double, 500, 8, 256, false, R4_5_5_5),
    D2D_VARIANTS(double, 500, 4, 128, true, R4_5_5_5),
    D2D_VARIANTS(double, 600, 8, 512, false, R8_15_5),
    D2D_VARIANTS(double, 600, 4, 256, true, R8_15_5),
    D2D_VARIANTS(float, 100, 32, 256, false, R4_5_5),
    D2D_VARIANTS(float, 100, 16, 128, true, R4_5_5),
    D2D_VARIANTS(float, 120, 32, 256, false, R8_15),
    D2D_VARIANTS(float, 120, 16, 128, true, R8_15),
    D2D_VARIANTS(float, 200, 16, 256, false, R8_5_5),
    D2D_VARIANTS(float, 200, 8, 128, true, R8_5_5),
    D2D_VARIANTS(float, 240, 16, 256, false, R16_15),
    D2D_VARIANTS(float, 240, 8, 128, true, R16_15),
    D2D_VARIANTS(float, 300, 16, 512, false, R12_5_5),
    D2D_VARIANTS(float, 300, 4, 128, true, R12_5_5),
    D2D_VARIANTS(float, 360, 16, 512, false, R8_9_5),
    D2D_VARIANTS(float, 360, 4, 128, true, R8_9_5),
    D2D_VARIANTS(float, 400, 16, 512, false, R16_5_5),
    D2D_VARIANTS(float, 400, 4, 128, true, R16_5_5),
    D2D_VARIANTS(float, 480, 16, 512, false, R16_6_5),
    D2D_VARIANTS(float, 480, 4, 128, true, R16_6_5),
    D2D_VARIANTS(float, 500, 16, 512, false, R4_5_5_5),
    D2D_VARIANTS(float, 500, 4, 128, true, R4_5_5_5),
    D2D_VARIANTS(float, 600, 16, 512, false, R8_15_5),
    D2D_VARIANTS(float, 600, 4, 256, true, R8_15_5),
    D2D_VARIANTS(double, 896, 8, 512, false, R16_8_7),
    D2D_VARIANTS(double, 896, 4, 256, false, R16_8_7),
    D2D_VARIANTS(double, 896, 4, 256, true, R16_8_7),
    D2D_SPLIT_C2C(double, 1200, 8, 1024, false, R16_15_5, 4),
    D2D_SPLIT_C2C(double, 1200, 8, 512, false, R16_15_5, 4),
    D2D_VARIANTS(double, 1200, 4, 256, false, R16_15_5),
    D2D_VARIANTS(double, 1200, 4, 256, true, R16_15_5),
    D2D_VARIANTS(double, 1200, 2, 256, true, R16_15_5),
    D2D_SPLIT_C2C(double, 1600, 8, 1024, false, R16_10_10, 4),
    make_variant<double, 1600, 8, 1024, IO_C2C, false, R8_8_5_5, 1, 4>(
        "double n=1600 tx=8 nt=1024 c2c xs w4 r8855"),
    make_variant<double, 1600, 2, 256, IO_C2C, true, R8_8_5_5>("double n=1600 tx=2 nt=256 c2c r8855"),
    make_variant<double, 1600, 2, 256, IO_R2C, true, R8_8_5_5>("double n=1600 tx=2 nt=256 r2c r8855"),
    make_variant<double, 1600, 2, 256, IO_C2R, true, R8_8_5_5>("double n=1600 tx=2 nt=256 c2r r8855"),
    make_variant<double, 1600, 2, 256, IO_R2C, true, R8_8_5_5, 1, 4>(
        "double n=1600 tx=2 nt=256 r2c xs w4 r8855"),
    make_variant<double, 1600, 2, 256, IO_C2R, true, R8_8_5_5, 1, 4>(
        "double n=1600 tx=2 nt=256 c2r xs w4 r8855"),
    D2D_SPLIT_C2C(double, 1600, 8, 512, false, R16_10_10, 4),
    D2D_VARIANTS(double, 1600, 4, 256, false, R16_10_10),
    D2D_VARIANTS(double, 1600, 2, 256, true, R16_10_10),
    D2D_VARIANTS(float, 896, 16, 512, false, R16_8_7),
    D2D_VARIANTS(float, 896, 4, 256, true, R16_8_7),
    D2D_VARIANTS(float, 1200, 8, 512, false, R16_15_5),
    D2D_VARIANTS(float, 1200, 4, 256, true, R16_15_5),
    D2D_VARIANTS(float, 1600, 8, 512, false, R16_10_10),
    D2D_VARIANTS(float, 1600, 4, 256, true, R16_10_10),
  };
}

}  // namespace d2dhip
