// psc_shapes_f32_small.hip -- launch shapes of the FP32 lattices, N = 16 .. 128.
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_f32_small(std::vector<ShapeEntry>& v) {
    v.push_back(PSC_SHAPE(PSC_FP32, float, 16, "r4.4_tc16_l64", Radices<4, 4>, 16, 64));
    v.push_back(PSC_SHAPE(PSC_FP32, float, 32, "r8.4_tc32_l64", Radices<8, 4>, 32, 64));
    v.push_back(PSC_SHAPE(PSC_FP32, float, 64, "r8.8_tc32_l32", Radices<8, 8>, 32, 32));
    v.push_back(PSC_SHAPE(PSC_FP32, float, 128, "r16.8_tc32_l32", Radices<16, 8>, 32, 32));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 128, "r16.8_tc16_l16", Radices<16, 8>, 16, 16)));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 128, "r8.4.4_tc32_l16", Radices<8, 4, 4>, 32, 16)));
}
}  // namespace psc
