// psc_shapes_f32_512b.hip -- further FP32 512-point shapes for the autotuner.
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_f32_512b(std::vector<ShapeEntry>& v) {
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 512, "r8.8.8_tc16_l4", Radices<8, 8, 8>, 16, 4)));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 512, "r16.16.2_tc16_l8", Radices<16, 16, 2>, 16, 8)));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 512, "r8.8.8_tc8_l8", Radices<8, 8, 8>, 8, 8)));
    (void)v;
}
}  // namespace psc
