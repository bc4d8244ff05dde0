// psc_shapes_f32_256.hip -- launch shapes of the FP32 256^3 lattice (and the
// 2N shapes of the dealiased 128^3 stepper).
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_f32_256(std::vector<ShapeEntry>& v) {
    v.push_back(PSC_SHAPE(PSC_FP32, float, 256, "r16.16_tc32_l16", Radices<16, 16>, 32, 16));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 256, "r16.16_tc16_l16", Radices<16, 16>, 16, 16)));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 256, "r16.16_tc32_l8", Radices<16, 16>, 32, 8)));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP32, float, 256, "r8.8.4_tc16_l8", Radices<8, 8, 4>, 16, 8)));
}
}  // namespace psc
