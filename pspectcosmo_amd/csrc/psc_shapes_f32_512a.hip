// psc_shapes_f32_512a.hip -- default FP32 512-point shapes: the dealiased
// 256^3 stepper's 2N passes (z rows on the 8.8.8 plan) and the 512^3
// lattice (y passes on the 16.16.2 plan with 32 columns).
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_f32_512a(std::vector<ShapeEntry>& v) {
    v.push_back(PSC_SHAPE_PAD(PSC_FP32, "r16.16.2_tc16_l8_z8.8.8_l4", F512_1616, F512_888));
    v.push_back(PSC_SHAPE_Y(PSC_FP32, "r8.8.8_tc16_l4_y16.16.2_tc32", F512_888, F512_Y));
}
}  // namespace psc
