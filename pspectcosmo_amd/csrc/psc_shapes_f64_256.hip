// psc_shapes_f64_256.hip -- launch shapes of the FP64 256^3 lattice.
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_f64_256(std::vector<ShapeEntry>& v) {
    v.push_back(PSC_SHAPE(PSC_FP64, double, 256, "r16.16_tc16_l8", Radices<16, 16>, 16, 8));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP64, double, 256, "r8.8.4_tc16_l8", Radices<8, 8, 4>, 16, 8)));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP64, double, 256, "r16.16_tc8_l8", Radices<16, 16>, 8, 8)));
}
}  // namespace psc
