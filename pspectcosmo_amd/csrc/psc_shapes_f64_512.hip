// psc_shapes_f64_512.hip -- launch shapes of the FP64 512^3 lattice.
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_f64_512(std::vector<ShapeEntry>& v) {
    v.push_back(PSC_SHAPE_Y(PSC_FP64, "r8.8.8_tc8_l4_y16.16.2_tc16", D512_888, D512_Y));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP64, double, 512, "r8.8.8_tc8_l4", Radices<8, 8, 8>, 8, 4)));
    PSC_EXTRA(v.push_back(PSC_SHAPE(PSC_FP64, double, 512, "r16.16.2_tc8_l4", Radices<16, 16, 2>, 8, 4)));
}
}  // namespace psc
