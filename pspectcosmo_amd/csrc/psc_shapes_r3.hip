// psc_shapes_r3.hip -- launch shapes of the lengths 3 * 2^k (48, 96, 192,
// 384; the reference's FFTW takes any N): the first stage of every line
// transform is an in-register DFT of length 12 (a radix-3 step over length-4
// transforms) and the later radices divide it.  The same shapes serve as the
// padded lattices of the dealiased stepper: N_pad = 3N/2 for N = 64, 128, 256
// and N_pad = 2N for N = 48, 96, 192; the 48-point padded lattice of N = 32
// has its own entry, whose 16-line row blocks stay within one padded plane.
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_r3(std::vector<ShapeEntry>& v) {
    v.push_back(PSC_SHAPE_ONLY(PSC_FP32, float, 48, "r12.4_tc16_l32", Radices<12, 4>, 16, 32));
    v.push_back(PSC_SHAPE(PSC_FP32, float, 96, "r12.4.2_tc16_l16", Radices<12, 4, 2>, 16, 16));
    v.push_back(PSC_SHAPE(PSC_FP32, float, 192, "r12.4.4_tc16_l16", Radices<12, 4, 4>, 16, 16));
    v.push_back(PSC_SHAPE(PSC_FP32, float, 384, "r12.4.4.2_tc16_l8", Radices<12, 4, 4, 2>, 16, 8));
    v.push_back(PSC_SHAPE_ONLY(PSC_FP64, double, 48, "r12.4_tc16_l32", Radices<12, 4>, 16, 32));
    v.push_back(PSC_SHAPE(PSC_FP64, double, 96, "r12.4.2_tc16_l16", Radices<12, 4, 2>, 16, 16));
    v.push_back(PSC_SHAPE(PSC_FP64, double, 192, "r12.4.4_tc16_l16", Radices<12, 4, 4>, 16, 16));
    v.push_back(PSC_SHAPE(PSC_FP64, double, 384, "r12.4.4.2_tc16_l8", Radices<12, 4, 4, 2>, 16, 8));
    using F48 = S<float, 48, Radices<12, 4>, 16, 16>;
    using D48 = S<double, 48, Radices<12, 4>, 16, 16>;
    v.push_back(PSC_SHAPE_PAD(PSC_FP32, "r12.4_tc16_l16", F48, F48));
    v.push_back(PSC_SHAPE_PAD(PSC_FP64, "r12.4_tc16_l16", D48, D48));
}
}  // namespace psc
