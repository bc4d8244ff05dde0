// psc_shapes_pad768.hip -- the 768-point padded lattice of the dealiased
// stepper: N_pad = 3N/2 for N = 512 and N_pad = 2N for N = 384.  A padded
// shape only (768 is not an unpadded lattice length); the first stage of every
// line is the in-register length-12 DFT, then three radix-4 stages (the later
// radices divide the first, so 64 threads per line keep whole butterflies).
// FP64 keeps its column tile at 8 columns (96 KiB of LDS).
#include "psc_shapes.hpp"

namespace psc {
void register_shapes_pad768(std::vector<ShapeEntry>& v) {
    using F768 = S<float, 768, Radices<12, 4, 4, 4>, 16, 4>;
    using D768 = S<double, 768, Radices<12, 4, 4, 4>, 8, 4>;
    // FP32 default: x pass on 8-column tiles, 512-thread blocks with room for
    // its two-tile prefetch (24.9 ms/step for 512^3 on 768^3 against 27.7 with
    // 16 columns, whose 1024-thread x pass spills at one tile); y passes on 16
    // columns, whose 128-byte rows also take non-temporal accesses (1295 /
    // 1326 -> 1114 / 1150 us per pass)
    using F768_8 = S<float, 768, Radices<12, 4, 4, 4>, 8, 4>;
    v.push_back(PSC_SHAPE_PAD_Y(PSC_FP32, "r12.4.4.4_tc8_l4_y16", F768_8, F768, F768));
    PSC_EXTRA(v.push_back(PSC_SHAPE_PAD(PSC_FP32, "r12.4.4.4_tc8_l4", F768_8, F768)));
    PSC_EXTRA(v.push_back(PSC_SHAPE_PAD(PSC_FP32, "r12.4.4.4_tc16_l4", F768, F768)));
    v.push_back(PSC_SHAPE_PAD(PSC_FP64, "r12.4.4.4_tc8_l4", D768, D768));
}
}  // namespace psc
