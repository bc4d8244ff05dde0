// psc_shapes.hpp -- the launch-shape table: each psc_shapes_<group>.hip
// instantiates the engines of one group of shapes and registers them; the
// library concatenates the groups in a fixed order, so that the first entry
// of every (precision, N) is its default.  Built with
// PSC_DEFAULT_SHAPES_ONLY, a group registers its defaults only.
#pragma once
#include "psc_engine.hpp"

#ifdef PSC_DEFAULT_SHAPES_ONLY
#define PSC_EXTRA(stmt)
#else
#define PSC_EXTRA(stmt) stmt
#endif

namespace psc {

using F512_888 = S<float, 512, Radices<8, 8, 8>, 16, 4>;
using F512_Y = S<float, 512, Radices<16, 16, 2>, 32, 4>;
using F512_1616 = S<float, 512, Radices<16, 16, 2>, 16, 8>;
using D512_888 = S<double, 512, Radices<8, 8, 8>, 8, 4>;
using D512_Y = S<double, 512, Radices<16, 16, 2>, 16, 4>;

void register_shapes_f32_small(std::vector<ShapeEntry>& v);
void register_shapes_f32_256(std::vector<ShapeEntry>& v);
void register_shapes_f32_512a(std::vector<ShapeEntry>& v);
void register_shapes_f32_512b(std::vector<ShapeEntry>& v);
void register_shapes_f64_small(std::vector<ShapeEntry>& v);
void register_shapes_f64_256(std::vector<ShapeEntry>& v);
void register_shapes_f64_512(std::vector<ShapeEntry>& v);
void register_shapes_r3(std::vector<ShapeEntry>& v);
void register_shapes_pad768(std::vector<ShapeEntry>& v);

}  // namespace psc
