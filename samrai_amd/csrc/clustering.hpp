// Berger-Rigoutsos clustering of tagged cells (host; clustering.cpp).
#pragma once

#include <vector>

namespace samrai_amd {
namespace mesh {

// Boxes (lo..., hi... per box, 2*dim ints each) covering the nonzero cells
// of tags, an array over [lo, hi] with the first index fastest.
std::vector<int> bergerRigoutsos(int dim, const int* lo, const int* hi,
                                 const unsigned char* tags, double efficiency_tolerance,
                                 double combine_efficiency, const int* min_box,
                                 const int* max_box);

}  // namespace mesh
}  // namespace samrai_amd
