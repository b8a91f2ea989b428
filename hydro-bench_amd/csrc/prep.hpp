// prep.hpp -- host-side mesh preparation: geometry preprocessing and the
// cell reordering of the fused paths.
#pragma once
#include <cstdint>
#include <vector>

namespace hydro {

// Reverse Cuthill-McKee order of the cell graph given by the 1-based slot
// neighbour array NAC[4*C] (0 = none).  Returns perm with perm[new] = old.
std::vector<int32_t> rcm_order(int C, const int32_t *NAC);

// Cells grouped into tiles of B consecutive positions, each grown breadth-
// first from a seed so that a tile is a compact patch holding most of its
// cells' neighbours.  The next tile starts from the free neighbour of the
// previous tile that comes first in `along` (a permutation, e.g. RCM), else
// from the first free cell of `along`; a tile whose growth runs out of
// connected cells continues from such a seed, so every tile but the last
// has exactly B cells.  perm[new] = old.
std::vector<int32_t> tile_order(int C, const int32_t *NAC, int B, const std::vector<int32_t> &along);

// Hilbert-curve order of cell centroids (xc, yc), perm[new] = old.
std::vector<int32_t> hilbert_order(int C, const double *xc, const double *yc);

// MeshData::preCalculate (cuda_native_impl/hydro-cal-src/src/mesh.cpp:280-372),
// in Real = double or float.  See hydro_precalculate in include/hydro.h.
template <class R>
void precalculate(int CELL, int NOD, const int32_t *NAP, R *XP, R *YP, const R *ZBC, R *Z, const R *FNC0,
                  R HM1, R *SIDE, R *SINF, R *COSF, R *SLCOS, R *SLSIN, R *AREA, R *ZB1, R *H, R *FNC);

}  // namespace hydro
