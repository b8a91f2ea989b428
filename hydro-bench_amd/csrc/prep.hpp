// prep.hpp -- host-side mesh preparation: geometry preprocessing and the
// cell reordering of the fused paths.
#pragma once
#include <cstdint>
#include <string>
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
// Within each tile of B consecutive positions, cells with a neighbour in
// another tile first (their updates finish, and are published, first).
// With `shallow` (by original id, nonzero for a cell that is shallow or dry
// at planning time), each of the two groups lists its deep cells first, so
// the shallow ones share wavefronts, which then skip the friction terms.
void edge_cells_first(int C, const int32_t *NAC, int B, std::vector<int32_t> &perm,
                      const uint8_t *shallow = nullptr);

// Hilbert-curve order of cell centroids (xc, yc), perm[new] = old.
std::vector<int32_t> hilbert_order(int C, const double *xc, const double *yc);

// MeshData::preCalculate (cuda_native_impl/hydro-cal-src/src/mesh.cpp:280-372),
// in Real = double or float.  See hydro_precalculate in include/hydro.h.
template <class R>
void precalculate(int CELL, int NOD, const int32_t *NAP, R *XP, R *YP, const R *ZBC, R *Z, const R *FNC0,
                  R HM1, R *SIDE, R *SINF, R *COSF, R *SLCOS, R *SLSIN, R *AREA, R *ZB1, R *H, R *FNC);

// Inputs of take_boundary: the fields of hydro_boundary_desc (include/hydro.h).
struct BoundaryIn {
    int CELL, NDAYS, rows, MDT, NHQ;
    double DT, HM2;
    int NZ, NNZ0;
    const int32_t *MBZ, *NNZ;
    int NQ, NNQ0;
    const int32_t *MBQ, *NNQ;
    int NZQ, NNZQ0;
    const int32_t *MBZQ, *NNZQ;
    const void *ZBC, *SIDE;
    const int32_t *KLAS;
};

// MeshData::take_boundary_for_two_d (cuda_native_impl/hydro-cal-src/src/
// mesh.cpp:374-521) with the expansion to cells of cell.cpp:92-126: reads
// <bounde_dir>/NZ/NZnnnn.DAT, NQ/NQnnnn.DAT and NZQ/NZQnnnn.DAT and writes
// rows x CELL column-major series and CELL x NHQ rating curves.  Returns false
// with `err` set when a file is missing or malformed.
template <class R>
bool take_boundary(const BoundaryIn &d, const char *bounde_dir, R *QT, R *DQT, R *ZT, R *DZT, int32_t *NHQ1, R *ZW,
                   R *QW, std::string &err);

// D[jt][c] = (V[jt+1][c] - V[jt][c]) / K0 for jt < rows_d, 0 where V has no
// row jt+1 (mesh.cpp:449-464 on an expanded series).
template <class R> void boundary_increments(int CELL, int rows_v, int rows_d, int K0, const R *V, R *D);

}  // namespace hydro
