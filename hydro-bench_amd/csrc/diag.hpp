// diag.hpp -- device diagnostics of a mesh's state (diag.hip).
#pragma once
#include "kernels.hpp"

namespace hydro {

// Sums carried as double-double: mass, momentum x, momentum y, kinetic
// energy, potential energy, KLAS 10 inflow, KLAS 1 inflow.
constexpr int kDiagSums = 7;
// A workgroup's record: the sums as (hi, lo) pairs, then the counts of
// non-finite and of infinite depths, then the finite depths' min and max.
constexpr int kDiagRecord = 2 * kDiagSums + 4;
// Cells (and their four slots) per workgroup record.
constexpr int kDiagCellsPerBlock = 1024;

// Records of the allocation-free entry point's fixed buffer (meshes of up to
// kDiagScratchRecords * kDiagCellsPerBlock = 1,048,576 cells; 144 KiB).
constexpr int kDiagScratchRecords = 1024;

// Device address of that buffer on the current device (looked up on the
// first call per device).
hipError_t diag_scratch(double **out);

// Records the first pass writes for a mesh of C cells.
int diag_records(int C);

// out[0 .. kDiagSums + 4) on the device: the rounded sums, the two counts,
// H min and max.  `part` holds diag_records(CELL) records.
template <class R> hipError_t launch_diagnostics(const ExtView<R> &m, double *part, double *out, hipStream_t s);

}  // namespace hydro
