// Variational MAC viscosity on the GPU (native/include/fluidity/host/viscosity.h:47-160;
// native/src/viscosity.cc:762-946, 1331-1650).
//
// AssembleFullCellMacStrain3D builds the symmetric-gradient strain rows of uniform full cells (three
// normal rows per cell, then the interior shear rows with harmonic edge viscosity), in the reference's
// order and with its quadrature weights, bit for bit. SolveMacViscousVelocityBlock3D solves the
// backward-Euler velocity block (M + dt E^T W E) u = rhs by Jacobi PCG in FP64 with the reference's
// stopping rule on the true residual. Same records, statuses and error order as the reference.
#pragma once

#include <array>
#include <cstddef>
#include <span>
#include <vector>

#include "fluidity/base/mac_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"

namespace fluidity {

struct MacViscosityCell3D {
  CellIndex3D cell_index;
  double dynamic_viscosity;  // Pa s, finite and nonnegative
};

struct MacStrainTerm3D {
  std::size_t face_slot;  // owned velocity slot of the grid
  double coefficient;     // 1/m
};

struct MacStrainRow3D {
  CellIndex3D low_cell;  // the cell of a normal row; the low-low cell of a shear edge
  FaceNormalAxis first_axis;
  FaceNormalAxis second_axis;  // equal to first_axis on a normal row
  std::array<MacStrainTerm3D, 8> terms;
  std::size_t term_count;    // active terms, 1 to 8
  double quadrature_weight;  // kg m^2/s, shear multiplicity included
  double prescribed_strain = 0.0;  // 1/s
};

// Rows of uniform full cells; cells unique and (z,y,x) ordered, every incident face owned.
// InvalidArgument for unordered cells or invalid viscosities, FailedPrecondition for a missing owner,
// OutOfRange for unrepresentable arithmetic or indices. Errors carry no rows.
StatusOr<std::vector<MacStrainRow3D>> AssembleFullCellMacStrain3D(SparseMacGridView3D grid,
                                                                  std::span<const MacViscosityCell3D> cells);
Status HipAssembleFullCellMacStrain3D(HipOperatorRuntime3D& runtime, SparseMacGridView3D grid,
                                      std::span<const MacViscosityCell3D> cells, std::vector<MacStrainRow3D>& rows,
                                      HipOperatorMeasurement& measurement);

struct MacViscousVelocityBlockInputs3D {
  std::span<const MacStrainRow3D> strain_rows;
  std::span<const double> face_inertias;  // kg, slot order; zero marks an eliminated face
  double step_duration;                   // nonnegative seconds
};

struct MacViscousVelocityBlockOptions3D {
  std::size_t max_iterations = 256;
  double relative_tolerance = 1e-10;  // relative to the largest |rhs|
};

struct MacViscousVelocityBlockSolution3D {
  std::vector<double> velocity;  // m/s, slot order; eliminated slots are zero
  std::size_t iterations;
  double max_residual;  // largest |true momentum residual|, kg m/s
};

// InvalidArgument for inconsistent extents, data or options, OutOfRange for unrepresentable arithmetic,
// Aborted for a non-positive curvature or an exhausted budget. Errors carry no solution.
StatusOr<MacViscousVelocityBlockSolution3D> SolveMacViscousVelocityBlock3D(const MacViscousVelocityBlockInputs3D& inputs,
                                                                           std::span<const double> rhs,
                                                                           MacViscousVelocityBlockOptions3D options = {});
StatusOr<MacViscousVelocityBlockSolution3D> HipSolveMacViscousVelocityBlock3D(
    HipOperatorRuntime3D& runtime, const MacViscousVelocityBlockInputs3D& inputs, std::span<const double> rhs,
    MacViscousVelocityBlockOptions3D options, HipOperatorMeasurement& measurement);

}  // namespace fluidity
