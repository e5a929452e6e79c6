// The viscosity operators over the C-ABI (fl_assemble_full_cell_strain, fl_solve_viscous_block).
#include "fluidity/host/viscosity.h"

#include <cstring>
#include <mutex>

#include "internal.h"

namespace fluidity {
namespace {

static_assert(sizeof(MacViscosityCell3D) == sizeof(fl_viscosity_cell) &&
              offsetof(MacViscosityCell3D, dynamic_viscosity) == offsetof(fl_viscosity_cell, dynamic_viscosity));
static_assert(sizeof(MacStrainRow3D) == sizeof(fl_strain_row) &&
              offsetof(MacStrainRow3D, first_axis) == offsetof(fl_strain_row, first_axis) &&
              offsetof(MacStrainRow3D, second_axis) == offsetof(fl_strain_row, second_axis) &&
              offsetof(MacStrainRow3D, terms) == offsetof(fl_strain_row, terms) &&
              offsetof(MacStrainRow3D, term_count) == offsetof(fl_strain_row, term_count) &&
              offsetof(MacStrainRow3D, quadrature_weight) == offsetof(fl_strain_row, quadrature_weight) &&
              offsetof(MacStrainRow3D, prescribed_strain) == offsetof(fl_strain_row, prescribed_strain));
static_assert(sizeof(MacStrainTerm3D) == sizeof(fl_strain_term) && sizeof(FaceNormalAxis) == 1);

void Copy(const fl_measurement& m, HipOperatorMeasurement& out) {
  std::memcpy(out.architecture, m.architecture, sizeof(out.architecture));
  out.kernel_ms = m.kernel_ms;
  out.copy_ms = m.copy_ms;
  out.launches = m.launches;
}

int Assemble(HipOperatorRuntime3D& rt, SparseMacGridView3D grid, std::span<const MacViscosityCell3D> cells,
             std::vector<MacStrainRow3D>& rows, fl_measurement* m) {
  const fl_geometry g = internal::ToC(grid.geometry());
  const auto blocks = grid.active_block_indices();
  rows.resize(6 * cells.size());
  std::size_t count = rows.size();
  const int code = fl_assemble_full_cell_strain(rt.handle(), &g, internal::ToC(blocks), blocks.size(),
                                                reinterpret_cast<const fl_viscosity_cell*>(cells.data()),
                                                cells.size(), reinterpret_cast<fl_strain_row*>(rows.data()), &count,
                                                m);
  if (code != FL_OK) {
    rows.clear();
    return code;
  }
  rows.resize(count);
  return FL_OK;
}

int Solve(HipOperatorRuntime3D& rt, const MacViscousVelocityBlockInputs3D& in, std::span<const double> rhs,
          MacViscousVelocityBlockOptions3D options, MacViscousVelocityBlockSolution3D& out, fl_measurement* m) {
  if (rhs.size() != in.face_inertias.size()) return FL_INVALID_ARGUMENT;
  std::vector<double> v(in.face_inertias.size());
  std::uint64_t iterations = 0;
  double max_residual = 0.0;
  const int code = fl_solve_viscous_block(rt.handle(), reinterpret_cast<const fl_strain_row*>(in.strain_rows.data()),
                                          in.strain_rows.size(), in.face_inertias.data(), in.face_inertias.size(),
                                          in.step_duration, rhs.data(), options.max_iterations,
                                          options.relative_tolerance, v.data(), &iterations, &max_residual, m);
  if (code != FL_OK) return code;
  out.velocity = std::move(v);
  out.iterations = iterations;
  out.max_residual = max_residual;
  return FL_OK;
}

}  // namespace

StatusOr<std::vector<MacStrainRow3D>> AssembleFullCellMacStrain3D(SparseMacGridView3D grid,
                                                                  std::span<const MacViscosityCell3D> cells) {
  HipOperatorCapacities3D need;
  const auto blocks = grid.active_block_indices();
  need.blocks = std::max<std::size_t>(blocks.size(), 1);
  need.lookup_blocks = internal::LookupNeed(blocks);
  need.pressure_cells = std::max<std::size_t>(cells.size(), 1);
  std::unique_lock<std::mutex> lock;
  auto rt = SharedHipRuntime3D::Acquire(need, lock);
  if (!rt.ok()) return rt.status();
  std::vector<MacStrainRow3D> rows;
  const int code = Assemble(**rt, grid, cells, rows, nullptr);
  if (code != FL_OK) return FromCode(code);
  return rows;
}

Status HipAssembleFullCellMacStrain3D(HipOperatorRuntime3D& runtime, SparseMacGridView3D grid,
                                      std::span<const MacViscosityCell3D> cells, std::vector<MacStrainRow3D>& rows,
                                      HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  const int code = Assemble(runtime, grid, cells, rows, &m);
  if (code == FL_OK) Copy(m, measurement);
  return FromCode(code);
}

StatusOr<MacViscousVelocityBlockSolution3D> SolveMacViscousVelocityBlock3D(const MacViscousVelocityBlockInputs3D& inputs,
                                                                           std::span<const double> rhs,
                                                                           MacViscousVelocityBlockOptions3D options) {
  HipOperatorCapacities3D need;
  need.blocks = std::max<std::size_t>((inputs.face_inertias.size() + 191) / 192, 1);
  need.pressure_cells = std::max<std::size_t>((inputs.strain_rows.size() + 5) / 6, 1);
  std::unique_lock<std::mutex> lock;
  auto rt = SharedHipRuntime3D::Acquire(need, lock);
  if (!rt.ok()) return rt.status();
  MacViscousVelocityBlockSolution3D out;
  const int code = Solve(**rt, inputs, rhs, options, out, nullptr);
  if (code != FL_OK) return FromCode(code);
  return out;
}

StatusOr<MacViscousVelocityBlockSolution3D> HipSolveMacViscousVelocityBlock3D(
    HipOperatorRuntime3D& runtime, const MacViscousVelocityBlockInputs3D& inputs, std::span<const double> rhs,
    MacViscousVelocityBlockOptions3D options, HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  MacViscousVelocityBlockSolution3D out;
  const int code = Solve(runtime, inputs, rhs, options, out, &m);
  if (code != FL_OK) return FromCode(code);
  Copy(m, measurement);
  return out;
}

}  // namespace fluidity
