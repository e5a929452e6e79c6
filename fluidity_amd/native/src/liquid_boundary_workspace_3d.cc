// Bounded boundary workspace: freeze, links, compatible deposit, normalization and boundary gather over
// the C-ABI into banks reserved at Create, published on success.
#include "fluidity/host/liquid_boundary_workspace_3d.h"

#include <cstring>
#include <mutex>
#include <new>

#include "internal.h"
#include "records.h"

namespace fluidity {
namespace {

int Other(int published) { return published == 0 ? 1 : 0; }

// Runs `call` on the process-wide runtime, growing its dense-table extent when the call reports that
// the grid's block box does not fit (as the owning operators do).
template <class Call>
int OnSharedRuntime(HipOperatorCapacities3D need, Call&& call) {
  std::unique_lock<std::mutex> lock;
  for (int attempt = 0;; ++attempt) {
    auto rt = SharedHipRuntime3D::Acquire(need, lock);
    if (!rt.ok()) return static_cast<int>(rt.status().code());
    const int code = call(**rt);
    if (code == FL_RESOURCE_EXHAUSTED && attempt < 12 && (*rt)->capacities().lookup_blocks < (std::size_t{1} << 30)) {
      need.lookup_blocks = (*rt)->capacities().lookup_blocks * 4;
      lock.unlock();
      continue;
    }
    return code;
  }
}

HipOperatorCapacities3D GridNeed(SparseMacGridView3D grid, std::size_t particles, std::size_t cells,
                                 std::size_t faces) {
  HipOperatorCapacities3D need;
  const auto blocks = grid.active_block_indices();
  need.particles = std::max<std::size_t>(particles, 1);
  need.blocks = std::max<std::size_t>(blocks.size(), 1);
  need.lookup_blocks = internal::LookupNeed(blocks);
  need.pressure_cells = std::max<std::size_t>(cells, 1);
  need.pressure_faces = std::max<std::size_t>(faces, 1);
  return need;
}

}  // namespace

StatusOr<LiquidBoundaryWorkspace3D> LiquidBoundaryWorkspace3D::Create(LiquidStateCapacities3D c) {
  LiquidBoundaryWorkspace3D w;
  w.capacities_ = c;
  try {
    const std::size_t stencils = std::min(c.particles, kLiquidStencilParticleCapacity3D);
    for (int k = 0; k < 2; ++k) {
      w.geometry_[k].centers.reserve(27 * c.cells);
      w.geometry_[k].cells.reserve(c.cells);
      w.geometry_[k].faces.reserve(c.faces);
      w.links_[k].cells.reserve(c.cells);
      w.links_[k].faces.reserve(c.faces);
      w.deposits_[k].quanta.reserve(c.faces);
      w.deposits_[k].stencils.reserve(stencils);
      w.normalized_[k].masses.reserve(c.faces);
      w.normalized_[k].velocities.reserve(c.faces);
      w.normalized_[k].validity.reserve(c.faces);
      w.normalized_[k].predicted.reserve(c.faces);
      w.gathered_[k].reserve(c.particles);
    }
    Staging& s = w.staging_;
    s.candidates.reserve(c.cells);
    s.centers.reserve(27 * c.cells);
    s.corners.reserve(8 * c.cells);
    s.frozen_cells.reserve(c.cells);
    s.frozen_faces.reserve(6 * c.cells);
    s.fields.reserve(c.cells);
    s.wetted.reserve(c.faces);
    s.air.reserve(c.cells);
    s.interfaces.reserve(c.faces);
    s.link_cells.reserve(c.cells);
    s.link_faces.reserve(c.faces);
  } catch (const std::bad_alloc&) {
    return Status(StatusCode::kResourceExhausted);
  }
  return w;
}

// ---------------------------------------------------------------- freeze

StatusOr<FrozenLiquidBoundaryGeometryView3D> LiquidBoundaryWorkspace3D::HipFreeze(
    HipOperatorRuntime3D& runtime, const LiquidBoundaryGeometryViewInputs3D& in) {
  const auto& c = capacities_;
  if (in.cell_indices.size() > c.cells || in.center_distances.size() > 27 * c.cells ||
      in.corner_distances.size() > 8 * c.cells)
    return Status(StatusCode::kResourceExhausted);
  Staging& s = staging_;
  s.candidates.clear();
  s.centers.clear();
  s.corners.clear();
  for (const auto& x : in.cell_indices) s.candidates.push_back(internal::ToC3(x));
  for (const auto& x : in.center_distances) s.centers.push_back(internal::ToC(x));
  for (const auto& x : in.corner_distances) s.corners.push_back(internal::ToC(x));
  fl_freeze_inputs r{};
  const auto blocks = in.mac_grid.active_block_indices();
  r.blocks = internal::ToC(blocks);
  r.block_count = blocks.size();
  r.geometry = internal::ToC(in.mac_grid.geometry());
  r.cell_indices = s.candidates.data();
  r.cell_count = s.candidates.size();
  r.center_distances = s.centers.data();
  r.center_count = s.centers.size();
  r.corner_distances = s.corners.data();
  r.corner_count = s.corners.size();
  r.obstacle = internal::ToC(in.obstacle);
  r.step_duration = in.step_duration;
  r.liquid_density = in.liquid_density;
  r.surface_tension = in.surface_tension;
  r.has_exact_curvature = in.exact_outward_interface_curvature.has_value();
  r.exact_curvature = in.exact_outward_interface_curvature.value_or(0.0);
  r.max_cells = c.cells;
  r.max_faces = c.faces;
  internal::SetSampledSolid(in.sampled_solid, r);
  s.frozen_cells.resize(s.candidates.size());
  s.frozen_faces.resize(6 * s.candidates.size());
  std::uint64_t nc = 0, nf = 0;
  double bound = 0;
  fl_measurement m{};
  const int code = fl_freeze_geometry(runtime.handle(), &r, s.frozen_cells.data(), &nc, s.frozen_faces.data(), &nf,
                                      &bound, &m);
  if (code != FL_OK) return FromCode(code);
  if (nf > c.faces) return Status(StatusCode::kResourceExhausted);
  const int k = Other(geometry_bank_);
  GeometryBank& b = geometry_[k];
  b.centers.assign(in.center_distances.begin(), in.center_distances.end());
  b.cells.resize(nc);
  for (std::size_t i = 0; i < nc; ++i) b.cells[i] = internal::FromC(s.frozen_cells[i]);
  b.faces.resize(nf);
  for (std::size_t i = 0; i < nf; ++i) b.faces[i] = internal::FromC(s.frozen_faces[i]);
  capillary_bound_[k] = bound;
  geometry_bank_ = k;
  return FrozenLiquidBoundaryGeometryView3D{in.mac_grid,        b.centers,         b.cells,
                                            b.faces,            in.step_duration,  in.liquid_density,
                                            in.surface_tension, capillary_bound_[k]};
}

StatusOr<FrozenLiquidBoundaryGeometryView3D> LiquidBoundaryWorkspace3D::Freeze(
    const LiquidBoundaryGeometryViewInputs3D& inputs) {
  // Workspace capacity first, so that only the device's dense-table refusal grows the shared runtime.
  if (inputs.cell_indices.size() > capacities_.cells || inputs.center_distances.size() > 27 * capacities_.cells ||
      inputs.corner_distances.size() > 8 * capacities_.cells)
    return Status(StatusCode::kResourceExhausted);
  StatusOr<FrozenLiquidBoundaryGeometryView3D> out = Status(StatusCode::kInternal);
  const int code = OnSharedRuntime(GridNeed(inputs.mac_grid, 1, inputs.cell_indices.size(), capacities_.faces),
                                   [&](HipOperatorRuntime3D& rt) {
                                     out = HipFreeze(rt, inputs);
                                     return static_cast<int>(out.status().code());
                                   });
  if (code != FL_OK) return FromCode(code);
  return out;
}

// ---------------------------------------------------------------- links

StatusOr<LiquidPressureProblemView3D> LiquidBoundaryWorkspace3D::HipBuildLinks(
    HipOperatorRuntime3D& runtime, const FrozenLiquidBoundaryGeometryView3D& g, const LiquidPressureFields3D& f) {
  const auto& c = capacities_;
  if (g.cells.size() > c.cells || g.faces.size() > c.faces || g.center_distances.size() > 27 * c.cells ||
      f.cells.size() > c.cells || f.wetted_piston_face_fluxes.size() > c.faces || f.air_cell_pressures.size() > c.cells ||
      f.internal_interfaces.size() > c.faces)
    return Status(StatusCode::kResourceExhausted);
  Staging& s = staging_;
  s.centers.clear();
  s.frozen_cells.clear();
  s.frozen_faces.clear();
  s.fields.clear();
  s.wetted.clear();
  s.air.clear();
  s.interfaces.clear();
  for (const auto& x : g.center_distances) s.centers.push_back(internal::ToC(x));
  for (const auto& x : g.cells) s.frozen_cells.push_back(internal::ToC(x));
  for (const auto& x : g.faces) s.frozen_faces.push_back(internal::ToC(x));
  for (const auto& x : f.cells) s.fields.push_back(internal::ToC(x));
  for (const auto& x : f.wetted_piston_face_fluxes) s.wetted.push_back(internal::ToC(x));
  for (const auto& x : f.air_cell_pressures) s.air.push_back(internal::ToC(x));
  for (const auto& x : f.internal_interfaces) s.interfaces.push_back(internal::ToC(x));
  const auto blocks = g.mac_grid.active_block_indices();
  fl_frozen_view v{};
  v.blocks = internal::ToC(blocks);
  v.block_count = blocks.size();
  v.geometry = internal::ToC(g.mac_grid.geometry());
  v.center_distances = s.centers.data();
  v.center_count = s.centers.size();
  v.cells = s.frozen_cells.data();
  v.cell_count = s.frozen_cells.size();
  v.faces = s.frozen_faces.data();
  v.face_count = s.frozen_faces.size();
  v.step_duration = g.step_duration;
  v.liquid_density = g.liquid_density;
  v.surface_tension = g.surface_tension;
  v.capillary_step_bound = g.capillary_step_bound;
  fl_pressure_fields pf{};
  pf.cells = s.fields.data();
  pf.cell_count = s.fields.size();
  pf.predicted_face_velocities = f.predicted_face_velocities.data();
  pf.predicted_count = f.predicted_face_velocities.size();
  pf.ambient_pressure = f.ambient_pressure;
  pf.wetted_piston_face_fluxes = s.wetted.data();
  pf.wetted_count = s.wetted.size();
  pf.air_cell_pressures = s.air.data();
  pf.air_count = s.air.size();
  pf.internal_interfaces = s.interfaces.data();
  pf.interface_count = s.interfaces.size();
  s.link_cells.resize(g.cells.size());
  s.link_faces.resize(g.faces.size());
  std::uint64_t lower = 0;
  fl_measurement m{};
  const int code = fl_build_pressure_links(runtime.handle(), &v, &pf, s.link_cells.data(), s.link_faces.data(),
                                           &lower, &m);
  if (code != FL_OK) return FromCode(code);
  const int k = Other(link_bank_);
  LinkBank& b = links_[k];
  b.cells.resize(s.link_cells.size());
  for (std::size_t i = 0; i < b.cells.size(); ++i) b.cells[i] = internal::FromC(s.link_cells[i]);
  b.faces.resize(s.link_faces.size());
  for (std::size_t i = 0; i < b.faces.size(); ++i) b.faces[i] = internal::FromC(s.link_faces[i]);
  lower_theta_[k] = lower;
  link_bank_ = k;
  return LiquidPressureProblemView3D{b.cells, b.faces, lower, false};
}

StatusOr<LiquidPressureProblemView3D> LiquidBoundaryWorkspace3D::BuildLinks(const FrozenLiquidBoundaryGeometryView3D& g,
                                                                           const LiquidPressureFields3D& f) {
  const auto& c = capacities_;
  if (g.cells.size() > c.cells || g.faces.size() > c.faces || g.center_distances.size() > 27 * c.cells ||
      f.cells.size() > c.cells || f.wetted_piston_face_fluxes.size() > c.faces || f.air_cell_pressures.size() > c.cells ||
      f.internal_interfaces.size() > c.faces)
    return Status(StatusCode::kResourceExhausted);
  StatusOr<LiquidPressureProblemView3D> out = Status(StatusCode::kInternal);
  const int code = OnSharedRuntime(GridNeed(g.mac_grid, 1, g.cells.size(), g.faces.size()),
                                   [&](HipOperatorRuntime3D& rt) {
                                     out = HipBuildLinks(rt, g, f);
                                     return static_cast<int>(out.status().code());
                                   });
  if (code != FL_OK) return FromCode(code);
  return out;
}

// ---------------------------------------------------------------- compatible deposit

StatusOr<LiquidBoundaryFaceDepositsView3D> LiquidBoundaryWorkspace3D::HipDeposit(
    HipOperatorRuntime3D& runtime, const LiquidBoundaryDepositViewInputs3D& in) {
  const std::size_t nf = in.mac_grid.face_count(), n = in.particles.size();
  if (nf > capacities_.faces || n > capacities_.particles) return Status(StatusCode::kResourceExhausted);
  const int k = Other(deposit_bank_);
  DepositBank& b = deposits_[k];
  b.quanta.resize(nf);
  b.stencils.resize(n <= b.stencils.capacity() ? n : 0);
  const fl_geometry g = internal::ToC(in.mac_grid.geometry());
  const auto blocks = in.mac_grid.active_block_indices();
  const fl_quantization q{in.quantization.mass_per_quantum, in.quantization.pure_momentum_per_quantum,
                          in.quantization.declared_particle_capacity};
  fl_measurement m{};
  const int code = fl_p2g_compatible(
      runtime.handle(), &g, internal::ToC(blocks), blocks.size(),
      reinterpret_cast<const fl_boundary_particle*>(in.particles.data()), n,
      reinterpret_cast<const fl_face_tag*>(in.face_tags.data()), in.face_tags.size(), q,
      reinterpret_cast<fl_face_quanta*>(b.quanta.data()), nf,
      b.stencils.empty() ? nullptr : reinterpret_cast<fl_compatible_sample*>(b.stencils.data()), &m);
  if (code != FL_OK) return FromCode(code);
  deposit_bank_ = k;
  return LiquidBoundaryFaceDepositsView3D{b.quanta, b.stencils};
}

StatusOr<LiquidBoundaryFaceDepositsView3D> LiquidBoundaryWorkspace3D::Deposit(
    const LiquidBoundaryDepositViewInputs3D& inputs) {
  if (inputs.mac_grid.face_count() > capacities_.faces || inputs.particles.size() > capacities_.particles)
    return Status(StatusCode::kResourceExhausted);
  StatusOr<LiquidBoundaryFaceDepositsView3D> out = Status(StatusCode::kInternal);
  auto need = GridNeed(inputs.mac_grid, std::max<std::size_t>(inputs.particles.size(),
                                                               inputs.quantization.declared_particle_capacity),
                       1, 1);
  const int code = OnSharedRuntime(need, [&](HipOperatorRuntime3D& rt) {
    out = HipDeposit(rt, inputs);
    return static_cast<int>(out.status().code());
  });
  if (code != FL_OK) return FromCode(code);
  return out;
}

// ---------------------------------------------------------------- normalization

StatusOr<NormalizedLiquidMacFacesView3D> LiquidBoundaryWorkspace3D::HipNormalize(
    HipOperatorRuntime3D& runtime, std::span<const MacFaceDepositQuanta3D> quanta, MacDepositQuantization3D q,
    double minimum_face_mass) {
  const std::size_t nf = quanta.size();
  if (nf > capacities_.faces) return Status(StatusCode::kResourceExhausted);
  const int k = Other(normalization_bank_);
  NormalizationBank& b = normalized_[k];
  b.masses.resize(nf);
  b.velocities.resize(nf);
  b.validity.resize(nf);
  b.predicted.resize(nf);
  fl_measurement m{};
  const int code = fl_normalize_forces(
      runtime.handle(), reinterpret_cast<const fl_face_quanta*>(quanta.data()), nf,
      fl_quantization{q.mass_per_quantum, q.pure_momentum_per_quantum, q.declared_particle_capacity}, minimum_face_mass,
      fl_vec3{0.0, 0.0, 0.0}, 0.0, b.masses.data(), b.velocities.data(),
      reinterpret_cast<std::uint8_t*>(b.validity.data()), b.predicted.data(), &m);
  if (code != FL_OK) return FromCode(code);
  normalization_bank_ = k;
  return NormalizedLiquidMacFacesView3D{b.masses, b.velocities, b.validity};
}

StatusOr<NormalizedLiquidMacFacesView3D> LiquidBoundaryWorkspace3D::Normalize(
    std::span<const MacFaceDepositQuanta3D> quanta, MacDepositQuantization3D q, double minimum_face_mass) {
  if (quanta.size() > capacities_.faces) return Status(StatusCode::kResourceExhausted);
  HipOperatorCapacities3D need;
  need.blocks = std::max<std::size_t>(quanta.size() / 192, 1);
  StatusOr<NormalizedLiquidMacFacesView3D> out = Status(StatusCode::kInternal);
  const int code = OnSharedRuntime(need, [&](HipOperatorRuntime3D& rt) {
    out = HipNormalize(rt, quanta, q, minimum_face_mass);
    return static_cast<int>(out.status().code());
  });
  if (code != FL_OK) return FromCode(code);
  return out;
}

// ---------------------------------------------------------------- boundary gather

StatusOr<LiquidBoundaryGatherView3D> LiquidBoundaryWorkspace3D::HipGather(HipOperatorRuntime3D& runtime,
                                                                         const LiquidBoundaryGatherViewInputs3D& in) {
  const std::size_t n = in.particles.size();
  if (n > capacities_.particles) return Status(StatusCode::kResourceExhausted);
  const auto history = [](const LiquidBoundaryVelocityHistory3D& h, bool& mismatch) {
    const std::size_t nf = h.face_velocities.size();
    if (h.validity.size() != nf || h.face_tags.size() != nf || h.wall_samples.size() != nf) mismatch = true;
    return fl_velocity_history{h.velocity_sample_time,
                               h.wall_sample_time,
                               h.face_velocities.data(),
                               reinterpret_cast<const std::uint8_t*>(h.validity.data()),
                               reinterpret_cast<const fl_face_tag*>(h.face_tags.data()),
                               reinterpret_cast<const fl_wall_sample*>(h.wall_samples.data()),
                               nf};
  };
  bool mismatch = false;
  const fl_velocity_history oh = history(in.old_history, mismatch), nh = history(in.new_history, mismatch);
  if (mismatch) return Status(StatusCode::kInvalidArgument);
  const int k = Other(gather_bank_);
  auto& out = gathered_[k];
  out.resize(n);
  const fl_geometry g = internal::ToC(in.mac_grid.geometry());
  const auto blocks = in.mac_grid.active_block_indices();
  std::uint64_t counts[3] = {0, 0, 0};
  fl_measurement m{};
  const int code = fl_g2p_boundary(runtime.handle(), &g, internal::ToC(blocks), blocks.size(),
                                   reinterpret_cast<const fl_boundary_particle*>(in.particles.data()), n, in.old_time,
                                   in.new_time, &oh, &nh, reinterpret_cast<fl_boundary_gather*>(out.data()), counts,
                                   &m);
  if (code != FL_OK) return FromCode(code);
  fallbacks_[k] = {counts[0], counts[1], counts[2]};
  gather_bank_ = k;
  return LiquidBoundaryGatherView3D{out, counts[0], counts[1], counts[2]};
}

StatusOr<LiquidBoundaryGatherView3D> LiquidBoundaryWorkspace3D::Gather(const LiquidBoundaryGatherViewInputs3D& inputs) {
  if (inputs.particles.size() > capacities_.particles) return Status(StatusCode::kResourceExhausted);
  StatusOr<LiquidBoundaryGatherView3D> out = Status(StatusCode::kInternal);
  const int code = OnSharedRuntime(GridNeed(inputs.mac_grid, inputs.particles.size(), 1, 1),
                                   [&](HipOperatorRuntime3D& rt) {
                                     out = HipGather(rt, inputs);
                                     return static_cast<int>(out.status().code());
                                   });
  if (code != FL_OK) return FromCode(code);
  return out;
}

// ---------------------------------------------------------------- free functions

StatusOr<FrozenLiquidBoundaryGeometryView3D> FreezeLiquidBoundaryGeometry3D(
    const LiquidBoundaryGeometryViewInputs3D& inputs, LiquidBoundaryWorkspace3D& workspace) {
  return workspace.Freeze(inputs);
}
StatusOr<LiquidPressureProblemView3D> BuildLiquidPressureLinks3D(const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                                const LiquidPressureFields3D& fields,
                                                                LiquidBoundaryWorkspace3D& workspace) {
  return workspace.BuildLinks(geometry, fields);
}
StatusOr<LiquidBoundaryFaceDepositsView3D> DepositCompatibleMacQuanta3D(const LiquidBoundaryDepositViewInputs3D& inputs,
                                                                        LiquidBoundaryWorkspace3D& workspace) {
  return workspace.Deposit(inputs);
}
StatusOr<NormalizedLiquidMacFacesView3D> NormalizeLiquidMacFaces3D(std::span<const MacFaceDepositQuanta3D> quanta,
                                                                   MacDepositQuantization3D quantization,
                                                                   double minimum_face_mass,
                                                                   LiquidBoundaryWorkspace3D& workspace) {
  return workspace.Normalize(quanta, quantization, minimum_face_mass);
}
StatusOr<LiquidBoundaryGatherView3D> GatherLiquidBoundaryParticles3D(const LiquidBoundaryGatherViewInputs3D& inputs,
                                                                     LiquidBoundaryWorkspace3D& workspace) {
  return workspace.Gather(inputs);
}

}  // namespace fluidity
