// Topology preparation over the C-ABI (fl_prepare_topology), with the reference's two-bank
// publication.
#include "fluidity/host/liquid_step_workspace_3d.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "internal.h"

namespace fluidity {
namespace {

static_assert(sizeof(LiquidBlockState3D) == sizeof(fl_block_state) &&
              offsetof(LiquidBlockState3D, generation) == offsetof(fl_block_state, generation) &&
              offsetof(LiquidBlockState3D, inactive_age) == offsetof(fl_block_state, inactive_age));
static_assert(sizeof(LiquidCellState3D) == sizeof(fl_cell_state) &&
              offsetof(LiquidCellState3D, occupancy) == offsetof(fl_cell_state, occupancy) &&
              offsetof(LiquidCellState3D, previous_pressure) == offsetof(fl_cell_state, previous_pressure));
static_assert(sizeof(LiquidFaceState3D) == sizeof(fl_face_state) &&
              offsetof(LiquidFaceState3D, face_index) + offsetof(FaceIndex3D, lattice_index) ==
                  offsetof(fl_face_state, lattice) &&
              offsetof(LiquidFaceState3D, face_slot) == offsetof(fl_face_state, face_slot) &&
              offsetof(LiquidFaceState3D, validity) == offsetof(fl_face_state, validity) &&
              offsetof(LiquidFaceState3D, side_tag) == offsetof(fl_face_state, side_tag) &&
              offsetof(LiquidFaceState3D, structurally_eligible) == offsetof(fl_face_state, structurally_eligible));
static_assert(sizeof(LiquidCellOccupancy3D) == 1 && sizeof(LiquidFaceValidity3D) == 1 && sizeof(bool) == 1);

// Runtime extent for one preparation: the block box of the positions' sort bins (plus the bin
// after) and of the cells' owners, with a block of margin.
HipOperatorCapacities3D Need(UniformMacGeometry3D g, std::span<const Vec3> positions,
                             std::span<const CellIndex3D> cells, const LiquidTopologyHistory3D* previous) {
  HipOperatorCapacities3D need;
  need.particles = std::max<std::size_t>(positions.size(), 1);
  need.pressure_cells = std::max<std::size_t>({cells.size(), previous ? previous->cells.size() : 0, 1});
  need.blocks = std::max<std::size_t>(previous ? previous->blocks.size() : 1, 1);
  long double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  bool any = false;
  const auto add = [&](long double x, long double y, long double z) {
    const long double v[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = any ? std::min(lo[a], v[a]) : v[a];
      hi[a] = any ? std::max(hi[a], v[a]) : v[a];
    }
    any = true;
  };
  for (const Vec3& p : positions) {
    const long double x = (p.x - g.world_origin.x) / g.cell_width, y = (p.y - g.world_origin.y) / g.cell_width,
                      z = (p.z - g.world_origin.z) / g.cell_width;
    if (std::isfinite(static_cast<double>(x)) && std::isfinite(static_cast<double>(y)) &&
        std::isfinite(static_cast<double>(z)))
      add(std::floor(x / 4), std::floor(y / 4), std::floor(z / 4));
  }
  for (const CellIndex3D& c : cells) add(std::floor(c.x / 4.0L), std::floor(c.y / 4.0L), std::floor(c.z / 4.0L));
  long double vol = 64;
  if (any) {
    vol = 1;
    for (int a = 0; a < 3; ++a) vol *= hi[a] - lo[a] + 4;
  }
  need.lookup_blocks = vol > 0x7fffffff ? std::size_t{0x7fffffff} : static_cast<std::size_t>(vol);
  need.blocks = std::max<std::size_t>(need.blocks, std::min<std::size_t>(need.lookup_blocks, 1 << 22));
  return need;
}

}  // namespace

bool LiquidBlockState3D::operator==(const LiquidBlockState3D& o) const {
  return block_index.x == o.block_index.x && block_index.y == o.block_index.y && block_index.z == o.block_index.z &&
         generation == o.generation && activation_age == o.activation_age && inactive_age == o.inactive_age;
}
bool LiquidCellState3D::operator==(const LiquidCellState3D& o) const {
  return cell_index.x == o.cell_index.x && cell_index.y == o.cell_index.y && cell_index.z == o.cell_index.z &&
         occupancy == o.occupancy && previous_pressure == o.previous_pressure;
}
bool LiquidFaceState3D::operator==(const LiquidFaceState3D& o) const {
  const auto& a = face_index.lattice_index;
  const auto& b = o.face_index.lattice_index;
  return face_index.axis == o.face_index.axis && a.x == b.x && a.y == b.y && a.z == b.z && face_slot == o.face_slot &&
         carrier_velocity == o.carrier_velocity && validity == o.validity && side_tag == o.side_tag &&
         structurally_eligible == o.structurally_eligible;
}

StatusOr<LiquidStepWorkspace3D> LiquidStepWorkspace3D::Create(LiquidStateCapacities3D capacities,
                                                              UniformMacGeometry3D geometry) {
  const auto& o = geometry.world_origin;
  if (!std::isfinite(o.x) || !std::isfinite(o.y) || !std::isfinite(o.z) || !std::isfinite(geometry.cell_width) ||
      !(geometry.cell_width > 0))
    return Status(StatusCode::kInvalidArgument);
  LiquidStepWorkspace3D w(capacities, geometry);
  // Every bank and the position staging at the capacities: preparations then resize within them.
  try {
    const std::size_t blocks = std::min(capacities.blocks, capacities.faces / 192);
    for (Bank& b : w.banks_) {
      b.particle_order.reserve(capacities.particles);
      b.particles.reserve(capacities.particles);
      b.block_keys.reserve(blocks);
      b.blocks.reserve(blocks);
      b.cells.reserve(capacities.cells);
      b.faces.reserve(192 * blocks);
    }
    w.positions_.reserve(capacities.particles);
  } catch (const std::bad_alloc&) {
    return Status(StatusCode::kResourceExhausted);
  }
  return w;
}

LiquidTopologyHistory3D TopologyHistoryOf(const LiquidState3D& s) {
  return {s.geometry, s.clock.epoch, s.blocks, s.cells, s.faces};
}

int LiquidStepWorkspace3D::PrepareRecords(HipOperatorRuntime3D& rt, std::span<const LiquidParticleState3D> particles,
                                          std::span<const CellIndex3D> cells, const LiquidState3D* previous,
                                          Bank& bank, void* measurement) {
  if (particles.size() > capacities_.particles) return FL_RESOURCE_EXHAUSTED;
  positions_.resize(particles.size());
  for (std::size_t i = 0; i < particles.size(); ++i) positions_[i] = particles[i].world_position;
  const LiquidTopologyHistory3D h = previous ? TopologyHistoryOf(*previous) : LiquidTopologyHistory3D{};
  const int code = Prepare(rt, positions_, cells, previous ? &h : nullptr, bank, measurement);
  if (code != FL_OK) return code;
  bank.particles.resize(particles.size());
  for (std::size_t i = 0; i < particles.size(); ++i) bank.particles[i] = particles[bank.particle_order[i].original_ordinal];
  return FL_OK;
}

Status LiquidStepWorkspace3D::PrepareTrialTopology(std::span<const LiquidParticleState3D> particles,
                                                   std::span<const CellIndex3D> pressure_cells,
                                                   const LiquidState3D* previous_state) {
  if (particles.size() > capacities_.particles) return Status(StatusCode::kResourceExhausted);
  positions_.resize(particles.size());
  for (std::size_t i = 0; i < particles.size(); ++i) positions_[i] = particles[i].world_position;
  const LiquidTopologyHistory3D h = previous_state ? TopologyHistoryOf(*previous_state) : LiquidTopologyHistory3D{};
  const int candidate = published_ == 0 ? 1 : 0;
  HipOperatorCapacities3D need = Need(geometry_, positions_, pressure_cells, previous_state ? &h : nullptr);
  std::unique_lock<std::mutex> lock;
  for (int attempt = 0;; ++attempt) {
    auto rt = SharedHipRuntime3D::Acquire(need, lock);
    if (!rt.ok()) return rt.status();
    const int code = PrepareRecords(**rt, particles, pressure_cells, previous_state, banks_[candidate], nullptr);
    if (code == FL_RESOURCE_EXHAUSTED && attempt < 6 && (*rt)->capacities().lookup_blocks < (std::size_t{1} << 30)) {
      need.lookup_blocks = (*rt)->capacities().lookup_blocks * 4;
      lock.unlock();
      continue;
    }
    if (code != FL_OK) return FromCode(code);
    published_ = candidate;
    return Status();
  }
}

Status LiquidStepWorkspace3D::HipPrepareTrialTopology(HipOperatorRuntime3D& runtime,
                                                      std::span<const LiquidParticleState3D> particles,
                                                      std::span<const CellIndex3D> pressure_cells,
                                                      const LiquidState3D* previous_state,
                                                      HipOperatorMeasurement& measurement) {
  const int candidate = published_ == 0 ? 1 : 0;
  fl_measurement m{};
  const int code = PrepareRecords(runtime, particles, pressure_cells, previous_state, banks_[candidate], &m);
  if (code != FL_OK) return FromCode(code);
  published_ = candidate;
  std::memcpy(measurement.architecture, m.architecture, sizeof(measurement.architecture));
  measurement.kernel_ms = m.kernel_ms;
  measurement.copy_ms = m.copy_ms;
  measurement.launches = m.launches;
  return Status();
}

Status LiquidStepWorkspace3D::ApplyTrialTopology(LiquidState3D& trial) const {
  if (published_ < 0) return Status(StatusCode::kFailedPrecondition);
  const Bank& b = banks_[published_];
  const auto& o = trial.geometry.world_origin;
  if (!(o.x == geometry_.world_origin.x && o.y == geometry_.world_origin.y && o.z == geometry_.world_origin.z &&
        trial.geometry.cell_width == geometry_.cell_width))
    return Status(StatusCode::kInvalidArgument);
  if (trial.particles.capacity() < b.particles.size() || trial.blocks.capacity() < b.blocks.size() ||
      trial.cells.capacity() < b.cells.size() || trial.faces.capacity() < b.faces.size())
    return Status(StatusCode::kResourceExhausted);
  trial.particles.assign(b.particles.begin(), b.particles.end());
  trial.blocks.assign(b.blocks.begin(), b.blocks.end());
  trial.cells.assign(b.cells.begin(), b.cells.end());
  trial.faces.assign(b.faces.begin(), b.faces.end());
  return Status();
}

int LiquidStepWorkspace3D::Prepare(HipOperatorRuntime3D& rt, std::span<const Vec3> positions,
                                   std::span<const CellIndex3D> cells, const LiquidTopologyHistory3D* previous,
                                   Bank& bank, void* measurement) {
  const fl_geometry g = internal::ToC(geometry_);
  const fl_topology_capacities caps{capacities_.particles, capacities_.blocks, capacities_.cells, capacities_.faces};
  fl_topology_history h{};
  if (previous) {
    h.geometry = internal::ToC(previous->geometry);
    h.epoch = previous->epoch;
    h.blocks = reinterpret_cast<const fl_block_state*>(previous->blocks.data());
    h.block_count = previous->blocks.size();
    h.cells = reinterpret_cast<const fl_cell_state*>(previous->cells.data());
    h.cell_count = previous->cells.size();
    h.faces = reinterpret_cast<const fl_face_state*>(previous->faces.data());
    h.face_count = previous->faces.size();
  }
  if (positions.size() > capacities_.particles || cells.size() > capacities_.cells) return FL_RESOURCE_EXHAUSTED;
  // Output extents: the capacities, bounded by the block box the inputs can activate.
  const std::size_t box = Need(geometry_, positions, cells, nullptr).lookup_blocks;
  const std::size_t block_cap = std::min({capacities_.blocks, capacities_.faces / 192, box});
  bank.particle_order.resize(positions.size());
  bank.blocks.resize(block_cap);
  bank.cells.resize(cells.size());
  bank.faces.resize(192 * block_cap);
  std::uint64_t nb = 0, nc = 0, nf = 0;
  const int code = fl_prepare_topology(
      rt.handle(), &g, &caps, reinterpret_cast<const fl_vec3*>(positions.data()), positions.size(),
      reinterpret_cast<const fl_int3*>(cells.data()), cells.size(), previous ? &h : nullptr,
      reinterpret_cast<fl_sort_entry*>(bank.particle_order.data()), reinterpret_cast<fl_block_state*>(bank.blocks.data()),
      &nb, reinterpret_cast<fl_cell_state*>(bank.cells.data()), &nc,
      reinterpret_cast<fl_face_state*>(bank.faces.data()), &nf, static_cast<fl_measurement*>(measurement));
  if (code != FL_OK) return code;
  bank.blocks.resize(nb);
  bank.cells.resize(nc);
  bank.faces.resize(nf);
  bank.block_keys.resize(nb);
  for (std::size_t i = 0; i < nb; ++i) bank.block_keys[i] = bank.blocks[i].block_index;
  return FL_OK;
}

Status LiquidStepWorkspace3D::PrepareTrialTopology(std::span<const Vec3> world_positions,
                                                   std::span<const CellIndex3D> pressure_cells,
                                                   const LiquidTopologyHistory3D* previous) {
  const int candidate = published_ == 0 ? 1 : 0;
  HipOperatorCapacities3D need = Need(geometry_, world_positions, pressure_cells, previous);
  std::unique_lock<std::mutex> lock;
  for (int attempt = 0;; ++attempt) {
    auto rt = SharedHipRuntime3D::Acquire(need, lock);
    if (!rt.ok()) return rt.status();
    const int code = Prepare(**rt, world_positions, pressure_cells, previous, banks_[candidate], nullptr);
    if (code == FL_RESOURCE_EXHAUSTED && attempt < 6 && (*rt)->capacities().lookup_blocks < (std::size_t{1} << 30)) {
      need.lookup_blocks = (*rt)->capacities().lookup_blocks * 4;
      lock.unlock();
      continue;
    }
    if (code != FL_OK) return FromCode(code);
    banks_[candidate].particles.clear();
    published_ = candidate;
    return Status();
  }
}

Status LiquidStepWorkspace3D::HipPrepareTrialTopology(HipOperatorRuntime3D& runtime,
                                                      std::span<const Vec3> world_positions,
                                                      std::span<const CellIndex3D> pressure_cells,
                                                      const LiquidTopologyHistory3D* previous,
                                                      HipOperatorMeasurement& measurement) {
  const int candidate = published_ == 0 ? 1 : 0;
  fl_measurement m{};
  const int code = Prepare(runtime, world_positions, pressure_cells, previous, banks_[candidate], &m);
  if (code != FL_OK) return FromCode(code);
  banks_[candidate].particles.clear();
  published_ = candidate;
  std::memcpy(measurement.architecture, m.architecture, sizeof(measurement.architecture));
  measurement.kernel_ms = m.kernel_ms;
  measurement.copy_ms = m.copy_ms;
  measurement.launches = m.launches;
  return Status();
}

StatusOr<LiquidPreparedTopologyView3D> LiquidStepWorkspace3D::prepared() const {
  if (published_ < 0) return Status(StatusCode::kFailedPrecondition);
  const Bank& b = banks_[published_];
  auto grid = SparseMacGridView3D::Create(geometry_, b.block_keys);
  if (!grid.ok()) return grid.status();
  return LiquidPreparedTopologyView3D{*grid, b.particle_order, b.particles, b.blocks, b.cells, b.faces};
}

}  // namespace fluidity
