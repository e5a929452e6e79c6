// Tests of the host C++ operator layer (fluidity/host/*.h), shaped like the reference's operator
// tests (native/tests/test_sparse_mac_grid_3d.cc, test_transfers_3d.cc, test_mac_pressure_3d.cc):
// layout queries, status codes, publication on failure, and parity of every device operator with
// the oracle (bit-exact quanta, sort order and forces; 1e-12 relative gathers; projection gates).
//
// Usage: host_operators_test cpu   (layout and status tests, no device)
//        host_operators_test gpu   (device operators against the oracle)
#include <array>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "fluidity/host/hip_liquid_substep_backend_3d.h"
#include "fluidity/host/liquid_boundary_workspace_3d.h"
#include "fluidity/host/mac_pressure_workspace_3d.h"
#include "fluidity/host/liquid_boundary_transfer_3d.h"
#include "fluidity/host/liquid_boundary_geometry_3d.h"
#include "fluidity/host/liquid_particle_update_3d.h"
#include "fluidity/host/liquid_step_workspace_3d.h"
#include "fluidity/host/mac_pressure_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/host/transfers_3d.h"
#include "fluidity/host/viscosity.h"
#include "oracle.h"

using namespace fluidity;

namespace {

int g_failures = 0;
#define EXPECT(cond)                                                                 \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      std::printf("  expectation failed at line %d: %s\n", __LINE__, #cond);         \
      ++g_failures;                                                                  \
      return;                                                                        \
    }                                                                                \
  } while (0)

const UniformMacGeometry3D kGeometry{{0.0, 0.0, 0.0}, 1.0 / 32.0};

std::vector<BlockIndex3D> DenseBlocks(int nx, int ny, int nz) {
  std::vector<BlockIndex3D> b;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) b.push_back({x, y, z});
  return b;
}

const flo_int3* O(std::span<const BlockIndex3D> b) { return reinterpret_cast<const flo_int3*>(b.data()); }
const flo_geometry* O(const UniformMacGeometry3D* g) { return reinterpret_cast<const flo_geometry*>(g); }

// Particles of a liquid box of cells [lo, hi): 8 per cell, one jittered point per octant, random velocity
// and affine slopes, trial = sorted position.
std::vector<MacTransferParticle3D> BoxParticles(int lo, int hi, std::uint64_t seed, double speed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 0.5);
  std::normal_distribution<double> n(0.0, 1.0);
  const double h = kGeometry.cell_width;
  std::vector<MacTransferParticle3D> p;
  for (int z = lo; z < hi; ++z)
    for (int y = lo; y < hi; ++y)
      for (int x = lo; x < hi; ++x)
        for (int o = 0; o < 8; ++o) {
          MacTransferParticle3D q{};
          q.sorted_world_position = {(x + 0.5 * (o & 1) + u(rng)) * h, (y + 0.5 * ((o >> 1) & 1) + u(rng)) * h,
                                     (z + 0.5 * (o >> 2) + u(rng)) * h};
          q.trial_world_position = q.sorted_world_position;
          q.particle_mass = 1000.0 * h * h * h / 8.0;
          q.particle_velocity = {speed * n(rng), speed * n(rng), speed * n(rng)};
          q.particle_velocity_slopes = {{0.1 * n(rng), 0.1 * n(rng), 0.1 * n(rng)},
                                        {0.1 * n(rng), 0.1 * n(rng), 0.1 * n(rng)},
                                        {0.1 * n(rng), 0.1 * n(rng), 0.1 * n(rng)}};
          p.push_back(q);
        }
  return p;
}

// ---------------------------------------------------------------- layout (no device)

void GridCreateSortsAndRejects() {
  std::vector<BlockIndex3D> keys = {{1, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 0, 0}};
  auto grid = SparseMacGrid3D::Create(kGeometry, keys, 8);
  EXPECT(grid.ok());
  const auto sorted = grid->active_block_indices();
  EXPECT(sorted.size() == 4 && sorted[0].x == 0 && sorted[1].x == 1 && sorted[2].y == 1 && sorted[3].z == 1);
  EXPECT(grid->face_count() == 4 * 192);
  keys.push_back({1, 0, 0});
  EXPECT(SparseMacGrid3D::Create(kGeometry, keys, 8).status().code() == StatusCode::kInvalidArgument);
  EXPECT(SparseMacGrid3D::Create(kGeometry, keys, 3).status().code() == StatusCode::kResourceExhausted);
  const std::vector<BlockIndex3D> far = {{0x7fffffff / 4 + 1, 0, 0}};
  EXPECT(SparseMacGrid3D::Create(kGeometry, far, 8).status().code() == StatusCode::kOutOfRange);
  for (double h : {0.0, -1.0, 1e-320, std::nan("")})
    EXPECT(SparseMacGrid3D::Create({{0, 0, 0}, h}, {}, 8).status().code() == StatusCode::kInvalidArgument);
  EXPECT(SparseMacGrid3D::Create({{std::nan(""), 0, 0}, 1.0}, {}, 8).status().code() == StatusCode::kInvalidArgument);
  const std::vector<BlockIndex3D> unsorted = {{1, 0, 0}, {0, 0, 0}};
  EXPECT(SparseMacGridView3D::Create(kGeometry, unsorted).status().code() == StatusCode::kInvalidArgument);
}

void FaceSlotsMatchOracle() {
  const std::vector<BlockIndex3D> keys = {{-1, 0, 0}, {0, 0, 0}, {0, -2, 1}, {3, 1, 1}};
  auto grid = SparseMacGrid3D::Create(kGeometry, keys, 8);
  EXPECT(grid.ok());
  const auto blocks = grid->active_block_indices();
  for (std::size_t slot = 0; slot < grid->face_count(); ++slot) {
    auto face = grid->FaceIndex(slot);
    EXPECT(face.ok());
    flo_face_index of{};
    EXPECT(flo_slot_face(O(blocks), blocks.size(), slot, &of) == 0);
    EXPECT(of.axis == static_cast<std::uint8_t>(face->axis) && of.lattice.x == face->lattice_index.x &&
           of.lattice.y == face->lattice_index.y && of.lattice.z == face->lattice_index.z);
    auto back = grid->FaceSlot(*face);
    EXPECT(back.ok() && *back == slot);
    auto anchor = grid->FaceWorldAnchor(*face);
    flo_vec3 oa{};
    EXPECT(anchor.ok() && flo_face_world_anchor(O(&kGeometry), of, &oa) == 0);
    EXPECT(anchor->x == oa.x && anchor->y == oa.y && anchor->z == oa.z);
  }
  EXPECT(grid->FaceIndex(grid->face_count()).status().code() == StatusCode::kOutOfRange);
  EXPECT(grid->FaceSlot({FaceNormalAxis::kX, {100, 0, 0}}).status().code() == StatusCode::kFailedPrecondition);
  EXPECT(grid->FaceSlot({static_cast<FaceNormalAxis>(3), {0, 0, 0}}).status().code() ==
         StatusCode::kInvalidArgument);
  const BlockIndex3D o = MacOwnerBlock3D({-1, -4, -5});
  EXPECT(o.x == -1 && o.y == -1 && o.z == -2);
  auto center = grid->CellWorldCenter({1, 2, 3});
  EXPECT(center.ok() && center->x == 1.5 / 32 && center->y == 2.5 / 32 && center->z == 3.5 / 32);
}

void OperatorsWithoutDeviceFailPrecondition() {
  auto rt = HipOperatorRuntime3D::Create({}, 0);
  EXPECT(rt.status().code() == StatusCode::kFailedPrecondition);
  auto grid = SparseMacGrid3D::Create(kGeometry, DenseBlocks(2, 2, 2), 8);
  EXPECT(grid.ok());
  const auto particles = BoxParticles(1, 3, 1, 1.0);
  std::vector<MacFaceDepositQuanta3D> out(grid->face_count(), MacFaceDepositQuanta3D{7, 7, 7});
  const Status s = DepositMacQuanta3D(*grid, particles, {0x1p-40, 0x1p-40, particles.size()}, out);
  EXPECT(s.code() == StatusCode::kFailedPrecondition);
  EXPECT(out[0].mass_quanta_sum == 7 && out.back().stress_force_quanta_sum == 7);
}

// ---------------------------------------------------------------- device operators

void DepositMatchesOracleBitExact() {
  auto grid = SparseMacGrid3D::Create(kGeometry, DenseBlocks(4, 4, 4), 64);
  EXPECT(grid.ok());
  const auto particles = BoxParticles(1, 14, 11, 2.0);
  const MacDepositQuantization3D q{0x1p-40, 0x1p-40, particles.size()};
  std::vector<MacFaceDepositQuanta3D> got(grid->face_count()), want(grid->face_count());
  EXPECT(DepositMacQuanta3D(*grid, particles, q, got).ok());
  const auto blocks = grid->active_block_indices();
  EXPECT(flo_deposit_quanta(O(&kGeometry), O(blocks), blocks.size(),
                            reinterpret_cast<const flo_particle*>(particles.data()), particles.size(),
                            {q.mass_per_quantum, q.pure_momentum_per_quantum, q.declared_particle_capacity},
                            std::uint64_t{1} << 40, reinterpret_cast<flo_face_quanta*>(want.data()), want.size(),
                            4) == 0);
  EXPECT(std::memcmp(got.data(), want.data(), got.size() * sizeof(got[0])) == 0);
  // Small declared capacity: the general rounding path (limit above 2^50).
  const std::vector<MacTransferParticle3D> few(particles.begin(), particles.begin() + 64);
  const MacDepositQuantization3D qs{0x1p-40, 0x1p-40, 64};
  EXPECT(DepositMacQuanta3D(*grid, few, qs, got).ok());
  EXPECT(flo_deposit_quanta(O(&kGeometry), O(blocks), blocks.size(), reinterpret_cast<const flo_particle*>(few.data()),
                            few.size(), {qs.mass_per_quantum, qs.pure_momentum_per_quantum, 64}, std::uint64_t{1} << 40,
                            reinterpret_cast<flo_face_quanta*>(want.data()), want.size(), 1) == 0);
  EXPECT(std::memcmp(got.data(), want.data(), got.size() * sizeof(got[0])) == 0);
}

// DepositCompatibleMacQuanta3D: two sides split at x = 0.5 of the box with a band of ineligible slots,
// half the particles with affine off; quanta and stencils against the oracle, field by field.
void CompatibleDepositMatchesOracle() {
  auto grid = SparseMacGrid3D::Create(kGeometry, DenseBlocks(4, 4, 4), 64);
  EXPECT(grid.ok());
  const auto transfer = BoxParticles(1, 14, 41, 2.0);
  std::vector<LiquidBoundaryParticle3D> particles(transfer.size());
  for (std::size_t i = 0; i < transfer.size(); ++i)
    particles[i] = {transfer[i], transfer[i].sorted_world_position.x < 0.25 ? 1 : 2, (i % 2) == 0};
  std::vector<LiquidBoundaryFaceTag3D> tags(grid->face_count());
  for (std::size_t f = 0; f < tags.size(); ++f) {
    const auto face = grid->FaceIndex(f);
    EXPECT(face.ok());
    const int x = face->lattice_index.x;
    tags[f] = {x < 8 ? 1 : 2, !(x == 8 && face->axis == FaceNormalAxis::kX)};
  }
  const MacDepositQuantization3D q{0x1p-40, 0x1p-40, particles.size()};
  auto got = DepositCompatibleMacQuanta3D({*grid, particles, tags, q});
  EXPECT(got.ok());
  std::vector<MacFaceDepositQuanta3D> want(grid->face_count());
  std::vector<flo_compatible_sample> want_st(81 * particles.size());
  const auto blocks = grid->active_block_indices();
  EXPECT(flo_deposit_compatible(O(&kGeometry), O(blocks), blocks.size(),
                                reinterpret_cast<const flo_boundary_particle*>(particles.data()), particles.size(),
                                reinterpret_cast<const flo_face_tag*>(tags.data()), tags.size(),
                                {q.mass_per_quantum, q.pure_momentum_per_quantum, q.declared_particle_capacity},
                                std::uint64_t{1} << 40, reinterpret_cast<flo_face_quanta*>(want.data()), want.size(),
                                want_st.data()) == 0);
  EXPECT(std::memcmp(got->face_quanta.data(), want.data(), want.size() * sizeof(want[0])) == 0);
  for (std::size_t i = 0; i < particles.size(); ++i)
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 27; ++k) {
        const auto& g = got->compatible_stencils[i][a][k];
        const auto& w = want_st[(i * 3 + a) * 27 + k];
        EXPECT(g.face_slot == w.face_slot && g.compatible == (w.compatible != 0));
        EXPECT(std::memcmp(&g.normalized_weight, &w.normalized_weight, 7 * sizeof(double)) == 0);
      }
  // No positive support for one particle: FailedPrecondition, nothing published through the twin.
  particles[5].side_tag = 77;
  std::vector<MacFaceDepositQuanta3D> out(grid->face_count(), MacFaceDepositQuanta3D{1, 2, 3});
  HipOperatorMeasurement m{};
  auto rt = HipOperatorRuntime3D::Create({particles.size(), 64, 1024, 16, 16});
  EXPECT(rt.ok());
  EXPECT(HipDepositCompatibleMacQuanta3D(**rt, {*grid, particles, tags, q}, out, {}, m).code() ==
         StatusCode::kFailedPrecondition);
  EXPECT(out[0].mass_quanta_sum == 1 && out.back().stress_force_quanta_sum == 3);
}

// GatherLiquidBoundaryParticles3D on the reference fixture (one block, unit cells, particle at 1.5) with one
// missing face per axis (continued from same-side donors) and an oblique wall at one of them.
void BoundaryGatherMatchesOracle() {
  const UniformMacGeometry3D unit{{0.0, 0.0, 0.0}, 1.0};
  const std::vector<BlockIndex3D> one = {{0, 0, 0}};
  auto grid = SparseMacGrid3D::Create(unit, one, 1);
  EXPECT(grid.ok());
  const std::size_t nf = grid->face_count();
  LiquidBoundaryParticle3D particle{};
  particle.transfer.sorted_world_position = {1.5, 1.5, 1.5};
  particle.transfer.trial_world_position = {1.5, 1.5, 1.5};
  particle.transfer.particle_mass = 16.0;
  particle.transfer.particle_velocity = {3.0, 4.0, 5.0};
  particle.side_tag = 7;
  std::vector<LiquidBoundaryFaceTag3D> tags(nf, {7, true});
  std::vector<double> old_u(nf), new_u(nf);
  std::vector<LiquidFaceValidity3D> validity(nf, LiquidFaceValidity3D::kValid);
  std::vector<LiquidPrescribedWallSample3D> old_w(nf), new_w(nf);
  for (std::size_t f = 0; f < nf; ++f) {
    const auto face = grid->FaceIndex(f);
    EXPECT(face.ok());
    const auto a = face->lattice_index;
    old_u[f] = 1.0 + a.x + 2.0 * a.y - a.z;
    new_u[f] = 4.0 + 2.0 * a.x - a.y + 3.0 * a.z;
  }
  for (auto axis : {FaceNormalAxis::kX, FaceNormalAxis::kY, FaceNormalAxis::kZ}) {
    const auto slot = grid->FaceSlot({axis, {1, 1, 1}});
    EXPECT(slot.ok());
    validity[*slot] = LiquidFaceValidity3D::kMissing;
  }
  const auto target = *grid->FaceSlot({FaceNormalAxis::kX, {1, 1, 1}});
  old_w[target] = {true, 7, {std::sqrt(0.5), std::sqrt(0.5), 0.0}, {2.0, -1.0, 4.0}};
  new_w[target] = {true, 7, {std::sqrt(0.5), std::sqrt(0.5), 0.0}, {5.0, 1.0, -2.0}};
  const LiquidBoundaryGatherInputs3D in{*grid, {&particle, 1}, 0.0, 1.0, {0.0, 0.0, old_u, validity, tags, old_w},
                                        {1.0, 1.0, new_u, validity, tags, new_w}};
  auto got = GatherLiquidBoundaryParticles3D(in);
  EXPECT(got.ok());
  const auto blocks = grid->active_block_indices();
  flo_history oh{0.0, 0.0, old_u.data(), reinterpret_cast<const std::uint8_t*>(validity.data()),
                 reinterpret_cast<const flo_face_tag*>(tags.data()), reinterpret_cast<const flo_wall_sample*>(old_w.data()),
                 nf};
  flo_history nh{1.0, 1.0, new_u.data(), reinterpret_cast<const std::uint8_t*>(validity.data()),
                 reinterpret_cast<const flo_face_tag*>(tags.data()), reinterpret_cast<const flo_wall_sample*>(new_w.data()),
                 nf};
  flo_boundary_gather want{};
  std::uint64_t counts[3];
  EXPECT(flo_gather_boundary(reinterpret_cast<const flo_geometry*>(&unit), O(blocks), blocks.size(),
                             reinterpret_cast<const flo_boundary_particle*>(&particle), 1, 0.0, 1.0, &oh, &nh, &want,
                             counts) == 0);
  const auto& g = got->particles[0];
  const double* a = &g.new_pic_velocity.x;
  const double* b = &want.new_pic_velocity.x;
  for (int i = 0; i < 27; ++i) EXPECT(std::fabs(a[i] - b[i]) <= 2e-12 + 1e-12 * std::fabs(b[i]));
  EXPECT(g.affine_enabled == (want.affine_enabled != 0) && g.old_pic_fallback == (want.old_pic_fallback != 0));
  EXPECT(got->old_pic_fallback_particle_count == counts[0] && got->low_mass_fallback_particle_count == counts[2]);
}

void Fp64DepositMatchesOracle() {
  auto grid = SparseMacGrid3D::Create(kGeometry, DenseBlocks(4, 4, 4), 64);
  EXPECT(grid.ok());
  const auto particles = BoxParticles(1, 14, 31, 2.0);
  std::vector<MacFaceDeposit3D> got(grid->face_count()), want(grid->face_count());
  EXPECT(DepositMac3D(*grid, particles, got).ok());
  const auto blocks = grid->active_block_indices();
  EXPECT(flo_deposit(O(&kGeometry), O(blocks), blocks.size(), reinterpret_cast<const flo_particle*>(particles.data()),
                     particles.size(), reinterpret_cast<flo_face_deposit*>(want.data()), want.size()) == 0);
  double mscale = 0, qscale = 0;
  for (const auto& w : want) {
    mscale = std::max(mscale, std::fabs(w.deposited_mass));
    qscale = std::max(qscale, std::fabs(w.deposited_pure_momentum));
  }
  for (std::size_t i = 0; i < got.size(); ++i) {
    EXPECT(std::fabs(got[i].deposited_mass - want[i].deposited_mass) <= 1e-12 * mscale);
    EXPECT(std::fabs(got[i].deposited_pure_momentum - want[i].deposited_pure_momentum) <= 1e-12 * qscale);
  }
}

void DepositFailuresPreserveOutput() {
  auto grid = SparseMacGrid3D::Create(kGeometry, DenseBlocks(4, 4, 4), 64);
  EXPECT(grid.ok());
  auto particles = BoxParticles(2, 6, 12, 1.0);
  std::vector<MacFaceDepositQuanta3D> out(grid->face_count(), MacFaceDepositQuanta3D{5, 6, 7});
  const MacDepositQuantization3D q{0x1p-40, 0x1p-40, particles.size()};
  auto moved = particles;
  moved[17].trial_world_position.x += 0.75 * kGeometry.cell_width;  // more than half a cell
  EXPECT(DepositMacQuanta3D(*grid, moved, q, out).code() == StatusCode::kFailedPrecondition);
  EXPECT(out[0].mass_quanta_sum == 5 && out[100].pure_momentum_quanta_sum == 6);
  auto heavy = particles;
  heavy[3].particle_mass = 1e30;  // scaled mass beyond 2^(62 - bit_width(cap - 1))
  EXPECT(DepositMacQuanta3D(*grid, heavy, q, out).code() == StatusCode::kOutOfRange);
  EXPECT(out[0].mass_quanta_sum == 5);
  auto negative = particles;
  negative[0].particle_mass = -1.0;
  EXPECT(DepositMacQuanta3D(*grid, negative, q, out).code() == StatusCode::kInvalidArgument);
  EXPECT(DepositMacQuanta3D(*grid, particles, {0x1p-40, 0x1p-40, particles.size() - 1}, out).code() ==
         StatusCode::kResourceExhausted);
  EXPECT(DepositMacQuanta3D(*grid, particles, {3e-12, 0x1p-40, particles.size()}, out).code() ==
         StatusCode::kInvalidArgument);
  EXPECT(out[0].mass_quanta_sum == 5 && out.back().stress_force_quanta_sum == 7);
}

void GatherMatchesOracle() {
  auto grid = SparseMacGrid3D::Create(kGeometry, DenseBlocks(4, 4, 4), 64);
  EXPECT(grid.ok());
  const auto particles = BoxParticles(1, 14, 13, 1.0);
  std::mt19937_64 rng(5);
  std::normal_distribution<double> n(0.0, 1.0);
  std::vector<double> u_old(grid->face_count()), u_new(grid->face_count());
  for (auto& v : u_old) v = n(rng);
  for (auto& v : u_new) v = n(rng);
  std::vector<MacParticleGatherResult3D> got(particles.size()), want(particles.size());
  EXPECT(GatherMac3D(*grid, particles, u_old, u_new, got).ok());
  const auto blocks = grid->active_block_indices();
  EXPECT(flo_gather(O(&kGeometry), O(blocks), blocks.size(), reinterpret_cast<const flo_particle*>(particles.data()),
                    particles.size(), u_old.data(), u_new.data(), u_old.size(),
                    reinterpret_cast<flo_gather_result*>(want.data()), 4) == 0);
  const double* a = reinterpret_cast<const double*>(got.data());
  const double* b = reinterpret_cast<const double*>(want.data());
  double scale = 0;
  for (std::size_t i = 0; i < 24 * got.size(); ++i) scale = std::max(scale, std::fabs(b[i]));
  for (std::size_t i = 0; i < 24 * got.size(); ++i) EXPECT(std::fabs(a[i] - b[i]) <= 1e-12 * std::fabs(b[i]) + 1e-12 * scale);
  std::vector<double> bad = u_new;
  bad[3] = std::nan("");
  std::vector<MacParticleGatherResult3D> keep(particles.size());
  keep[0].new_pic_velocity.x = 42.0;
  EXPECT(GatherMac3D(*grid, particles, u_old, bad, keep).code() == StatusCode::kInvalidArgument);
  EXPECT(keep[0].new_pic_velocity.x == 42.0);
}

// One stage sequence of a liquid substep through the table, each stage against the oracle on the same
// inputs: topology (sorted particles, activation, unique cells), compatible deposit, normalization,
// forces, frozen geometry of a box level set, typed links, projection, boundary gather and the particle
// update; then the publication hooks and a rejected stage that keeps the published view.
void SubstepTableDrivesStagesAgainstOracle() {
  const int lo = 4, hi = 12;  // liquid cells [4, 12)^3 of the h = 1/32 lattice
  const double h = kGeometry.cell_width, dt = 2e-4;
  const auto transfer = BoxParticles(lo, hi, 23, 0.5);
  const std::size_t n = transfer.size();
  LiquidState3D committed;
  committed.geometry = kGeometry;
  committed.particles.resize(n);
  for (std::size_t i = 0; i < n; ++i) {
    auto& p = committed.particles[i];
    p.id = i + 1;
    p.world_position = transfer[i].sorted_world_position;
    p.world_velocity = transfer[i].particle_velocity;
    p.velocity_slopes = transfer[i].particle_velocity_slopes;
    p.mass = transfer[i].particle_mass;
    p.volume_ratio = 1.0;
    p.pic_initialized = true;
    p.constituent_count = 1;
    p.constituents[0].volume_fraction = 1.0;
  }
  std::vector<CellIndex3D> candidates;
  for (int z = lo; z < hi; ++z)
    for (int y = lo; y < hi; ++y)
      for (int x = lo; x < hi; ++x) candidates.push_back({x, y, z});
  const LiquidStateCapacities3D caps{8192, 512, 4096, 192 * 512, 0};
  auto context = HipLiquidSubstepContext3D::Create(caps);
  EXPECT(context.ok());
  (*context)->projection_settings.relative_divergence_target = 1e-9;
  (*context)->projection_settings.levels = 3;
  const HipLiquidSubstepBackend3D table = MakeHipLiquidSubstepBackend3D(**context);
  EXPECT(table.sample_field == nullptr && table.project_viscous != nullptr);
  auto step_ws = LiquidStepWorkspace3D::Create(caps, kGeometry);
  auto boundary_ws = LiquidBoundaryWorkspace3D::Create(caps);
  auto pressure_ws = MacPressureWorkspace3D::Create(caps.cells, 6 * caps.cells);
  EXPECT(step_ws.ok() && boundary_ws.ok() && pressure_ws.ok());
  EXPECT(table.initialize(table.context, committed).ok());

  // topology
  EXPECT(table.topology(table.context, committed.particles, candidates, committed, *step_ws).ok());
  auto prepared = step_ws->prepared();
  EXPECT(prepared.ok() && prepared->particles.size() == n && prepared->cells.size() == candidates.size());
  std::vector<Vec3> positions(n);
  for (std::size_t i = 0; i < n; ++i) positions[i] = committed.particles[i].world_position;
  std::vector<LiquidParticleSortEntry3D> want_order(n);
  EXPECT(flo_sort_particles(O(&kGeometry), reinterpret_cast<const flo_vec3*>(positions.data()), n,
                            reinterpret_cast<flo_sort_entry*>(want_order.data())) == 0);
  for (std::size_t i = 0; i < n; ++i)
    EXPECT(prepared->particle_order[i].original_ordinal == want_order[i].original_ordinal &&
           prepared->particles[i].id == committed.particles[want_order[i].original_ordinal].id);
  const SparseMacGridView3D grid = prepared->grid;
  const auto blocks = grid.active_block_indices();
  const std::size_t nf = grid.face_count();

  // compatible deposit with every slot eligible on side 0
  std::vector<LiquidBoundaryParticle3D> bp(n);
  for (std::size_t i = 0; i < n; ++i) {
    const auto& p = prepared->particles[i];
    bp[i].transfer = {p.world_position, p.world_position, p.mass, p.world_velocity, p.velocity_slopes};
  }
  const std::vector<LiquidBoundaryFaceTag3D> tags(nf);
  const MacDepositQuantization3D q{0x1p-40, 0x1p-40, n};
  auto deposits = table.deposit(table.context, {grid, bp, tags, q}, *boundary_ws);
  EXPECT(deposits.ok() && deposits->face_quanta.size() == nf && deposits->compatible_stencils.size() == n);
  std::vector<MacFaceDepositQuanta3D> want_q(nf);
  EXPECT(flo_deposit_compatible(O(&kGeometry), O(blocks), blocks.size(),
                                reinterpret_cast<const flo_boundary_particle*>(bp.data()), n,
                                reinterpret_cast<const flo_face_tag*>(tags.data()), nf,
                                {q.mass_per_quantum, q.pure_momentum_per_quantum, q.declared_particle_capacity},
                                std::uint64_t{1} << 40, reinterpret_cast<flo_face_quanta*>(want_q.data()), nf,
                                nullptr) == 0);
  EXPECT(std::memcmp(deposits->face_quanta.data(), want_q.data(), nf * sizeof(MacFaceDepositQuanta3D)) == 0);
  // a rejected deposit (tags of the wrong extent) keeps the published quanta
  const std::vector<LiquidBoundaryFaceTag3D> short_tags(nf - 1);
  EXPECT(table.deposit(table.context, {grid, bp, short_tags, q}, *boundary_ws).status().code() ==
         StatusCode::kInvalidArgument);
  EXPECT(std::memcmp(deposits->face_quanta.data(), want_q.data(), nf * sizeof(MacFaceDepositQuanta3D)) == 0);

  // normalize and forces
  auto normalized = table.normalize(table.context, deposits->face_quanta, q, 1e-12, *boundary_ws);
  EXPECT(normalized.ok());
  std::vector<double> want_mass(nf), want_u0(nf), pred(nf), want_pred(nf);
  std::vector<std::uint8_t> want_validity(nf);
  EXPECT(flo_normalize(reinterpret_cast<const flo_face_quanta*>(want_q.data()), nf,
                       {q.mass_per_quantum, q.pure_momentum_per_quantum, q.declared_particle_capacity}, 1e-12,
                       want_mass.data(), want_u0.data(), want_validity.data()) == 0);
  EXPECT(std::memcmp(normalized->face_masses.data(), want_mass.data(), nf * 8) == 0 &&
         std::memcmp(normalized->face_velocities.data(), want_u0.data(), nf * 8) == 0 &&
         std::memcmp(normalized->validity.data(), want_validity.data(), nf) == 0);
  const Vec3 g{0.0, -9.81, 0.0};
  EXPECT(table.forces(table.context, grid, normalized->face_velocities, g, dt, pred).ok());
  EXPECT(flo_apply_forces(O(blocks), blocks.size(), want_u0.data(), nf, {g.x, g.y, g.z}, dt, want_pred.data()) == 0);
  EXPECT(std::memcmp(pred.data(), want_pred.data(), nf * 8) == 0);

  // frozen geometry of the box level set: centre samples of cells [lo - 1, hi + 1)^3, corners [lo, hi]^3
  const auto box = [&](double x, double y, double z) {
    const double c = 0.5 * (lo + hi) * h, half = 0.5 * (hi - lo) * h;
    return std::max({std::fabs(x - c), std::fabs(y - c), std::fabs(z - c)}) - half;
  };
  std::vector<LiquidCenterDistance3D> centers;
  std::vector<LiquidCornerDistance3D> corners;
  for (int z = lo - 1; z <= hi; ++z)
    for (int y = lo - 1; y <= hi; ++y)
      for (int x = lo - 1; x <= hi; ++x) centers.push_back({{x, y, z}, box((x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h)});
  for (int z = lo; z <= hi; ++z)
    for (int y = lo; y <= hi; ++y)
      for (int x = lo; x <= hi; ++x) corners.push_back({{x, y, z}, box(x * h, y * h, z * h)});
  const double rho = 1000.0;
  auto frozen = table.freeze(table.context, {grid, candidates, centers, corners, {}, dt, rho, 0.0}, *boundary_ws);
  EXPECT(frozen.ok() && frozen->cells.size() == candidates.size());
  {
    std::vector<flo_int3> fc;
    for (const auto& c : candidates) fc.push_back({c.x, c.y, c.z});
    std::vector<flo_center_distance> fce, fco;
    for (const auto& c : centers) fce.push_back({{c.cell_index.x, c.cell_index.y, c.cell_index.z}, 0, c.signed_distance});
    for (const auto& c : corners)
      fco.push_back({{c.corner_index.x, c.corner_index.y, c.corner_index.z}, 0, c.signed_distance});
    flo_freeze_inputs fi{};
    fi.blocks = O(blocks);
    fi.block_count = blocks.size();
    fi.geometry = *O(&kGeometry);
    fi.cell_indices = fc.data();
    fi.cell_count = fc.size();
    fi.center_distances = fce.data();
    fi.center_count = fce.size();
    fi.corner_distances = fco.data();
    fi.corner_count = fco.size();
    fi.step_duration = dt;
    fi.liquid_density = rho;
    std::vector<flo_frozen_cell> wc(fc.size());
    std::vector<flo_frozen_face> wf(6 * fc.size());
    std::uint64_t wnc = 0, wnf = 0;
    double bound = 0;
    EXPECT(flo_freeze_geometry(&fi, wc.data(), &wnc, wf.data(), &wnf, &bound) == 0);
    EXPECT(wnc == frozen->cells.size() && wnf == frozen->faces.size());
    for (std::size_t i = 0; i < wnc && i < frozen->cells.size(); ++i)
      EXPECT(frozen->cells[i].fluid_volume == wc[i].fluid_volume && frozen->cells[i].curvature == wc[i].curvature);
    for (std::size_t i = 0; i < wnf && i < frozen->faces.size(); ++i)
      EXPECT(frozen->faces[i].face_slot == wf[i].face_slot && frozen->faces[i].open_area == wf[i].open_area &&
             frozen->faces[i].blocked_area == wf[i].blocked_area);
  }

  // typed links, then the projection against the FP64 oracle
  const std::vector<LiquidPressureCellFields3D> fields(frozen->cells.size(),
                                                       {0.0, std::numeric_limits<double>::infinity(), 0.0});
  std::vector<double> face_pred(frozen->faces.size());
  for (std::size_t i = 0; i < face_pred.size(); ++i) face_pred[i] = pred[frozen->faces[i].face_slot];
  auto problem = table.links(table.context, *frozen, {fields, face_pred, 0.0}, *boundary_ws);
  EXPECT(problem.ok() && problem->faces.size() == frozen->faces.size());
  auto projection = table.project(table.context, grid, dt, *problem, {256, 1.0}, *pressure_ws);
  EXPECT(projection.ok());
  std::vector<flo_pressure_cell> oc;
  std::vector<flo_pressure_face> of;
  for (const auto& c : problem->cells)
    oc.push_back({{c.cell_index.x, c.cell_index.y, c.cell_index.z}, 0, c.fluid_volume, c.target_divergence,
                  c.bulk_modulus, c.previous_pressure, c.integrated_inward_wall_flux});
  for (const auto& f : problem->faces) {
    flo_pressure_face r{};
    r.axis = static_cast<std::int32_t>(f.face_index.axis);
    r.lattice = {f.face_index.lattice_index.x, f.face_index.lattice_index.y, f.face_index.lattice_index.z};
    r.open_area = f.open_area;
    r.inverse_density = f.inverse_density;
    r.pressure_gradient_distance = f.pressure_gradient_distance;
    r.predicted_normal_velocity = f.predicted_normal_velocity;
    r.wall_normal_velocity = f.wall_normal_velocity;
    if (const auto* in = std::get_if<InternalMacPressureLink3D>(&f.pressure_link)) {
      r.link_kind = 1;
      r.lift = in->pressure_jump;
    } else if (const auto* d = std::get_if<DirichletMacPressureLink3D>(&f.pressure_link)) {
      r.link_kind = 2;
      r.lift = d->exterior_pressure;
    }
    of.push_back(r);
  }
  std::vector<double> wp(oc.size()), wu(of.size()), wr(oc.size()), wres(oc.size()), wd(oc.size());
  std::vector<std::uint64_t> wg(oc.size());
  flo_pressure_diagnostics wdiag{};
  const flo_pressure_options po{256, 1.0, 1e-12, 3, 2, 16, 0, 0.5, 0, 0};
  EXPECT(flo_project(O(&kGeometry), O(blocks), blocks.size(), dt, oc.data(), oc.size(), of.data(), of.size(), &po,
                     wp.data(), wu.data(), wr.data(), wres.data(), wd.data(), wg.data(), &wdiag) == 0);
  double pmax = 0, perr = 0;
  for (std::size_t i = 0; i < wp.size(); ++i) {
    pmax = std::max(pmax, std::fabs(wp[i]));
    perr = std::max(perr, std::fabs(projection->cell_pressures[i] - wp[i]));
  }
  EXPECT(projection->diagnostics.iteration_count > 0 && perr <= 1e-6 * pmax);
  EXPECT(std::abs(projection->frozen_measurement.integrated_rhs_l2_m3_per_s - wdiag.rhs_l2) <=
         1e-12 * wdiag.rhs_l2);
  std::printf("  table projection: %zu cells, %zu iterations, pressure error %.2e of %.3e Pa\n", oc.size(),
              projection->diagnostics.iteration_count, perr, pmax);

  // The viscous projection member (0.5 Pa s) against the oracle's coupled solve (SolveCoupled with the
  // reference's Jacobi diagonal) on full-cell strain rows: the links' problem with every cell made full
  // (the box's high corner cells are cut, which the member refuses as Unimplemented, below).
  {
    const double mu = 0.5, full = h * h * h;
    const LiquidBoundaryGeometryViewInputs3D gin{grid, candidates, centers, corners, {}, dt, rho, 0.0};
    std::vector<MacPressureCell3D> full_cells(problem->cells.begin(), problem->cells.end());
    for (auto& c : full_cells) c.fluid_volume = full;
    const LiquidPressureProblemView3D uncut{full_cells, problem->faces};
    auto viscous = table.project_viscous(table.context, gin, uncut, mu, {}, {256, 1.0}, *pressure_ws);
    if (!viscous.ok()) std::printf("  viscous projection status %d\n", static_cast<int>(viscous.status().code()));
    EXPECT(viscous.ok());
    std::vector<flo_pressure_cell> voc(oc);
    for (auto& c : voc) c.fluid_volume = full;
    std::vector<flo_viscosity_cell> vc;
    for (const auto& c : problem->cells) vc.push_back({{c.cell_index.x, c.cell_index.y, c.cell_index.z}, 0, mu});
    std::vector<flo_strain_row> rows(6 * vc.size());
    std::size_t nr = 0;
    EXPECT(flo_assemble_full_cell_strain(O(&kGeometry), O(blocks), blocks.size(), vc.data(), vc.size(), rows.data(),
                                         rows.size(), &nr) == 0);
    std::vector<double> vp(voc.size()), vu(of.size()), vd(voc.size());
    flo_viscous_diagnostics vdiag{};
    EXPECT(flo_project_viscous(O(&kGeometry), O(blocks), blocks.size(), dt, voc.data(), voc.size(), of.data(),
                               of.size(), rows.data(), nr, 256, vp.data(), vu.data(), vd.data(), &vdiag) == 0);
    if (viscous.ok()) {
      double vmax = 0, verr = 0, umax = 0, uerr = 0;
      for (std::size_t i = 0; i < vp.size(); ++i) {
        vmax = std::max(vmax, std::fabs(vp[i]));
        verr = std::max(verr, std::fabs(viscous->cell_pressures[i] - vp[i]));
      }
      for (std::size_t i = 0; i < vu.size(); ++i) {
        umax = std::max(umax, std::fabs(vu[i]));
        uerr = std::max(uerr, std::fabs(viscous->face_velocities[i] - vu[i]));
      }
      EXPECT(verr <= 1e-3 * vmax && uerr <= 1e-3 * umax);
      EXPECT(viscous->max_cell_constraint_defect <= 2e-5);
      std::printf("  table viscous projection: %zu rows, %zu GCR iterations (oracle %llu), pressure error %.2e of %.3e Pa\n",
                  nr, viscous->pressure_iterations, static_cast<unsigned long long>(vdiag.pressure_iterations), verr,
                  vmax);
    }
    EXPECT(table.project_viscous(table.context, gin, *problem, mu, {}, {256, 1.0}, *pressure_ws).status().code() ==
           StatusCode::kUnimplemented);
  }

  // boundary gather: the old history is the normalized field, the new one the predicted field with the
  // corrected pressure-face velocities
  std::vector<double> u_new(pred.begin(), pred.end());
  for (std::size_t i = 0; i < frozen->faces.size(); ++i)
    u_new[frozen->faces[i].face_slot] = projection->corrected_face_velocities[i];
  const std::vector<LiquidPrescribedWallSample3D> walls(nf);
  const LiquidBoundaryVelocityHistory3D old_h{0.0, 0.0, normalized->face_velocities, normalized->validity, tags, walls};
  const LiquidBoundaryVelocityHistory3D new_h{dt, dt, u_new, normalized->validity, tags, walls};
  auto gathered = table.gather(table.context, {grid, bp, 0.0, dt, old_h, new_h}, *boundary_ws);
  EXPECT(gathered.ok() && gathered->particles.size() == n);
  std::vector<LiquidBoundaryParticleGather3D> want_g(n);
  {
    const flo_history oh{0.0, 0.0, want_u0.data(), want_validity.data(), reinterpret_cast<const flo_face_tag*>(tags.data()),
                         reinterpret_cast<const flo_wall_sample*>(walls.data()), nf};
    const flo_history nh{dt, dt, u_new.data(), want_validity.data(), reinterpret_cast<const flo_face_tag*>(tags.data()),
                         reinterpret_cast<const flo_wall_sample*>(walls.data()), nf};
    std::uint64_t counts[3] = {};
    EXPECT(flo_gather_boundary(O(&kGeometry), O(blocks), blocks.size(),
                               reinterpret_cast<const flo_boundary_particle*>(bp.data()), n, 0.0, dt, &oh, &nh,
                               reinterpret_cast<flo_boundary_gather*>(want_g.data()), counts) == 0);
    EXPECT(counts[0] == gathered->old_pic_fallback_particle_count && counts[2] == gathered->low_mass_fallback_particle_count);
  }
  double gscale = 0;
  for (const auto& w : want_g) gscale = std::max({gscale, std::fabs(w.new_pic_velocity.x), std::fabs(w.new_pic_velocity.y)});
  for (std::size_t i = 0; i < n; ++i) {
    const auto& a = gathered->particles[i];
    const auto& b = want_g[i];
    EXPECT(std::fabs(a.new_pic_velocity.y - b.new_pic_velocity.y) <= 1e-12 * gscale &&
           std::fabs(a.grid_velocity_increment.x - b.grid_velocity_increment.x) <= 1e-12 * gscale &&
           a.affine_enabled == b.affine_enabled && a.low_mass_fallback == b.low_mass_fallback);
  }

  // particle update of the sorted trial particles against the oracle, on the same gather records
  std::vector<double> distances(n);
  for (std::size_t i = 0; i < n; ++i) {
    const Vec3 x = prepared->particles[i].world_position;
    distances[i] = box(x.x, x.y, x.z);
  }
  std::vector<LiquidParticleState3D> trial(prepared->particles.begin(), prepared->particles.end());
  std::vector<LiquidParticleState3D> want_p = trial;
  const LiquidParticleUpdateInputs3D ui{gathered->particles, distances, {}, h, dt, 1.0, 0.0, 0.0};
  EXPECT(table.update_particles(table.context, ui, trial).ok());
  const flo_update_options uo{h, dt, 1.0, 0.0, 0.0, {}};
  EXPECT(flo_update_particles(&uo, reinterpret_cast<const flo_boundary_gather*>(gathered->particles.data()),
                              distances.data(), n, reinterpret_cast<flo_liquid_particle*>(want_p.data())) == 0);
  std::size_t moved = 0;
  for (std::size_t i = 0; i < n; ++i) {
    EXPECT(trial[i].world_position.x == want_p[i].world_position.x &&
           trial[i].world_position.y == want_p[i].world_position.y &&
           trial[i].world_velocity.y == want_p[i].world_velocity.y);
    moved += trial[i].world_position.y != prepared->particles[i].world_position.y;
  }
  EXPECT(moved > n / 2);

  // publication: the prepared topology into a reserved trial state, then the hooks
  LiquidState3D next;
  next.geometry = kGeometry;
  next.particles.reserve(caps.particles);
  next.blocks.reserve(caps.blocks);
  next.cells.reserve(caps.cells);
  next.faces.reserve(caps.faces);
  EXPECT(step_ws->ApplyTrialTopology(next).ok());
  EXPECT(next.particles.size() == n && next.blocks.size() == blocks.size() && next.faces.size() == nf);
  LiquidState3D small;
  small.geometry = kGeometry;
  EXPECT(step_ws->ApplyTrialTopology(small).code() == StatusCode::kResourceExhausted && small.particles.empty());
  LiquidClock3D clock;
  clock.epoch = 1;
  EXPECT(table.prepare_publication(table.context, next, clock).ok());
  table.commit_publication(table.context);
  table.discard_publication(table.context);
  // the second step's topology with the first as history: the surviving cells keep their records
  next.particles.assign(trial.begin(), trial.end());
  next.clock = clock;
  EXPECT(table.topology(table.context, next.particles, candidates, next, *step_ws).ok());
  auto second = step_ws->prepared();
  EXPECT(second.ok() && second->cells.size() == candidates.size());
}

// Liquid box [2, 10)^3 in a sealed 16^3 tank: liquid-liquid faces internal, liquid-air Dirichlet p = 0 at
// distance h, tank walls sealed. Faces are the union of incident faces in owned-slot order.
void ProjectionMeetsGatesAndMatchesOracle() {
  const int n = 16, lo = 2, hi = 10;
  const double h = 1.0 / n, dt = 1e-3, rho = 1000.0;
  const UniformMacGeometry3D geom{{0, 0, 0}, h};
  auto grid = SparseMacGrid3D::Create(geom, DenseBlocks(5, 5, 5), 256);
  EXPECT(grid.ok());
  auto liquid = [&](int x, int y, int z) { return x >= lo && x < hi && y >= lo && y < hi && z >= lo && z < hi; };
  std::vector<MacPressureCell3D> cells;
  for (int z = lo; z < hi; ++z)
    for (int y = lo; y < hi; ++y)
      for (int x = lo; x < hi; ++x) cells.push_back({{x, y, z}, h * h * h, 0.0, INFINITY, 0.0, 0.0});
  std::mt19937_64 rng(3);
  std::normal_distribution<double> nd(0.0, 0.5);
  std::vector<std::pair<std::size_t, MacPressureFace3D>> faces;
  for (const auto& c : cells)
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        FaceLatticeIndex3D l{c.cell_index.x, c.cell_index.y, c.cell_index.z};
        (a == 0 ? l.x : (a == 1 ? l.y : l.z)) += side;
        const FaceIndex3D f{static_cast<FaceNormalAxis>(a), l};
        auto slot = grid->FaceSlot(f);
        EXPECT(slot.ok());
        bool seen = false;
        for (const auto& e : faces) seen = seen || e.first == *slot;
        if (seen) continue;
        const int t[3] = {l.x - (a == 0), l.y - (a == 1), l.z - (a == 2)};
        const bool tail = liquid(t[0], t[1], t[2]), head = liquid(l.x, l.y, l.z);
        const int li = a == 0 ? l.x : (a == 1 ? l.y : l.z);
        MacPressureFace3D face{f, h * h, 1.0 / rho, h, nd(rng) - 9.81 * dt * (a == 1), 0.0, SealedMacPressureLink3D{}};
        if (li == 0 || li == n) {
          face.open_area = 0.0;
        } else if (tail && head) {
          face.pressure_link = InternalMacPressureLink3D{0.0};
        } else {
          face.pressure_link = DirichletMacPressureLink3D{0.0};
        }
        faces.push_back({*slot, face});
      }
  std::sort(faces.begin(), faces.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<MacPressureFace3D> face_list;
  for (const auto& e : faces) face_list.push_back(e.second);
  const MacPressureInputs3D inputs{*grid, dt, cells, face_list};
  auto got = ProjectMacPressure3D(inputs);
  EXPECT(got.ok());
  const auto& d = got->diagnostics;
  double b2 = 0;
  for (double b : got->integrated_rhs) b2 += b * b;
  EXPECT(d.true_residual_l2 <= 2e-6 + 2e-6 * std::sqrt(b2));
  EXPECT(d.max_volume_scaled_residual <= 2e-5 && d.max_cell_constraint_defect <= 2e-5);
  // Oracle projection (FP64 V-cycle) of the same system: pressures agree to the solve tolerance.
  std::vector<flo_pressure_cell> oc(cells.size());
  std::vector<flo_pressure_face> of(face_list.size());
  for (std::size_t i = 0; i < cells.size(); ++i)
    oc[i] = {{cells[i].cell_index.x, cells[i].cell_index.y, cells[i].cell_index.z}, 0, cells[i].fluid_volume,
             0.0, cells[i].bulk_modulus, 0.0, 0.0};
  for (std::size_t i = 0; i < face_list.size(); ++i) {
    const auto& f = face_list[i];
    const int kind = std::holds_alternative<InternalMacPressureLink3D>(f.pressure_link)
                         ? 1
                         : (std::holds_alternative<DirichletMacPressureLink3D>(f.pressure_link) ? 2 : 0);
    of[i] = {static_cast<std::int32_t>(f.face_index.axis),
             {f.face_index.lattice_index.x, f.face_index.lattice_index.y, f.face_index.lattice_index.z},
             f.open_area, f.inverse_density, f.pressure_gradient_distance, f.predicted_normal_velocity,
             f.wall_normal_velocity, kind, 0, 0.0, 0.0};
  }
  const flo_pressure_options opt{256, 1.0, 1e-12, 4, 2, 16, 0, 0.5};
  std::vector<double> want(cells.size());
  flo_pressure_diagnostics od{};
  const auto blocks = grid->active_block_indices();
  EXPECT(flo_project(O(&geom), O(blocks), blocks.size(), dt, oc.data(), oc.size(), of.data(), of.size(), &opt,
                     want.data(), nullptr, nullptr, nullptr, nullptr, nullptr, &od) == 0);
  double pmax = 0, err = 0;
  for (std::size_t i = 0; i < want.size(); ++i) {
    pmax = std::max(pmax, std::fabs(want[i]));
    err = std::max(err, std::fabs(got->cell_pressures[i] - want[i]));
  }
  std::printf("  projection: %zu cells, %zu iterations, max |dp| / max |p| = %.3g\n", cells.size(),
              d.iteration_count, err / pmax);
  EXPECT(err <= 1e-5 * pmax);
  // Invalid data: a nonpositive step duration is an argument error and publishes nothing.
  EXPECT(ProjectMacPressure3D({*grid, -1.0, cells, face_list}).status().code() == StatusCode::kInvalidArgument);
}

void SharedRuntimeGrows() {
  auto small = SparseMacGrid3D::Create(kGeometry, DenseBlocks(2, 2, 2), 8);
  auto large = SparseMacGrid3D::Create(kGeometry, DenseBlocks(6, 6, 6), 512);
  EXPECT(small.ok() && large.ok());
  for (auto* grid : {&*small, &*large}) {
    const int cells = static_cast<int>(grid->active_block_indices().size() == 8 ? 6 : 22);
    const auto particles = BoxParticles(1, cells, 21, 1.0);
    const MacDepositQuantization3D q{0x1p-40, 0x1p-40, particles.size()};
    std::vector<MacFaceDepositQuanta3D> got(grid->face_count()), want(grid->face_count());
    EXPECT(DepositMacQuanta3D(*grid, particles, q, got).ok());
    const auto blocks = grid->active_block_indices();
    EXPECT(flo_deposit_quanta(O(&kGeometry), O(blocks), blocks.size(),
                              reinterpret_cast<const flo_particle*>(particles.data()), particles.size(),
                              {q.mass_per_quantum, q.pure_momentum_per_quantum, q.declared_particle_capacity},
                              std::uint64_t{1} << 40, reinterpret_cast<flo_face_quanta*>(want.data()), want.size(),
                              4) == 0);
    EXPECT(std::memcmp(got.data(), want.data(), got.size() * sizeof(got[0])) == 0);
  }
}

// LiquidStepWorkspace3D.RemapsHistoryByIdentityAcrossShiftedBlockOrdinals
// (native/tests/test_liquid_step_workspace_3d.cc:142-190), then a random history against the oracle.
void WorkspaceRemapsHistoryAndMatchesOracle() {
  const UniformMacGeometry3D unit{{0.0, 0.0, 0.0}, 1.0};
  const LiquidStateCapacities3D caps{16, 32, 16, 32 * 192, 0};
  auto ws = LiquidStepWorkspace3D::Create(caps, unit);
  EXPECT(ws.ok());
  EXPECT(ws->prepared().status().code() == StatusCode::kFailedPrecondition);
  const std::vector<CellIndex3D> old_cells{{1, 1, 1}, {9, 1, 1}};
  EXPECT(ws->PrepareTrialTopology({}, old_cells).ok());
  auto view = ws->prepared();
  EXPECT(view.ok() && view->blocks.size() == 2 && view->faces.size() == 384);
  std::vector<LiquidBlockState3D> blocks(view->blocks.begin(), view->blocks.end());
  std::vector<LiquidCellState3D> cells(view->cells.begin(), view->cells.end());
  std::vector<LiquidFaceState3D> faces(view->faces.begin(), view->faces.end());
  blocks[0].generation = 4;
  blocks[0].activation_age = 7;
  blocks[0].inactive_age = 2;
  cells[0].previous_pressure = 123.5;
  cells[0].occupancy = LiquidCellOccupancy3D::kSolid;
  for (auto& f : faces) {
    f.carrier_velocity = 10.0 + static_cast<double>(f.face_slot);
    f.validity = LiquidFaceValidity3D::kLowMass;
    f.side_tag = -17;
    f.structurally_eligible = false;
  }
  const LiquidTopologyHistory3D previous{unit, 11, blocks, cells, faces};
  const std::vector<CellIndex3D> new_cells{{-3, 1, 1}, {1, 1, 1}};
  EXPECT(ws->PrepareTrialTopology({}, new_cells, &previous).ok());
  view = ws->prepared();
  EXPECT(view.ok() && view->blocks.size() == 2);
  EXPECT(view->blocks[0].block_index.x == -1 && view->blocks[0].generation == 12 && view->blocks[0].activation_age == 0);
  EXPECT(view->blocks[1].generation == 4 && view->blocks[1].activation_age == 7 && view->blocks[1].inactive_age == 0);
  EXPECT(view->cells[0].previous_pressure == 0 && view->cells[0].occupancy == LiquidCellOccupancy3D::kLiquid);
  EXPECT(view->cells[1].previous_pressure == 123.5 && view->cells[1].occupancy == LiquidCellOccupancy3D::kSolid);
  for (std::size_t i = 0; i < 192; ++i) {
    EXPECT(view->faces[i].carrier_velocity == 0 && view->faces[i].validity == LiquidFaceValidity3D::kMissing);
    LiquidFaceState3D expected = faces[i];
    expected.face_slot += 192;
    EXPECT(view->faces[i + 192] == expected);
  }
  // A failed preparation keeps the published bank.
  const std::vector<CellIndex3D> bad{{2147483647, 0, 0}};
  EXPECT(ws->PrepareTrialTopology({}, bad).code() == StatusCode::kOutOfRange);
  EXPECT(ws->prepared()->blocks.size() == 2 && ws->prepared()->cells[1].previous_pressure == 123.5);
  LiquidTopologyHistory3D stale = previous;
  stale.epoch = UINT64_MAX;
  EXPECT(ws->PrepareTrialTopology({}, new_cells, &stale).code() == StatusCode::kOutOfRange);

  // Random particles and cells with a remapped history: records equal the oracle's.
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(-2.0, 2.0);
  std::uniform_int_distribution<int> ci(-10, 10);
  const UniformMacGeometry3D g{{-0.37, 0.11, 0.5}, 0.25};
  const LiquidStateCapacities3D big{1 << 16, 1 << 12, 1 << 16, 192 << 12, 0};
  std::vector<Vec3> pos(2000);
  for (auto& p : pos) p = {u(rng), u(rng), u(rng)};
  std::vector<CellIndex3D> cl(500);
  for (auto& c : cl) c = {ci(rng), ci(rng), ci(rng)};
  auto w2 = LiquidStepWorkspace3D::Create(big, g);
  EXPECT(w2.ok() && w2->PrepareTrialTopology(pos, cl).ok());
  auto v0 = w2->prepared();
  std::vector<LiquidBlockState3D> b0(v0->blocks.begin(), v0->blocks.end());
  std::vector<LiquidCellState3D> c0(v0->cells.begin(), v0->cells.end());
  std::vector<LiquidFaceState3D> f0(v0->faces.begin(), v0->faces.end());
  for (std::size_t i = 0; i < f0.size(); ++i) f0[i].carrier_velocity = std::sin(double(i));
  for (std::size_t i = 0; i < c0.size(); ++i) c0[i].previous_pressure = double(i);
  for (auto& p : pos) p = {p.x + 0.3, p.y - 0.2, p.z};
  for (auto& c : cl) c = {c.x + 1, c.y, c.z - 2};
  const LiquidTopologyHistory3D h0{g, 3, b0, c0, f0};
  HipOperatorCapacities3D rc;
  rc.particles = 1 << 16;
  rc.blocks = 1 << 12;
  rc.lookup_blocks = 1 << 16;
  rc.pressure_cells = 1 << 16;
  auto rt = HipOperatorRuntime3D::Create(rc);
  EXPECT(rt.ok());
  HipOperatorMeasurement m;
  EXPECT(w2->HipPrepareTrialTopology(**rt, pos, cl, &h0, m).ok());
  auto v1 = w2->prepared();
  const flo_topology_capacities oc{big.particles, big.blocks, big.cells, big.faces};
  const flo_topology_history oh{*O(&g), 3, reinterpret_cast<const flo_block_state*>(b0.data()), b0.size(),
                                reinterpret_cast<const flo_cell_state*>(c0.data()), c0.size(),
                                reinterpret_cast<const flo_face_state*>(f0.data()), f0.size()};
  std::vector<flo_sort_entry> wo(pos.size());
  std::vector<flo_block_state> wb(big.blocks);
  std::vector<flo_cell_state> wc(cl.size());
  std::vector<flo_face_state> wf(big.faces);
  std::uint64_t nb = 0, nc = 0, nf = 0;
  EXPECT(flo_prepare_topology(O(&g), &oc, reinterpret_cast<const flo_vec3*>(pos.data()), pos.size(),
                              reinterpret_cast<const flo_int3*>(cl.data()), cl.size(), &oh, wo.data(), wb.data(), &nb,
                              wc.data(), &nc, wf.data(), &nf) == 0);
  EXPECT(nb == v1->blocks.size() && nc == v1->cells.size() && nf == v1->faces.size());
  for (std::size_t i = 0; i < pos.size(); ++i)
    EXPECT(v1->particle_order[i].original_ordinal == wo[i].original_ordinal);
  for (std::size_t i = 0; i < nb; ++i) {
    const LiquidBlockState3D& a = v1->blocks[i];
    EXPECT(a.block_index.x == wb[i].block_index.x && a.block_index.z == wb[i].block_index.z &&
           a.generation == wb[i].generation && a.activation_age == wb[i].activation_age);
  }
  for (std::size_t i = 0; i < nc; ++i)
    EXPECT(v1->cells[i].cell_index.x == wc[i].cell_index.x && v1->cells[i].previous_pressure == wc[i].previous_pressure);
  for (std::size_t i = 0; i < nf; ++i)
    EXPECT(v1->faces[i].face_slot == wf[i].face_slot && v1->faces[i].carrier_velocity == wf[i].carrier_velocity &&
           v1->faces[i].face_index.lattice_index.y == wf[i].face_index.lattice.y);
  std::printf("  topology: %zu blocks, %zu cells, %.3f ms on %s\n", v1->blocks.size(), v1->cells.size(), m.kernel_ms,
              m.architecture);
}

// The backend table's update_particles stage against the oracle on random particles.
void BackendUpdateMatchesOracle() {
  std::mt19937_64 rng(3);
  std::normal_distribution<double> n(0.0, 1.0);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  const std::size_t count = 3000;
  std::vector<LiquidBoundaryParticleGather3D> gathered(count);
  std::vector<double> distances(count);
  std::vector<LiquidParticleState3D> particles(count);
  for (std::size_t i = 0; i < count; ++i) {
    auto& g = gathered[i];
    g.new_pic_velocity = {n(rng), n(rng), n(rng)};
    g.particle_minus_old_pic_velocity = {0.3 * n(rng), 0.3 * n(rng), 0.3 * n(rng)};
    g.velocity_gradient = {{n(rng), n(rng), n(rng)}, {n(rng), n(rng), n(rng)}, {n(rng), n(rng), n(rng)}};
    g.fitted_velocity_slopes = {{n(rng), 0, 0}, {0, n(rng), 0}, {0, 0, n(rng)}};
    g.affine_enabled = u(rng) < 0.7;
    g.old_pic_fallback = u(rng) < 0.1;
    distances[i] = -0.5 + u(rng);
    auto& p = particles[i];
    p.id = i + 1;
    p.world_position = {3 * u(rng), 3 * u(rng), 3 * u(rng)};
    p.volume_ratio = 0.5 + u(rng);
    p.pic_initialized = u(rng) < 0.8;
  }
  const LiquidParticleUpdateInputs3D in{gathered, distances, {}, 0.25, 0.01, 0.7, 0.0, 0.0};
  const std::vector<LiquidParticleState3D> original = particles;
  std::vector<LiquidParticleState3D> want = particles;
  const flo_update_options o{0.25, 0.01, 0.7, 0.0, 0.0};
  EXPECT(flo_update_particles(&o, reinterpret_cast<const flo_boundary_gather*>(gathered.data()), distances.data(),
                              count, reinterpret_cast<flo_liquid_particle*>(want.data())) == 0);
  auto context = HipLiquidSubstepContext3D::Create({1 << 12, 64, 256, 192 * 64, 0});
  EXPECT(context.ok());
  const auto backend = MakeHipLiquidSubstepBackend3D(**context);
  EXPECT(backend.update_particles(backend.context, in, particles).ok());
  for (std::size_t i = 0; i < count; ++i) {
    const auto& a = particles[i];
    const auto& b = want[i];
    EXPECT(a.world_position.x == b.world_position.x && a.world_position.y == b.world_position.y &&
           a.world_position.z == b.world_position.z && a.world_velocity.x == b.world_velocity.x &&
           a.velocity_slopes.y_row.y == b.velocity_slopes.y_row.y && a.affine_enabled == b.affine_enabled &&
           a.pic_initialized && std::abs(a.volume_ratio - b.volume_ratio) <= 4.5e-16 * b.volume_ratio);
  }
  // A moving half-space at x = 0: particles near it hit it within the step.
  std::vector<LiquidParticleState3D> start = original, walled = original;
  const LiquidParticleUpdateInputs3D wall_in{gathered, distances, TranslatingHalfspace3D{{1, 0, 0}, 0.0, {0.3, 0, 0}},
                                             0.25, 0.01, 0.7, 0.2, 0.0};
  flo_update_options ow{0.25, 0.01, 0.7, 0.2, 0.0, {}};
  ow.obstacle.kind = FLO_OBSTACLE_HALFSPACE;
  ow.obstacle.unit_normal = {1, 0, 0};
  ow.obstacle.wall_velocity = {0.3, 0, 0};
  EXPECT(flo_update_particles(&ow, reinterpret_cast<const flo_boundary_gather*>(gathered.data()), distances.data(),
                              count, reinterpret_cast<flo_liquid_particle*>(walled.data())) == 0);
  EXPECT(backend.update_particles(backend.context, wall_in, start).ok());
  std::size_t moved = 0;
  for (std::size_t i = 0; i < count; ++i) {
    const auto& a = start[i];
    const auto& b = walled[i];
    EXPECT(a.world_position.x == b.world_position.x && a.world_position.y == b.world_position.y &&
           a.world_position.z == b.world_position.z && a.world_velocity.x == b.world_velocity.x &&
           a.world_velocity.y == b.world_velocity.y && a.world_velocity.z == b.world_velocity.z);
    moved += a.world_velocity.x != want[i].world_velocity.x;
  }
  std::printf("  half-space contact changed %zu of %zu stored velocities\n", moved, count);
  // The open tank through the host mapping (liquid_wall_contact_3d_cases.h:38-145, first case): piston at
  // x = 0.25 moving at +0.25 m/s, walls x <= 2, floor y = 0.125, z in [0.125, 0.75]; the tied first hit
  // takes (0.5, 0.375, 0.375) to (0.5, 0.125, 0.625) with stored velocity (0.25, 0, 0.125).
  {
    std::vector<LiquidBoundaryParticleGather3D> tg(1);
    tg[0].new_pic_velocity = {-0.25, -0.5, 0.25};
    tg[0].particle_minus_old_pic_velocity = {-1.0 + 0.25, -2.0 + 0.5, 0.125 - 0.25};
    const std::vector<double> td(1, 0.0);
    std::vector<LiquidParticleState3D> tp(1);
    tp[0].world_position = {0.5, 0.375, 0.375};
    tp[0].volume_ratio = 1.0;
    tp[0].pic_initialized = true;
    tp[0].mass = 1.0;
    tp[0].constituent_count = 1;
    tp[0].constituents[0].volume_fraction = 1.0;
    const LiquidParticleUpdateInputs3D tank_in{tg, td, AxisAlignedLiquidTank3D{0.25, 2.0, 0.125, 0.125, 0.75, 0.25},
                                               1e3, 1.0, 1.0 / 0.99, 0.0, 0.0};
    EXPECT(UpdateLiquidParticles3D(tank_in, tp).ok());
    const auto near = [](double a, double b) { return std::abs(a - b) <= 2e-14; };
    EXPECT(near(tp[0].world_position.x, 0.5) && near(tp[0].world_position.y, 0.125) &&
           near(tp[0].world_position.z, 0.625));
    EXPECT(near(tp[0].world_velocity.x, 0.25) && near(tp[0].world_velocity.y, 0.0) &&
           near(tp[0].world_velocity.z, 0.125));
  }
  // A mismatched span is InvalidArgument and leaves the particles untouched.
  const LiquidParticleUpdateInputs3D short_in{std::span(gathered).first(10), distances, {}, 0.25, 0.01, 0.7, 0.0,
                                              0.0};
  EXPECT(UpdateLiquidParticles3D(short_in, particles).code() == StatusCode::kInvalidArgument);
}

// LiquidBoundaryWorkspace3D.FrozenGeometryAndTypedLinksMatchOwningAndUnitCube and
// MissingIncidentFacePreservesPublishedPressureLinks (native/tests/test_liquid_boundary_workspace_3d.cc:75-101,
// :601-620) on a hand-frozen unit cube, then the projection of those links.
void PressureLinksUnitCube() {
  const UniformMacGeometry3D unit{{0.0, 0.0, 0.0}, 1.0};
  const std::vector<BlockIndex3D> keys{{0, 0, 0}};
  auto grid = SparseMacGrid3D::Create(unit, keys, 1);
  EXPECT(grid.ok());
  std::vector<LiquidCenterDistance3D> centers;
  for (int z = 0; z <= 2; ++z)
    for (int y = 0; y <= 2; ++y)
      for (int x = 0; x <= 2; ++x) centers.push_back({{x, y, z}, x == 1 && y == 1 && z == 1 ? -0.5 : 0.5});
  const std::vector<FrozenLiquidCell3D> cells{{{1, 1, 1}, 1.0, -0.5, 0.0}};
  std::vector<FrozenLiquidFace3D> faces;
  for (auto [axis, lx, ly, lz] : std::vector<std::array<int, 4>>{{0, 1, 1, 1}, {0, 2, 1, 1}, {1, 1, 1, 1},
                                                                 {1, 1, 2, 1}, {2, 1, 1, 1}, {2, 1, 1, 2}}) {
    const FaceIndex3D f{static_cast<FaceNormalAxis>(axis), {lx, ly, lz}};
    auto slot = grid->FaceSlot(f);
    EXPECT(slot.ok());
    faces.push_back({f, *slot, 1.0, 0.0, 0.0});
  }
  const FrozenLiquidBoundaryGeometryView3D view{grid->view(), centers, cells, faces, 0.01, 1000.0, 0.0, 0.0};
  const std::vector<LiquidPressureCellFields3D> fields{{0.0, std::numeric_limits<double>::infinity(), 7.0}};
  const std::vector<double> velocities(6, 0.0);
  auto links = BuildLiquidPressureLinks3D(view, {fields, velocities, 9.0});
  EXPECT(links.ok() && links->faces.size() == 6 && links->lower_theta_regularization_count == 0);
  EXPECT(links->cells[0].previous_pressure == 7.0);
  for (const auto& f : links->faces) {
    EXPECT(f.pressure_gradient_distance == 0.5 && f.inverse_density == 0.001);
    EXPECT(std::get<DirichletMacPressureLink3D>(f.pressure_link).exterior_pressure == 9.0);
  }
  // With no flux the cell takes the exterior pressure.
  auto projected = ProjectMacPressure3D({*grid, 0.01, links->cells, links->faces});
  EXPECT(projected.ok() && std::abs(projected->cell_pressures[0] - 9.0) <= 1e-9);
  FrozenLiquidBoundaryGeometryView3D missing = view;
  missing.faces = std::span(faces).first(5);
  EXPECT(BuildLiquidPressureLinks3D(missing, {fields, std::span(velocities).first(5), 9.0}).status().code() ==
         StatusCode::kFailedPrecondition);
  FrozenLiquidBoundaryGeometryView3D heavy = view;
  heavy.liquid_density = std::numeric_limits<double>::max();
  EXPECT(BuildLiquidPressureLinks3D(heavy, {fields, velocities, 9.0}).status().code() == StatusCode::kOutOfRange);
}

// The reference's freeze cases on its 2x2x2-block grid (native/tests/liquid_boundary_geometry_3d_cases.h):
// centre samples at field(i + 2.5), corner samples at field(i + 2), candidate (-2, -2, -2).
struct FreezeSamples {
  std::vector<CellIndex3D> candidates{{-2, -2, -2}};
  std::vector<LiquidCenterDistance3D> centers;
  std::vector<LiquidCornerDistance3D> corners;
};
template <typename Field>
FreezeSamples SampleField(Field field) {
  FreezeSamples s;
  for (int z = -4; z <= 1; ++z)
    for (int y = -4; y <= 1; ++y)
      for (int x = -4; x <= 1; ++x) {
        s.centers.push_back({{x, y, z}, field(x + 2.5, y + 2.5, z + 2.5)});
        s.corners.push_back({{x, y, z}, field(x + 2.0, y + 2.0, z + 2.0)});
      }
  return s;
}

void FreezeMatchesReferenceCases() {
  std::vector<BlockIndex3D> keys;
  for (int z = -1; z <= 0; ++z)
    for (int y = -1; y <= 0; ++y)
      for (int x = -1; x <= 0; ++x) keys.push_back({x, y, z});
  auto grid = SparseMacGrid3D::Create({{2.0, 2.0, 2.0}, 1.0}, keys, keys.size());
  EXPECT(grid.ok());
  const auto liquid = SampleField([](double, double, double) { return -1.0; });
  // Plane fixture: a half-space at 0.25 along x leaves an open width of 0.75.
  auto plane = FreezeLiquidBoundaryGeometry3D(
      {*grid, liquid.candidates, liquid.centers, liquid.corners, TranslatingHalfspace3D{{1, 0, 0}, 0.25, {2, 3, 4}},
       0.01, 1000.0, 0.0});
  EXPECT(plane.ok() && plane->cells().size() == 1 && plane->faces().size() == 6);
  EXPECT(std::abs(plane->cells()[0].fluid_volume - 0.75) <= 2e-14);
  for (const auto& f : plane->faces()) {
    EXPECT(f.blocked_area == 1.0 - f.open_area);
    EXPECT(f.wall_normal_velocity == 2.0 + static_cast<int>(f.face_index.axis));
    if (f.face_index.axis != FaceNormalAxis::kX) EXPECT(std::abs(f.open_area - 0.75) <= 2e-14);
  }
  EXPECT(plane->capillary_step_bound() == std::numeric_limits<double>::infinity());
  // A thin slab leaves 15/32 of a cell whose liquid lies beyond it; an all-liquid cell is split.
  const TranslatingSlab3D slab{{1, 0, 0}, 0.5, 1.0 / 32.0, {0, 0, 0}};
  const auto beyond = SampleField([](double x, double, double) { return x - 0.52; });
  auto thin = FreezeLiquidBoundaryGeometry3D(
      {*grid, beyond.candidates, beyond.centers, beyond.corners, slab, 0.01, 1000.0, 0.0});
  EXPECT(thin.ok() && std::abs(thin->cells()[0].fluid_volume - 15.0 / 32.0) <= 2e-14);
  EXPECT(FreezeLiquidBoundaryGeometry3D({*grid, liquid.candidates, liquid.centers, liquid.corners, slab, 0.01, 1000.0,
                                         0.0})
             .status()
             .code() == StatusCode::kFailedPrecondition);
  // The capillary bound of a plane interface.
  const auto surface = SampleField([](double x, double, double) { return x - 0.75; });
  auto capillary = FreezeLiquidBoundaryGeometry3D(
      {*grid, surface.candidates, surface.centers, surface.corners, {}, 0.01, 1000.0, 0.072});
  EXPECT(capillary.ok() && std::abs(capillary->cells()[0].curvature) <= 1e-14);
  EXPECT(std::abs(capillary->capillary_step_bound() - std::sqrt(1000.0 / (4.0 * std::acos(-1.0) * 0.072))) <= 1e-13);
  // The open tank clips the liquid below y = 0.75 to 75/256 of the cell and puts the piston's flux 35/512
  // on the low-x face only (OpenTankPreservesNarrowAndLiquidClippedMeasures); a piston reaching the
  // high-x wall within the step is refused.
  const auto pool = SampleField([](double, double y, double) { return y - 0.75; });
  const AxisAlignedLiquidTank3D tank{0.25, 2.0, 0.125, 0.125, 0.75, 0.125};
  auto in_tank = FreezeLiquidBoundaryGeometry3D(
      {*grid, pool.candidates, pool.centers, pool.corners, tank, 0.01, 1000.0, 0.0});
  EXPECT(in_tank.ok() && in_tank->cells().size() == 1 &&
         std::abs(in_tank->cells()[0].fluid_volume - 0.29296875) <= 2e-14);
  int pistons = 0;
  for (const auto& face : in_tank->faces()) {
    const bool piston = face.face_index.axis == FaceNormalAxis::kX && face.face_index.lattice_index.x == -2;
    EXPECT(face.prescribed_wall_flux.has_value() && face.is_low_x_piston_face == piston);
    EXPECT(*face.prescribed_wall_flux == (piston ? 0.068359375 : 0.0));
    pistons += piston ? 1 : 0;
  }
  EXPECT(pistons == 1);
  EXPECT(FreezeLiquidBoundaryGeometry3D({*grid, pool.candidates, pool.centers, pool.corners,
                                         AxisAlignedLiquidTank3D{0.0, 1.0, 0.125, 0.125, 0.75, 1.0}, 1.0, 1000.0,
                                         0.0})
             .status()
             .code() == StatusCode::kFailedPrecondition);
  // More than 4096 candidates are refused first.
  std::vector<CellIndex3D> many(4097, CellIndex3D{-2, -2, -2});
  EXPECT(FreezeLiquidBoundaryGeometry3D({*grid, many, liquid.centers, liquid.corners, {}, 0.0, 1000.0, -1.0})
             .status()
             .code() == StatusCode::kResourceExhausted);

  // Sealed low face at the minimum signed index: freeze, links at ambient 2 Pa, projection gives 2 Pa
  // (SealedLowFaceAtMinimumSignedIndexNeedsNoExteriorSample).
  constexpr int kMin = std::numeric_limits<std::int32_t>::min();
  const std::vector<BlockIndex3D> edge{{kMin / 4, -1, -1}};
  auto low = SparseMacGrid3D::Create({{2147483648.0, 2.0, 2.0}, 1.0}, edge, 1);
  EXPECT(low.ok());
  auto iso = SampleField([](double, double, double) { return 0.99; });
  for (auto& c : iso.corners) c.signed_distance = -1.0;
  for (auto& c : iso.centers)
    if (c.cell_index.x == -2 && c.cell_index.y == -2 && c.cell_index.z == -2) c.signed_distance = -0.01;
  std::erase_if(iso.centers, [](const auto& c) { return c.cell_index.x < -2; });
  std::erase_if(iso.corners, [](const auto& c) { return c.corner_index.x < -2; });
  for (auto& c : iso.centers) c.cell_index.x = kMin + (c.cell_index.x + 2);
  for (auto& c : iso.corners) c.corner_index.x = kMin + (c.corner_index.x + 2);
  iso.candidates[0].x = kMin;
  auto frozen = FreezeLiquidBoundaryGeometry3D(
      {*low, iso.candidates, iso.centers, iso.corners, TranslatingHalfspace3D{{1, 0, 0}, 0.25, {0, 0, 0}}, 0.01,
       1000.0, 0.0});
  EXPECT(frozen.ok());
  if (!frozen.ok()) return;
  const std::vector<LiquidPressureCellFields3D> fields(frozen->cells().size(),
                                                       {0.0, std::numeric_limits<double>::infinity(), 0.0});
  const std::vector<double> velocities(frozen->faces().size(), 0.0);
  auto links = BuildLiquidPressureLinks3D(frozen->view(), {fields, velocities, 2.0});
  EXPECT(links.ok());
  if (!links.ok()) return;
  auto projected = ProjectMacPressure3D({*low, 0.01, links->cells, links->faces});
  EXPECT(projected.ok() && std::abs(projected->cell_pressures[0] - 2.0) <= 1e-6);
}

int Run(const char* name, const std::function<void()>& test) {
  std::printf("[ RUN      ] %s\n", name);
  std::fflush(stdout);
  const int before = g_failures;
  test();
  std::printf(g_failures == before ? "[       OK ] %s\n" : "[  FAILED  ] %s\n", name);
  std::fflush(stdout);
  return g_failures == before ? 0 : 1;
}

}  // namespace

// Strain rows of a box of cells with varying viscosity, bit-exact against the oracle; the velocity block
// of those rows against the oracle's FP64 Jacobi PCG (same iterations, 1e-10 of the largest velocity);
// an unsorted cell list is InvalidArgument with no rows.
void ViscosityMatchesOracle() {
  const auto keys = DenseBlocks(4, 4, 4);
  auto grid = SparseMacGrid3D::Create(kGeometry, keys, 64);
  EXPECT(grid.ok());
  const auto blocks = grid->active_block_indices();
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(0.5, 2.0);
  std::vector<MacViscosityCell3D> cells;
  for (int z = 2; z < 13; ++z)
    for (int y = 2; y < 11; ++y)
      for (int x = 3; x < 14; ++x) cells.push_back({{x, y, z}, u(rng)});
  auto rows = AssembleFullCellMacStrain3D(grid->view(), cells);
  if (!rows.ok()) std::printf("  assembly status %d\n", static_cast<int>(rows.status().code()));
  EXPECT(rows.ok());
  std::vector<flo_strain_row> expected(6 * cells.size());
  std::size_t count = 0;
  EXPECT(flo_assemble_full_cell_strain(O(&kGeometry), O(blocks), blocks.size(),
                                       reinterpret_cast<const flo_viscosity_cell*>(cells.data()), cells.size(),
                                       expected.data(), expected.size(), &count) == 0);
  EXPECT(rows->size() == count && count > 3 * cells.size());
  EXPECT(std::memcmp(rows->data(), expected.data(), count * sizeof(flo_strain_row)) == 0);

  std::vector<double> inertia(grid->face_count(), 0.0), rhs(grid->face_count(), 0.0);
  for (const auto& row : *rows)
    for (std::size_t t = 0; t < row.term_count; ++t) {
      const std::size_t s = row.terms[t].face_slot;
      inertia[s] = 1000.0 * std::pow(kGeometry.cell_width, 3) * (0.5 + 0.5 * std::sin(0.37 * s) * std::sin(0.37 * s));
      rhs[s] = inertia[s] * std::cos(0.11 * s);
    }
  MacViscousVelocityBlockInputs3D in{*rows, inertia, 0.01};
  auto solved = SolveMacViscousVelocityBlock3D(in, rhs, {256, 1e-10});
  EXPECT(solved.ok());
  std::vector<double> v(inertia.size());
  std::uint64_t iterations = 0;
  double max_residual = 0;
  EXPECT(flo_solve_viscous_block(expected.data(), count, inertia.data(), inertia.size(), 0.01, rhs.data(), 256,
                                 1e-10, v.data(), &iterations, &max_residual) == 0);
  double scale = 0, diff = 0;
  for (std::size_t i = 0; i < v.size(); ++i) {
    scale = std::max(scale, std::abs(v[i]));
    diff = std::max(diff, std::abs(v[i] - solved->velocity[i]));
  }
  std::printf("  viscous block: %zu rows, %llu iterations (oracle %llu), max |dv| %.3e of %.3e\n", count,
              static_cast<unsigned long long>(solved->iterations), static_cast<unsigned long long>(iterations), diff,
              scale);
  EXPECT(scale > 0 && diff <= 1e-10 * scale);
  EXPECT(solved->iterations + 1 >= iterations && solved->iterations <= iterations + 1);

  std::vector<MacViscosityCell3D> unsorted = {cells[1], cells[0]};
  auto bad = AssembleFullCellMacStrain3D(grid->view(), unsorted);
  EXPECT(bad.status().code() == StatusCode::kInvalidArgument);
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  int failed = 0;
  if (mode == "cpu") {
    failed += Run("GridCreateSortsAndRejects", GridCreateSortsAndRejects);
    failed += Run("FaceSlotsMatchOracle", FaceSlotsMatchOracle);
    failed += Run("OperatorsWithoutDeviceFailPrecondition", OperatorsWithoutDeviceFailPrecondition);
  } else {
    failed += Run("DepositMatchesOracleBitExact", DepositMatchesOracleBitExact);
    failed += Run("DepositFailuresPreserveOutput", DepositFailuresPreserveOutput);
    failed += Run("Fp64DepositMatchesOracle", Fp64DepositMatchesOracle);
    failed += Run("CompatibleDepositMatchesOracle", CompatibleDepositMatchesOracle);
    failed += Run("BoundaryGatherMatchesOracle", BoundaryGatherMatchesOracle);
    failed += Run("GatherMatchesOracle", GatherMatchesOracle);
    failed += Run("SubstepTableDrivesStagesAgainstOracle", SubstepTableDrivesStagesAgainstOracle);
    failed += Run("ProjectionMeetsGatesAndMatchesOracle", ProjectionMeetsGatesAndMatchesOracle);
    failed += Run("WorkspaceRemapsHistoryAndMatchesOracle", WorkspaceRemapsHistoryAndMatchesOracle);
    failed += Run("BackendUpdateMatchesOracle", BackendUpdateMatchesOracle);
    failed += Run("PressureLinksUnitCube", PressureLinksUnitCube);
    failed += Run("FreezeMatchesReferenceCases", FreezeMatchesReferenceCases);
    failed += Run("ViscosityMatchesOracle", ViscosityMatchesOracle);
    failed += Run("SharedRuntimeGrows", SharedRuntimeGrows);
  }
  std::printf("%d test(s) failed\n", failed);
  return failed ? 1 : 0;
}
