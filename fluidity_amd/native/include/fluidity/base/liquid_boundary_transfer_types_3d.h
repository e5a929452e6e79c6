// Records of the compatible (masked) liquid deposit.
//
// Same names, fields and layouts as native/include/fluidity/base/liquid_boundary_transfer_types_3d.h:4-20;
// each record is layout-identical to its C-ABI twin in fluidity_hip.h.
#pragma once

#include <cstddef>
#include <cstdint>

#include "fluidity/base/mac_types_3d.h"

namespace fluidity {

struct LiquidBoundaryParticle3D {
  MacTransferParticle3D transfer;
  // Material side of the particle: a face sample is compatible only when its tag holds the same value.
  std::int64_t side_tag = 0;
  // When false the particle deposits as if its affine matrix were zero.
  bool affine_enabled = true;
};

struct LiquidBoundaryFaceTag3D {
  std::int64_t side_tag = 0;
  bool structurally_eligible = true;
};

// One of the 27 samples of a face axis after the compatibility mask.
struct CompatibleMacFaceSample3D {
  // Owned slot of the sample; set whether or not the sample took part.
  std::size_t face_slot;
  // The slot is eligible and its tag equals the particle's side tag.
  bool compatible;
  // Weight divided by the sum S of compatible weights; 0 for a masked sample.
  double normalized_weight;
  // Gradient of the normalized weight (quotient rule), per meter.
  Vec3 normalized_weight_gradient;
  // Face offset from the particle after subtracting the normalized-weight mean offset, meters.
  Vec3 centered_face_offset;
};

struct LiquidPrescribedWallSample3D {
  bool available = false;
  // The one side tag allowed to continue from this wall sample.
  std::int64_t side_tag = 0;
  // Unit wall normal in world axes (meaningful only when available).
  Vec3 unit_normal{};
  // Wall velocity in world axes at the sample time of the history, meters per second.
  Vec3 wall_velocity{};
};

struct LiquidBoundaryParticleGather3D {
  // PIC velocity gathered from the filled stencil of the new field, meters per second.
  Vec3 new_pic_velocity{};
  // Gathered new field minus gathered old field, meters per second.
  Vec3 grid_velocity_increment{};
  // Particle velocity minus the gathered old field, meters per second.
  Vec3 particle_minus_old_pic_velocity{};
  // Affine slopes from the second-moment fit of the new field, per second.
  Mat3 fitted_velocity_slopes{};
  // Gradient of the gathered new field with respect to position, per second.
  Mat3 velocity_gradient{};
  // Cleared when the input disabled the affine part or any sample needed a fallback.
  bool affine_enabled = true;
  // The old-field continuation used the nearest donor instead of the fit.
  bool old_pic_fallback = false;
  // The new-field continuation used the nearest donor instead of the fit.
  bool new_pic_fallback = false;
  // Some sample of either field lies on a face whose mass is at or below the minimum.
  bool low_mass_fallback = false;
};

enum class LiquidFaceValidity3D : std::uint8_t {
  kMissing = 0,
  kValid = 1,
  // No mass, but a finite continued value that may serve as a donor.
  kPredicted = 2,
  // Mass above zero but not above the minimum mass.
  kLowMass = 3,
};

}  // namespace fluidity
