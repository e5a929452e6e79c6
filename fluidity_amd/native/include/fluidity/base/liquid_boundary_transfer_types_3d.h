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
  std::int64_t side_tag = 0;   // Equality identifies the same material/body side.
  bool affine_enabled = true;  // False deposits with a zero affine row.
};

struct LiquidBoundaryFaceTag3D {
  std::int64_t side_tag = 0;
  bool structurally_eligible = true;
};

struct CompatibleMacFaceSample3D {
  std::size_t face_slot;            // Original owned slot, kept for excluded samples too.
  bool compatible;                  // Eligible and on the particle's side.
  double normalized_weight;         // N / S; zero when excluded.
  Vec3 normalized_weight_gradient;  // Quotient-rule derivative, 1/m.
  Vec3 centered_face_offset;        // r minus the compatible weighted mean, m.
};

struct LiquidPrescribedWallSample3D {
  bool available = false;
  std::int64_t side_tag = 0;  // Wall continuation is usable only by this side.
  Vec3 unit_normal{};         // World Cartesian unit normal when available.
  Vec3 wall_velocity{};       // World m/s at the history's wall sample time.
};

struct LiquidBoundaryParticleGather3D {
  Vec3 new_pic_velocity{};                 // m/s from the complete new filled stencil.
  Vec3 grid_velocity_increment{};          // new PIC minus old PIC, m/s.
  Vec3 particle_minus_old_pic_velocity{};  // Retained particle residual, m/s.
  Mat3 fitted_velocity_slopes{};           // Complete second-moment fit, 1/s.
  Mat3 velocity_gradient{};                // True original-basis derivative, 1/s.
  bool affine_enabled = true;              // False on input affine-off or any fallback.
  bool old_pic_fallback = false;           // An old-stage continuation fell back from the fit.
  bool new_pic_fallback = false;           // A new-stage continuation fell back from the fit.
  bool low_mass_fallback = false;          // Either history touches a low-mass face.
};

enum class LiquidFaceValidity3D : std::uint8_t {
  kMissing = 0,
  kValid = 1,
  kPredicted = 2,  // Finite continuation usable as a donor, without mass.
  kLowMass = 3,    // Positive mass at or below the minimum.
};

}  // namespace fluidity
