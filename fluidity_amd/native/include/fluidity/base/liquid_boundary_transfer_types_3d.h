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

}  // namespace fluidity
