// Oracle: typed pressure links of a frozen boundary geometry (test infrastructure only).
// Restates BuildLiquidPressureLinks3D on a frozen view: ValidateFrozenView
// (native/src/liquid_boundary_geometry_3d.cc:1289-1383), then LinksCore (:1077-1287) with the kLink
// geometry query (native/include/fluidity/base/liquid_boundary_geometry_math_3d.h:676-689) and
// InternalLiquidCapillaryStepBound3D (liquid_boundary_geometry_3d.cc:1669-1701).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numbers>
#include <numeric>
#include <vector>

#include "grid_internal.h"

using namespace flo;

namespace {

bool PositiveFinite(double v) { return std::isfinite(v) && v > 0; }
bool NormalOrZero(double v) { return v == 0.0 || std::isnormal(v); }

// OffsetCellIndex: the cell shifted along axis, or OutOfRange outside the int32 lattice.
int Offset(flo_int3 c, int axis, int amount, flo_int3& out) {
  int64_t v[3] = {c.x, c.y, c.z};
  v[axis] += amount;
  if (v[axis] < kI32Min || v[axis] > kI32Max) return FLO_OUT_OF_RANGE;
  out = {int32_t(v[0]), int32_t(v[1]), int32_t(v[2])};
  return FLO_OK;
}

uint64_t FindCell(const flo_frozen_view& g, flo_int3 key) {
  const flo_frozen_cell* b = g.cells;
  const flo_frozen_cell* e = b + g.cell_count;
  const flo_frozen_cell* f =
      std::lower_bound(b, e, key, [](const flo_frozen_cell& c, const flo_int3& k) { return BlockLess(c.cell_index, k); });
  return (f != e && BlockEq(f->cell_index, key)) ? uint64_t(f - b) : g.cell_count;
}

// LookupSignedDistance: FailedPrecondition when the sample is missing.
int LookupDistance(const flo_frozen_view& g, flo_int3 key, double& out) {
  const flo_center_distance* b = g.center_distances;
  const flo_center_distance* e = b + g.center_count;
  const flo_center_distance* f = std::lower_bound(
      b, e, key, [](const flo_center_distance& s, const flo_int3& k) { return BlockLess(s.cell_index, k); });
  if (f == e || !BlockEq(f->cell_index, key)) return FLO_FAILED_PRECONDITION;
  out = f->signed_distance;
  return FLO_OK;
}

int ValidateView(const flo_frozen_view& g) {
  if (!PositiveFinite(g.step_duration) || !PositiveFinite(g.liquid_density) || !std::isfinite(g.surface_tension) ||
      g.surface_tension < 0)
    return FLO_INVALID_ARGUMENT;
  const double h = g.geometry.cell_width;
  const double full_area = h * h;
  const double full_volume = full_area * h;
  const double inverse_density = 1 / g.liquid_density;
  if (!std::isnormal(full_area) || !std::isnormal(full_volume) || !std::isnormal(inverse_density))
    return FLO_OUT_OF_RANGE;
  for (uint64_t i = 0; i < g.center_count; ++i)
    if (!std::isfinite(g.center_distances[i].signed_distance) ||
        (i > 0 && !BlockLess(g.center_distances[i - 1].cell_index, g.center_distances[i].cell_index)))
      return FLO_INVALID_ARGUMENT;
  for (uint64_t i = 0; i < g.cell_count; ++i) {
    const flo_frozen_cell& c = g.cells[i];
    if (!PositiveFinite(c.fluid_volume) || c.fluid_volume > full_volume || !std::isfinite(c.center_distance) ||
        c.center_distance >= 0 || !std::isfinite(c.curvature))
      return FLO_INVALID_ARGUMENT;
    if (i > 0 && !BlockLess(g.cells[i - 1].cell_index, c.cell_index)) return FLO_INVALID_ARGUMENT;
  }
  for (uint64_t i = 0; i < g.face_count; ++i) {
    const flo_frozen_face& f = g.faces[i];
    uint64_t slot = 0;
    if (int s = flo_face_slot(g.blocks, g.block_count, f.face_index, &slot)) return s;
    if (slot != f.face_slot || (i > 0 && f.face_slot <= g.faces[i - 1].face_slot)) return FLO_INVALID_ARGUMENT;
    if (!std::isfinite(f.open_area) || f.open_area < 0 || f.open_area > full_area ||
        f.blocked_area != full_area - f.open_area || !std::isfinite(f.wall_normal_velocity))
      return FLO_INVALID_ARGUMENT;
    if (f.has_prescribed_wall_flux &&
        (!std::isfinite(f.prescribed_wall_flux) || (f.blocked_area == 0 && f.prescribed_wall_flux != 0)))
      return FLO_INVALID_ARGUMENT;
    const flo_int3 head = f.face_index.lattice;
    flo_int3 tail{};
    const int tail_status = Offset(head, f.face_index.axis, -1, tail);
    const bool has_head = FindCell(g, head) != g.cell_count;
    const bool has_tail = tail_status == FLO_OK && FindCell(g, tail) != g.cell_count;
    if (!has_head && !has_tail) return FLO_FAILED_PRECONDITION;
    if (tail_status != FLO_OK && f.open_area > 0) return tail_status;
  }
  for (uint64_t i = 0; i < g.cell_count; ++i)
    for (int axis = 0; axis < 3; ++axis)
      for (int side = 0; side < 2; ++side) {
        flo_int3 shifted{};
        if (int s = Offset(g.cells[i].cell_index, axis, side, shifted)) return s;
        uint64_t slot = 0;
        if (int s = flo_face_slot(g.blocks, g.block_count, flo_face_index{uint8_t(axis), shifted}, &slot)) return s;
        const flo_frozen_face* b = g.faces;
        const flo_frozen_face* e = b + g.face_count;
        const flo_frozen_face* found =
            std::lower_bound(b, e, slot, [](const flo_frozen_face& x, uint64_t v) { return x.face_slot < v; });
        if (found == e || found->face_slot != slot) return FLO_FAILED_PRECONDITION;
      }
  return FLO_OK;
}

int CapillaryBound(double tail_density, double head_density, double h, double sigma, double& out) {
  if (!PositiveFinite(tail_density) || !PositiveFinite(head_density) || !PositiveFinite(h) || !std::isfinite(sigma) ||
      sigma < 0.0)
    return FLO_INVALID_ARGUMENT;
  if (sigma == 0.0) {
    out = std::numeric_limits<double>::infinity();
    return FLO_OK;
  }
  const double larger = std::max(tail_density, head_density);
  const double ratio = std::min(tail_density, head_density) / larger;
  int de = 0, we = 0, te = 0;
  const double sd = std::frexp(larger, &de) * (1.0 + ratio);
  const double sw = std::frexp(h, &we);
  const double st = std::frexp(sigma, &te);
  const double square = sd * sw * sw * sw / (4.0 * std::numbers::pi * st);
  const int exponent = de + 3 * we - te;
  const double bound = std::scalbn(std::sqrt(std::scalbn(square, exponent % 2)), exponent / 2);
  if (!PositiveFinite(bound)) return FLO_OUT_OF_RANGE;
  out = bound;
  return FLO_OK;
}

void SetLink(flo_pressure_face& f, int kind, double lift) {
  f.link_kind = kind;
  f.lift = lift;
}

}  // namespace

extern "C" int flo_build_pressure_links(const flo_frozen_view* view, const flo_pressure_fields* fields,
                                        flo_pressure_cell* out_cells, flo_pressure_face* out_faces,
                                        uint64_t* out_lower_theta_count) {
  const flo_frozen_view& g = *view;
  const flo_pressure_fields& F = *fields;
  if (int s = ValidateView(g)) return s;
  // LinksCore: field extents and records.
  const bool has_wetted = F.wetted_count != 0;
  if (F.cell_count != g.cell_count || F.predicted_count != g.face_count ||
      (has_wetted && F.wetted_count != g.face_count) || F.air_count > g.center_count ||
      F.interface_count > g.face_count || !std::isfinite(F.ambient_pressure))
    return FLO_INVALID_ARGUMENT;
  for (uint64_t i = 0; i < F.interface_count; ++i) {
    const flo_internal_interface& c = F.internal_interfaces[i];
    if (c.face_ordinal >= g.face_count || (i > 0 && F.internal_interfaces[i - 1].face_ordinal >= c.face_ordinal) ||
        !std::isfinite(c.tail_fraction) || c.tail_fraction < 0.0 || c.tail_fraction > 1.0 ||
        !std::isfinite(c.surface_tension) || c.surface_tension < 0.0 || !std::isfinite(c.curvature_tail_to_head) ||
        !std::isfinite(c.normal_viscous_jump_tail_to_head))
      return FLO_INVALID_ARGUMENT;
  }
  for (uint64_t i = 0; i < F.air_count; ++i) {
    const flo_air_pressure& a = F.air_cell_pressures[i];
    if (!std::isfinite(a.gas_gauge_pressure) ||
        (i > 0 && !BlockLess(F.air_cell_pressures[i - 1].cell_index, a.cell_index)))
      return FLO_INVALID_ARGUMENT;
    double d = 0;
    if (int s = LookupDistance(g, a.cell_index, d)) return s;
    if (d < 0.0 || FindCell(g, a.cell_index) != g.cell_count) return FLO_INVALID_ARGUMENT;
  }
  std::vector<flo_pressure_cell> cells(g.cell_count);
  for (uint64_t i = 0; i < g.cell_count; ++i) {
    const flo_cell_fields& f = F.cells[i];
    if (!std::isfinite(f.target_divergence) || !std::isfinite(f.previous_pressure) || std::isnan(f.bulk_modulus) ||
        f.bulk_modulus <= 0 || (f.has_density && !PositiveFinite(f.density)))
      return FLO_INVALID_ARGUMENT;
    flo_pressure_cell& c = cells[i];
    std::memset(&c, 0, sizeof c);
    c.cell_index = g.cells[i].cell_index;
    c.fluid_volume = g.cells[i].fluid_volume;
    c.target_divergence = f.target_divergence;
    c.bulk_modulus = f.bulk_modulus;
    c.previous_pressure = f.previous_pressure;
    c.integrated_inward_wall_flux = 0.0;
  }
  const double h = g.geometry.cell_width;
  std::vector<flo_pressure_face> faces(g.face_count);
  uint64_t lower = 0;
  for (uint64_t i = 0; i < g.face_count; ++i) {
    const flo_frozen_face& face = g.faces[i];
    if (!std::isfinite(F.predicted_face_velocities[i])) return FLO_INVALID_ARGUMENT;
    flo_pressure_face& link = faces[i];
    std::memset(&link, 0, sizeof link);
    link.axis = face.face_index.axis;
    link.lattice = face.face_index.lattice;
    link.open_area = face.open_area;
    link.inverse_density = 1 / g.liquid_density;
    link.pressure_gradient_distance = h;
    link.predicted_normal_velocity = F.predicted_face_velocities[i];
    link.wall_normal_velocity = face.wall_normal_velocity;
    SetLink(link, 0, 0.0);
    link.has_prescribed_wall_flux = face.has_prescribed_wall_flux;
    link.prescribed_wall_flux = face.has_prescribed_wall_flux ? face.prescribed_wall_flux : 0.0;
    const flo_int3 head = face.face_index.lattice;
    flo_int3 tail{};
    const int tail_status = Offset(head, face.face_index.axis, -1, tail);
    const uint64_t head_found = FindCell(g, head);
    const uint64_t tail_found = tail_status == FLO_OK ? FindCell(g, tail) : g.cell_count;
    const bool has_head = head_found != g.cell_count, has_tail = tail_found != g.cell_count;
    if (has_wetted) {
      const bool receiver = face.is_low_x_piston_face && face.face_index.axis == 0 && face.open_area == 0 &&
                            has_head && !has_tail && face.has_prescribed_wall_flux;
      const flo_optional_flux& flux = F.wetted_piston_face_fluxes[i];
      if (receiver != bool(flux.has_value)) return FLO_INVALID_ARGUMENT;
      if (flux.has_value) {
        const double w = face.wall_normal_velocity;
        const bool wrong = (flux.value > 0 && w <= 0) || (flux.value < 0 && w >= 0);
        if (!std::isfinite(flux.value) || wrong) return FLO_INVALID_ARGUMENT;
        link.has_prescribed_wall_flux = 1;
        link.prescribed_wall_flux = flux.value;
      }
    }
    const flo_internal_interface* ib = F.internal_interfaces;
    const flo_internal_interface* ie = ib + F.interface_count;
    const flo_internal_interface* crossing = std::lower_bound(
        ib, ie, i, [](const flo_internal_interface& c, uint64_t v) { return c.face_ordinal < v; });
    if (crossing != ie && crossing->face_ordinal == i) {
      if (!has_head || !has_tail || face.open_area <= 0.0 || !F.cells[head_found].has_density ||
          !F.cells[tail_found].has_density)
        return FLO_INVALID_ARGUMENT;
      const double tail_density = F.cells[tail_found].density;
      const double head_density = F.cells[head_found].density;
      double bound = 0;
      if (int s = CapillaryBound(tail_density, head_density, h, crossing->surface_tension, bound)) return s;
      if (g.step_duration > bound) return FLO_ABORTED;
      const double face_density = std::lerp(head_density, tail_density, crossing->tail_fraction);
      const double inverse_density = 1.0 / face_density;
      const double jump =
          std::fma(-crossing->surface_tension, crossing->curvature_tail_to_head, crossing->normal_viscous_jump_tail_to_head);
      if (!std::isnormal(inverse_density) || !std::isfinite(jump)) return FLO_OUT_OF_RANGE;
      link.inverse_density = inverse_density;
      SetLink(link, 1, jump);
      continue;
    }
    if (tail_status != FLO_OK) {
      if (face.open_area > 0) return tail_status;
      continue;
    }
    if (!has_head && !has_tail) return FLO_FAILED_PRECONDITION;
    if (face.open_area > 0) {
      const bool hd = has_head && F.cells[head_found].has_density;
      const bool td = has_tail && F.cells[tail_found].has_density;
      if (hd || td) {
        const double head_density = hd ? F.cells[head_found].density : g.liquid_density;
        const double tail_density = td ? F.cells[tail_found].density : g.liquid_density;
        const double face_density = has_head && has_tail ? std::midpoint(head_density, tail_density)
                                                         : (has_head ? head_density : tail_density);
        link.inverse_density = 1.0 / face_density;
        if (!std::isnormal(link.inverse_density)) return FLO_OUT_OF_RANGE;
      }
    }
    if (has_head && has_tail) {
      SetLink(link, 1, 0.0);
    } else if (face.open_area > 0) {
      const flo_int3 air_cell = has_head ? tail : head;
      double air_distance = 0;
      if (int s = LookupDistance(g, air_cell, air_distance)) return s;
      if (air_distance < 0) return FLO_FAILED_PRECONDITION;
      const flo_frozen_cell& liquid = g.cells[has_head ? head_found : tail_found];
      double ambient = F.ambient_pressure;
      {
        const flo_air_pressure* ab = F.air_cell_pressures;
        const flo_air_pressure* ae = ab + F.air_count;
        const flo_air_pressure* p = std::lower_bound(
            ab, ae, air_cell, [](const flo_air_pressure& a, const flo_int3& k) { return BlockLess(a.cell_index, k); });
        if (p != ae && BlockEq(p->cell_index, air_cell)) ambient = p->gas_gauge_pressure;
      }
      // kLink (liquid_boundary_geometry_math_3d.h:676-689)
      const double magnitude = std::max(-liquid.center_distance, air_distance);
      const double liquid_scaled = -liquid.center_distance / magnitude;
      const double theta = liquid_scaled / (liquid_scaled + air_distance / magnitude);
      const bool regularized = theta < .05;
      const double distance = std::clamp(theta, .05, 1.0) * h;
      const double lift = g.surface_tension * liquid.curvature;
      const double exterior = ambient + lift;
      if ((g.surface_tension > 0 && liquid.curvature != 0 && lift == 0) || !NormalOrZero(lift) ||
          !std::isfinite(exterior) || !std::isnormal(distance))
        return FLO_OUT_OF_RANGE;
      lower += regularized ? 1 : 0;
      link.pressure_gradient_distance = distance;
      SetLink(link, 2, exterior);
    }
  }
  std::copy(cells.begin(), cells.end(), out_cells);
  std::copy(faces.begin(), faces.end(), out_faces);
  *out_lower_theta_count = lower;
  return FLO_OK;
}
