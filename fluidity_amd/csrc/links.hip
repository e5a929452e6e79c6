// Typed pressure links of a frozen boundary geometry on the device (BuildLiquidPressureLinks3D on a
// frozen view: ValidateFrozenView, native/src/liquid_boundary_geometry_3d.cc:1289-1383, then LinksCore,
// :1077-1287, with the free-surface link of base/liquid_boundary_geometry_math_3d.h:676-689).
//
// The reference checks records in sequential loops; here each loop is one kernel over its records
// and every failure writes min(phase, ordinal, code) into the context's error word, so the first
// failing record of the first failing loop is reported. Lookups (blocks, cells, centre samples, faces
// by slot, interfaces by face ordinal, air pressures) are binary searches over the sorted inputs.
// One lane per face builds its link.
#include "keys.h"
#include "stdmath.h"

namespace fl {
namespace {

__device__ inline bool positive_finite(double v) { return finite(v) && v > 0; }
__device__ inline bool is_normal(double v) { return finite(v) && fabs(v) >= 2.2250738585072014e-308; }

struct View {
  const fl_int3* blocks;
  uint64_t nb;
  const fl_center_distance* centers;
  uint64_t ncenters;
  const fl_frozen_cell* cells;
  uint64_t ncells;
  const fl_frozen_face* faces;
  uint64_t nfaces;
  double h, step, density, sigma;
};

__device__ inline uint64_t find_cell(const View& v, fl_int3 k) {
  return find_sorted(v.cells, v.ncells, k, [](const fl_frozen_cell& c) { return c.cell_index; });
}

__device__ inline int face_slot(const View& v, int axis, fl_int3 l, uint64_t& slot) {
  return sorted_face_slot(v.blocks, v.nb, axis, l, slot);
}

__global__ void k_check_centers(View v, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= v.ncenters) return;
  if (!finite(v.centers[i].signed_distance) ||
      (i > 0 && !key_less(v.centers[i - 1].cell_index, v.centers[i].cell_index)))
    fail(err, 0, i, kInvalidArgument);
}

__global__ void k_check_frozen_cells(View v, double full_volume, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= v.ncells) return;
  const fl_frozen_cell c = v.cells[i];
  if (!positive_finite(c.fluid_volume) || c.fluid_volume > full_volume || !finite(c.center_distance) ||
      c.center_distance >= 0 || !finite(c.curvature) || (i > 0 && !key_less(v.cells[i - 1].cell_index, c.cell_index)))
    fail(err, 1, i, kInvalidArgument);
}

__global__ void k_check_frozen_faces(View v, double full_area, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= v.nfaces) return;
  const fl_frozen_face f = v.faces[i];
  uint64_t slot = 0;
  int code = face_slot(v, f.axis, f.lattice, slot);
  if (code) {
    fail(err, 2, i, code);
    return;
  }
  if (slot != f.face_slot || (i > 0 && f.face_slot <= v.faces[i - 1].face_slot) || !finite(f.open_area) ||
      f.open_area < 0 || f.open_area > full_area || f.blocked_area != full_area - f.open_area ||
      !finite(f.wall_normal_velocity) ||
      (f.has_prescribed_wall_flux &&
       (!finite(f.prescribed_wall_flux) || (f.blocked_area == 0 && f.prescribed_wall_flux != 0)))) {
    fail(err, 2, i, kInvalidArgument);
    return;
  }
  fl_int3 tail{};
  const bool tail_ok = offset(f.lattice, f.axis, -1, tail);
  const bool has_head = find_cell(v, f.lattice) != v.ncells;
  const bool has_tail = tail_ok && find_cell(v, tail) != v.ncells;
  if (!has_head && !has_tail) fail(err, 2, i, kFailedPrecondition);
  else if (!tail_ok && f.open_area > 0) fail(err, 2, i, kOutOfRange);
}

// Every face of every cell is in the face list (ordinal = 6 cell + 2 axis + side).
__global__ void k_check_incident(View v, uint64_t* err) {
  const uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (t >= 6 * v.ncells) return;
  const uint64_t i = t / 6;
  const int axis = int(t % 6) / 2, side = int(t % 2);
  fl_int3 shifted{};
  if (!offset(v.cells[i].cell_index, axis, side, shifted)) {
    fail(err, 3, t, kOutOfRange);
    return;
  }
  uint64_t slot = 0;
  const int code = face_slot(v, axis, shifted, slot);
  if (code) {
    fail(err, 3, t, code);
    return;
  }
  uint64_t lo = 0, hi = v.nfaces;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (v.faces[mid].face_slot < slot) lo = mid + 1;
    else hi = mid;
  }
  if (lo == v.nfaces || v.faces[lo].face_slot != slot) fail(err, 3, t, kFailedPrecondition);
}

struct Fields {
  const fl_cell_fields* cells;
  const double* predicted;
  double ambient;
  const fl_optional_flux* wetted;
  uint64_t nwetted;
  const fl_air_pressure* air;
  uint64_t nair;
  const fl_internal_interface* interfaces;
  uint64_t ninterfaces;
};

__global__ void k_check_interfaces(View v, Fields F, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= F.ninterfaces) return;
  const fl_internal_interface c = F.interfaces[i];
  if (c.face_ordinal >= v.nfaces || (i > 0 && F.interfaces[i - 1].face_ordinal >= c.face_ordinal) ||
      !finite(c.tail_fraction) || c.tail_fraction < 0.0 || c.tail_fraction > 1.0 || !finite(c.surface_tension) ||
      c.surface_tension < 0.0 || !finite(c.curvature_tail_to_head) || !finite(c.normal_viscous_jump_tail_to_head))
    fail(err, 4, i, kInvalidArgument);
}

__global__ void k_check_air(View v, Fields F, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= F.nair) return;
  const fl_air_pressure a = F.air[i];
  if (!finite(a.gas_gauge_pressure) || (i > 0 && !key_less(F.air[i - 1].cell_index, a.cell_index))) {
    fail(err, 5, i, kInvalidArgument);
    return;
  }
  const uint64_t s =
      find_sorted(v.centers, v.ncenters, a.cell_index, [](const fl_center_distance& c) { return c.cell_index; });
  if (s == v.ncenters) fail(err, 5, i, kFailedPrecondition);
  else if (v.centers[s].signed_distance < 0.0 || find_cell(v, a.cell_index) != v.ncells)
    fail(err, 5, i, kInvalidArgument);
}

__global__ void k_link_cells(View v, Fields F, fl_pressure_cell* out, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= v.ncells) return;
  const fl_cell_fields f = F.cells[i];
  if (!finite(f.target_divergence) || !finite(f.previous_pressure) || isnan(f.bulk_modulus) || f.bulk_modulus <= 0 ||
      (f.has_density && !positive_finite(f.density))) {
    fail(err, 6, i, kInvalidArgument);
    return;
  }
  fl_pressure_cell c{};
  c.cell_index = v.cells[i].cell_index;
  c.pad = 0;
  c.fluid_volume = v.cells[i].fluid_volume;
  c.target_divergence = f.target_divergence;
  c.bulk_modulus = f.bulk_modulus;
  c.previous_pressure = f.previous_pressure;
  c.integrated_inward_wall_flux = 0.0;
  out[i] = c;
}

// std::lerp and std::midpoint as libstdc++ defines them for double.
// InternalLiquidCapillaryStepBound3D (liquid_boundary_geometry_3d.cc:1669-1701).
__device__ inline int capillary_bound(double td, double hd, double h, double sigma, double& out) {
  if (!positive_finite(td) || !positive_finite(hd) || !positive_finite(h) || !finite(sigma) || sigma < 0.0)
    return kInvalidArgument;
  if (sigma == 0.0) {
    out = __builtin_huge_val();
    return kOk;
  }
  const double larger = td < hd ? hd : td;
  const double ratio = (hd < td ? hd : td) / larger;
  int de = 0, we = 0, te = 0;
  const double sd = frexp(larger, &de) * (1.0 + ratio);
  const double sw = frexp(h, &we);
  const double st = frexp(sigma, &te);
  const double square = sd * sw * sw * sw / (4.0 * 3.141592653589793 * st);
  const int e = de + 3 * we - te;
  const double bound = scalbn(sqrt(scalbn(square, e % 2)), e / 2);
  if (!positive_finite(bound)) return kOutOfRange;
  out = bound;
  return kOk;
}

__device__ inline void set_link(fl_pressure_face& f, int kind, double lift) {
  f.link_kind = kind;
  f.lift = lift;
}

// LinksCore's staged link of one face and, for an open face with air on one side, its free-surface
// geometry (the kLink query).
__global__ void k_link_faces(View v, Fields F, fl_pressure_face* out, unsigned long long* lower, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= v.nfaces) return;
  const fl_frozen_face face = v.faces[i];
  const double pv = F.predicted[i];
  if (!finite(pv)) {
    fail(err, 7, i, kInvalidArgument);
    return;
  }
  fl_pressure_face link{};
  link.axis = face.axis;
  link.lattice = face.lattice;
  link.open_area = face.open_area;
  link.inverse_density = 1 / v.density;
  link.pressure_gradient_distance = v.h;
  link.predicted_normal_velocity = pv;
  link.wall_normal_velocity = face.wall_normal_velocity;
  set_link(link, 0, 0.0);
  link.has_prescribed_wall_flux = face.has_prescribed_wall_flux;
  link.prescribed_wall_flux = face.has_prescribed_wall_flux ? face.prescribed_wall_flux : 0.0;
  const fl_int3 head = face.lattice;
  fl_int3 tail{};
  const bool tail_ok = offset(head, face.axis, -1, tail);
  const uint64_t head_found = find_cell(v, head);
  const uint64_t tail_found = tail_ok ? find_cell(v, tail) : v.ncells;
  const bool has_head = head_found != v.ncells, has_tail = tail_found != v.ncells;
  if (F.nwetted) {
    const bool receiver = face.is_low_x_piston_face && face.axis == 0 && face.open_area == 0 && has_head && !has_tail &&
                          face.has_prescribed_wall_flux;
    const fl_optional_flux flux = F.wetted[i];
    if (receiver != bool(flux.has_value)) {
      fail(err, 7, i, kInvalidArgument);
      return;
    }
    if (flux.has_value) {
      const double w = face.wall_normal_velocity;
      const bool wrong = (flux.value > 0 && w <= 0) || (flux.value < 0 && w >= 0);
      if (!finite(flux.value) || wrong) {
        fail(err, 7, i, kInvalidArgument);
        return;
      }
      link.has_prescribed_wall_flux = 1;
      link.prescribed_wall_flux = flux.value;
    }
  }
  uint64_t lo = 0, hi = F.ninterfaces;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (F.interfaces[mid].face_ordinal < i) lo = mid + 1;
    else hi = mid;
  }
  if (lo < F.ninterfaces && F.interfaces[lo].face_ordinal == i) {
    const fl_internal_interface c = F.interfaces[lo];
    if (!has_head || !has_tail || face.open_area <= 0.0 || !F.cells[head_found].has_density ||
        !F.cells[tail_found].has_density) {
      fail(err, 7, i, kInvalidArgument);
      return;
    }
    const double tail_density = F.cells[tail_found].density, head_density = F.cells[head_found].density;
    double bound = 0;
    const int code = capillary_bound(tail_density, head_density, v.h, c.surface_tension, bound);
    if (code) {
      fail(err, 7, i, code);
      return;
    }
    if (v.step > bound) {
      fail(err, 7, i, kAborted);
      return;
    }
    const double inverse_density = 1.0 / lerp_std(head_density, tail_density, c.tail_fraction);
    const double jump = fma(-c.surface_tension, c.curvature_tail_to_head, c.normal_viscous_jump_tail_to_head);
    if (!is_normal(inverse_density) || !finite(jump)) {
      fail(err, 7, i, kOutOfRange);
      return;
    }
    link.inverse_density = inverse_density;
    set_link(link, 1, jump);
    out[i] = link;
    return;
  }
  if (!tail_ok) {
    if (face.open_area > 0) fail(err, 7, i, kOutOfRange);
    else out[i] = link;
    return;
  }
  if (!has_head && !has_tail) {
    fail(err, 7, i, kFailedPrecondition);
    return;
  }
  if (face.open_area > 0) {
    const bool hd = has_head && F.cells[head_found].has_density;
    const bool td = has_tail && F.cells[tail_found].has_density;
    if (hd || td) {
      const double head_density = hd ? F.cells[head_found].density : v.density;
      const double tail_density = td ? F.cells[tail_found].density : v.density;
      const double face_density =
          has_head && has_tail ? midpoint_std(head_density, tail_density) : (has_head ? head_density : tail_density);
      link.inverse_density = 1.0 / face_density;
      if (!is_normal(link.inverse_density)) {
        fail(err, 7, i, kOutOfRange);
        return;
      }
    }
  }
  if (has_head && has_tail) {
    set_link(link, 1, 0.0);
  } else if (face.open_area > 0) {
    const fl_int3 air_cell = has_head ? tail : head;
    const uint64_t s =
        find_sorted(v.centers, v.ncenters, air_cell, [](const fl_center_distance& c) { return c.cell_index; });
    if (s == v.ncenters || v.centers[s].signed_distance < 0) {
      fail(err, 7, i, kFailedPrecondition);
      return;
    }
    const double air_distance = v.centers[s].signed_distance;
    const fl_frozen_cell liquid = v.cells[has_head ? head_found : tail_found];
    double ambient = F.ambient;
    const uint64_t a = find_sorted(F.air, F.nair, air_cell, [](const fl_air_pressure& x) { return x.cell_index; });
    if (a != F.nair) ambient = F.air[a].gas_gauge_pressure;
    const double magnitude = -liquid.center_distance < air_distance ? air_distance : -liquid.center_distance;
    const double liquid_scaled = -liquid.center_distance / magnitude;
    const double theta = liquid_scaled / (liquid_scaled + air_distance / magnitude);
    const bool regularized = theta < .05;
    const double clamped = theta < .05 ? .05 : (1.0 < theta ? 1.0 : theta);
    const double distance = clamped * v.h;
    const double lift = v.sigma * liquid.curvature;
    const double exterior = ambient + lift;
    if ((v.sigma > 0 && liquid.curvature != 0 && lift == 0) || !(lift == 0.0 || is_normal(lift)) || !finite(exterior) ||
        !is_normal(distance)) {
      fail(err, 7, i, kOutOfRange);
      return;
    }
    if (regularized) atomicAdd(lower, 1ull);
    link.pressure_gradient_distance = distance;
    set_link(link, 2, exterior);
  }
  out[i] = link;
}

}  // namespace

int build_links_device(fl_ctx* ctx, const LinkDevice& d, fl_pressure_cell* out_cells, fl_pressure_face* out_faces,
                       uint64_t* d_lower, bool validate_only) {
  hipStream_t s = ctx->stream;
  const View v{d.blocks, d.nb, d.centers, d.ncenters, d.cells, d.ncells, d.faces, d.nfaces,
               d.h, d.step, d.density, d.sigma};
  if (validate_only) {
    const double full_area = d.h * d.h, full_volume = full_area * d.h;
    if (v.ncenters) k_check_centers<<<grid_for(v.ncenters), kThreads, 0, s>>>(v, ctx->err);
    if (v.ncells) k_check_frozen_cells<<<grid_for(v.ncells), kThreads, 0, s>>>(v, full_volume, ctx->err);
    if (v.nfaces) k_check_frozen_faces<<<grid_for(v.nfaces), kThreads, 0, s>>>(v, full_area, ctx->err);
    if (v.ncells) k_check_incident<<<grid_for(6 * v.ncells), kThreads, 0, s>>>(v, ctx->err);
    FL_HIP(hipGetLastError());
    return FL_OK;
  }
  const Fields F{d.fields, d.predicted, d.ambient, d.wetted, d.nwetted, d.air, d.nair, d.interfaces, d.ninterfaces};
  if (F.ninterfaces) k_check_interfaces<<<grid_for(F.ninterfaces), kThreads, 0, s>>>(v, F, ctx->err);
  if (F.nair) k_check_air<<<grid_for(F.nair), kThreads, 0, s>>>(v, F, ctx->err);
  if (v.ncells) k_link_cells<<<grid_for(v.ncells), kThreads, 0, s>>>(v, F, out_cells, ctx->err);
  if (v.nfaces)
    k_link_faces<<<grid_for(v.nfaces), kThreads, 0, s>>>(v, F, out_faces, reinterpret_cast<unsigned long long*>(d_lower),
                                                        ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl
