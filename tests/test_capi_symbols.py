"""The C-ABI library loads and exports every entry point include/fluidity_hip.h declares (no GPU needed)."""
import ctypes
import os
import re

from fluidity_amd import hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fluidity_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|void|size_t)\s+(fl_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_entry_point():
    lib = H.lib()
    names = declared_symbols()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(H.EXPORTS) <= set(names)


def test_record_layouts_match_the_header():
    from fluidity_amd import records as R

    assert R.PARTICLE.itemsize == 152 and R.GATHER_RESULT.itemsize == 192
    assert R.FACE_QUANTA.itemsize == 24 and R.SORT_ENTRY.itemsize == 24
    assert R.PRESSURE_CELL.itemsize == 56 and R.PRESSURE_FACE.itemsize == 80
    assert R.PRESSURE_OPTIONS.itemsize == 56 and R.PRESSURE_DIAGNOSTICS.itemsize == 96
    assert R.BOUNDARY_PARTICLE.itemsize == 168 and R.FACE_TAG.itemsize == 16 and R.COMPATIBLE_SAMPLE.itemsize == 72
    assert R.WALL_SAMPLE.itemsize == 64 and R.BOUNDARY_GATHER.itemsize == 224
    assert R.BLOCK_STATE.itemsize == 32 and R.CELL_STATE.itemsize == 24 and R.FACE_STATE.itemsize == 56
    assert R.LIQUID_PARTICLE.itemsize == 328 and R.UPDATE_OPTIONS.itemsize == 160
    assert R.VISCOSITY_CELL.itemsize == 24 and R.STRAIN_ROW.itemsize == 168 and R.SOLID_CORNER.itemsize == 48


def test_record_offsets_match_the_c_compiler(tmp_path):
    """sizeof/offsetof of the topology and boundary records as gcc lays out the header."""
    import subprocess

    from fluidity_amd import records as R

    checks = {
        "fl_block_state": (R.BLOCK_STATE, {"generation": "generation", "activation_age": "activation_age",
                                           "inactive_age": "inactive_age"}),
        "fl_cell_state": (R.CELL_STATE, {"occupancy": "occupancy", "previous_pressure": "previous_pressure"}),
        "fl_face_state": (R.FACE_STATE, {"axis": "axis", "lattice": "lattice", "face_slot": "face_slot",
                                         "carrier_velocity": "carrier_velocity", "validity": "validity",
                                         "side_tag": "side_tag", "structurally_eligible": "structurally_eligible"}),
        "fl_liquid_particle": (R.LIQUID_PARTICLE, {n: n for n in R.LIQUID_PARTICLE.names}),
        "fl_particle_update_options": (R.UPDATE_OPTIONS, {n: n for n in R.UPDATE_OPTIONS.names}),
        "fl_obstacle": (R.OBSTACLE, {n: n for n in R.OBSTACLE.names}),
        "fl_frozen_face": (R.FROZEN_FACE, {"face_slot": "face_slot", "open_area": "open_area",
                                           "prescribed_wall_flux": "prescribed_wall_flux",
                                           "has_prescribed_wall_flux": "has_prescribed_wall_flux",
                                           "is_low_x_piston_face": "is_low_x_piston_face"}),
        "fl_frozen_cell": (R.FROZEN_CELL, {n: n for n in R.FROZEN_CELL.names}),
        "fl_cell_fields": (R.CELL_FIELDS, {n: n for n in R.CELL_FIELDS.names}),
        "fl_center_distance": (R.CENTER_DISTANCE, {n: n for n in R.CENTER_DISTANCE.names}),
        "fl_boundary_gather": (R.BOUNDARY_GATHER, {"affine_enabled": "affine_enabled",
                                                   "low_mass_fallback": "low_mass_fallback"}),
        "fl_viscosity_cell": (R.VISCOSITY_CELL, {n: n for n in R.VISCOSITY_CELL.names}),
        "fl_strain_row": (R.STRAIN_ROW, {n: n for n in R.STRAIN_ROW.names}),
        "fl_solid_corner": (R.SOLID_CORNER, {n: n for n in R.SOLID_CORNER.names}),
    }
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{ROOT}/include/fluidity_hip.h"', "int main(void) {"]
    for struct, (_, fields) in checks.items():
        lines.append(f'  printf("{struct} %zu\\n", sizeof({struct}));')
        for cfield in fields.values():
            lines.append(f'  printf("{struct}.{cfield} %zu\\n", offsetof({struct}, {cfield}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                           check=True).stdout.splitlines())
    for struct, (dtype, fields) in checks.items():
        assert int(got[struct]) == dtype.itemsize, struct
        for name, cfield in fields.items():
            assert int(got[f"{struct}.{cfield}"]) == dtype.fields[name][1], (struct, name)


def test_ctypes_structures_match_the_c_compiler(tmp_path):
    """The binding's argument structures (frozen geometry inputs, obstacle, frozen view, pressure fields)
    have gcc's size and member offsets."""
    import ctypes as C
    import subprocess

    from fluidity_amd import hip as H

    checks = {"fl_freeze_inputs": (H.FreezeInputs, {"origin": "geometry", "obstacle": "obstacle",
                                                    "step_duration": "step_duration",
                                                    "has_exact_curvature": "has_exact_curvature",
                                                    "max_cells": "max_cells", "max_faces": "max_faces",
                                                    "solid_corners": "solid_corners",
                                                    "face_velocity_count": "face_velocity_count",
                                                    "has_sampled_solid": "has_sampled_solid"}),
              "fl_obstacle": (H.Obstacle, {"unit_normal": "unit_normal", "coordinate": "coordinate",
                                           "half_width": "half_width", "wall_velocity": "wall_velocity"}),
              "fl_frozen_view": (H.FrozenView, {"origin": "geometry", "center_distances": "center_distances",
                                                "capillary_step_bound": "capillary_step_bound"}),
              "fl_pressure_fields": (H.PressureFields, {"ambient_pressure": "ambient_pressure",
                                                        "interface_count": "interface_count"})}
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{ROOT}/include/fluidity_hip.h"', "int main(void) {"]
    for struct, (_, fields) in checks.items():
        lines.append(f'  printf("{struct} %zu\\n", sizeof({struct}));')
        for cfield in fields.values():
            lines.append(f'  printf("{struct}.{cfield} %zu\\n", offsetof({struct}, {cfield}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "structs.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "structs"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                           check=True).stdout.splitlines())
    for struct, (cls, fields) in checks.items():
        assert int(got[struct]) == C.sizeof(cls), struct
        for name, cfield in fields.items():
            assert int(got[f"{struct}.{cfield}"]) == getattr(cls, name).offset, (struct, name)


def test_create_without_device_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        return
    handle = ctypes.c_void_p()
    cap = H.Capacities(16, 16, 64, 16, 16)
    assert H.lib().fl_create(ctypes.byref(cap), 0, ctypes.byref(handle)) == H.FAILED_PRECONDITION
