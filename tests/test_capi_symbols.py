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


def test_create_without_device_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        return
    handle = ctypes.c_void_p()
    cap = H.Capacities(16, 16, 64, 16, 16)
    assert H.lib().fl_create(ctypes.byref(cap), 0, ctypes.byref(handle)) == H.FAILED_PRECONDITION
