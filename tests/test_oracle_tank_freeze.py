"""Pins the oracle's open-tank freeze (flo_freeze_geometry with an AxisAlignedLiquidTank3D).

Known answers are the reference's (native/tests/liquid_boundary_geometry_3d_cases.h:90-302, cases
OpenTankClipsVolumeAndUsesPistonOnlyFlux and OpenTankPreservesNarrowAndLiquidClippedMeasures) on its
2x2x2-block grid with the candidate (-2, -2, -2): the tank-clipped volume 105/256 with the open high-x
and high-y apertures 35/64 and 15/32, the piston's flux 35/512 on the low-x face only (and its sign for a
retreating piston), a closed high-x wall, the liquid-clipped volume 75/256, a nearly dry cell
(15/32 2^-30), a tank 2^-45 wide (25/64 2^-45), the unresolvable local cut (OutOfRange) and the tank
step refusals (InvalidArgument, FailedPrecondition, OutOfRange). Closed form: a full cell inside a box
tank has the box's volume, in a scaled-exponent product.
"""
import numpy as np
import pytest

import pyoracle as O
from test_oracle_freeze import problem, sample

INF = float("inf")


def tank(start_low_x=0.25, high_x=2.0, floor_y=0.125, low_z=0.125, high_z=0.75, piston=0.125):
    return (O.OBSTACLE_TANK, (0.0, 0.0, 0.0), 0.0, 0.0, (0.0, 0.0, 0.0),
            (start_low_x, high_x, floor_y, low_z, high_z, piston))


def only_centre_liquid():
    centers, corners = sample(lambda x, y, z: -1.0)
    for i, c in enumerate(centers["cell_index"]):
        if (int(c["x"]), int(c["y"]), int(c["z"])) != (-2, -2, -2):
            centers["signed_distance"][i] = 1.0
    return centers, corners


def head(face):
    return tuple(int(v) for v in face["face_index"]["lattice"]) == (-2, -2, -2)


def aperture(face):
    if head(face):
        return 0.0
    return {0: 0.546875, 1: 0.46875, 2: 0.0}[int(face["face_index"]["axis"])]


def test_tank_clips_volume_and_uses_piston_only_flux():
    centers, corners = only_centre_liquid()
    s, cells, faces, _ = O.freeze_geometry(problem(centers=centers, corners=corners, obstacle=tank(), step=0.5,
                                                   density=1.0))
    assert s == O.OK and len(cells) == 1 and len(faces) == 6
    assert abs(cells[0]["fluid_volume"] - 0.41015625) <= 2e-14
    low_x = high_x = None
    for k, f in enumerate(faces):
        assert f["open_area"] == aperture(f) and f["blocked_area"] == 1 - aperture(f)
        assert f["has_prescribed_wall_flux"] == 1
        piston = head(f) and int(f["face_index"]["axis"]) == 0
        assert f["prescribed_wall_flux"] == (0.068359375 if piston else 0.0)
        assert f["wall_normal_velocity"] == (0.125 if piston else 0.0)
        assert f["is_low_x_piston_face"] == (1 if piston else 0)
        if int(f["face_index"]["axis"]) == 0:
            if head(f):
                low_x = k
            else:
                high_x = k
    s, _, faces, _ = O.freeze_geometry(problem(centers=centers, corners=corners, obstacle=tank(piston=-0.125),
                                               step=0.5, density=1.0))
    assert s == O.OK and faces[low_x]["prescribed_wall_flux"] == -0.068359375
    s, _, faces, _ = O.freeze_geometry(problem(centers=centers, corners=corners, obstacle=tank(high_x=1.0),
                                               step=0.5, density=1.0))
    assert s == O.OK
    assert faces[high_x]["open_area"] == 0 and faces[high_x]["wall_normal_velocity"] == 0
    assert faces[high_x]["prescribed_wall_flux"] == 0 and faces[high_x]["has_prescribed_wall_flux"] == 1


def test_tank_preserves_narrow_and_liquid_clipped_measures():
    centers, corners = sample(lambda x, y, z: y - 0.75)
    s, cells, _, _ = O.freeze_geometry(problem(centers=centers, corners=corners, obstacle=tank()))
    assert s == O.OK and len(cells) == 1 and abs(cells[0]["fluid_volume"] - 0.29296875) <= 2e-14
    thin_c, thin_k = sample(lambda x, y, z: -1.0 if y == 0.5 else y - (0.125 + 2.0 ** -30))
    s, cells, faces, _ = O.freeze_geometry(problem(centers=thin_c, corners=thin_k, obstacle=tank()))
    assert s == O.OK and len(cells) == 1 and abs(cells[0]["fluid_volume"] - 0.46875 * 2.0 ** -30) <= 2e-17
    piston = [f for f in faces if int(f["face_index"]["axis"]) == 0 and int(f["face_index"]["lattice"]["x"]) == -2]
    assert len(piston) == 1 and piston[0]["prescribed_wall_flux"] == 0.068359375
    narrow = tank(start_low_x=0.0, high_x=2.0 ** -45, piston=0.0)
    s, cells, _, _ = O.freeze_geometry(problem(centers=centers, corners=corners, obstacle=narrow))
    assert s == O.OK and len(cells) == 1 and abs(cells[0]["fluid_volume"] - 0.390625 * 2.0 ** -45) <= 1e-27
    collapsed = (-2.0 ** -60, 0.0, 0.125, 0.125, 0.75, 0.0)
    assert O.freeze_geometry(problem(centers=centers, corners=corners, candidates=((-3, -2, -2),),
                                     obstacle=tank(*collapsed)))[0] == O.OUT_OF_RANGE
    assert O.freeze_geometry(problem(centers=centers, corners=corners,
                                     obstacle=tank(start_low_x=0.0, high_x=0.0, piston=0.0)))[0] == O.INVALID_ARGUMENT
    assert O.freeze_geometry(problem(centers=centers, corners=corners,
                                     obstacle=tank(start_low_x=0.0, high_x=1.0, high_z=float("nan"), piston=0.0))
                             )[0] == O.INVALID_ARGUMENT
    assert O.freeze_geometry(problem(centers=centers, corners=corners, step=1.0,
                                     obstacle=tank(start_low_x=0.0, high_x=1.0, piston=1.0)))[0] == O.FAILED_PRECONDITION
    assert O.freeze_geometry(problem(centers=centers, corners=corners,
                                     obstacle=tank(start_low_x=0.0, high_x=1.0, piston=INF)))[0] == O.INVALID_ARGUMENT
    big = -np.finfo(np.float64).max
    assert O.freeze_geometry(problem(centers=centers, corners=corners, step=2.0,
                                     obstacle=tank(start_low_x=0.0, high_x=1.0, piston=big)))[0] == O.OUT_OF_RANGE


@pytest.mark.parametrize("box", [(0.25, 0.875, 0.0625, 0.125, 0.75), (-0.5, 0.5, 0.5, 0.3, 0.35),
                                 (0.1, 0.2, 0.95, 0.0, 1.0)])
def test_full_cell_in_box_tank_closed_form(box):
    lo_x, hi_x, floor_y, lo_z, hi_z = box
    centers, corners = sample(lambda x, y, z: -1.0)
    s, cells, _, _ = O.freeze_geometry(problem(centers=centers, corners=corners,
                                               obstacle=tank(lo_x, hi_x, floor_y, lo_z, hi_z, 0.0)))
    ex = max(0.0, min(hi_x, 1.0) - max(lo_x, 0.0))
    ey = max(0.0, 1.0 - max(floor_y, 0.0))
    ez = max(0.0, min(hi_z, 1.0) - max(lo_z, 0.0))
    want = ex * ey * ez
    assert s == O.OK
    got = cells[0]["fluid_volume"] if len(cells) else 0.0
    assert abs(got - want) <= 1e-15
