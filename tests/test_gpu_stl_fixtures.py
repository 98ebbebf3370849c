"""The voxelizer on the reference's own meshes and on tie-sensitive meshes,
against the oracle, bit for bit.

Fixtures: NACA0012/NACA0012.STL (12,996 triangles) and cascade/vane.STL
(3,584 triangles), copied from the reference as data.  Each is read from its
file by the library (readSTL), checked for open edges (checkSTLEdges, counts
equal to the oracle's) and voxelized (applyMarkersInsideSTL) on the grid its
driver builds -- NACA0012/NACA0012Main.cu:161-184 (rotated by the angle of
attack, 7143 x 7143 x 1 cells of 1.4 mm from (-3000, -5000)) and
cascade/cascadeMain.cu:154-178 (eleven vanes at pitch s = 10 OR-ed into one
2500 x 2000 x 1 mask of 0.02 mm) -- and on a second, three-dimensional grid at
another resolution with offset grid and mesh origins.
"""
import math
import os

import numpy as np
import pytest

from tests.stl_shapes import GOLDEN_STL, cap_area, read_binary_stl, tilted_box

pytestmark = pytest.mark.gpu


def _voxelize(P, stl, info, inside=True):
    n = info.cellCountX * info.cellCountY * info.cellCountZ
    m = P.BoolArray(n)
    P.applyMarkersInsideSTL(m, stl, inside, info)
    return m.get()


def _read_and_check(P, O, name):
    path = os.path.join(GOLDEN_STL, name)
    tri = read_binary_stl(path)
    stl = P.readSTL(path)
    assert stl.triangleCount == tri.shape[0]
    assert np.array_equal(stl.triangles(), tri)
    bad, counts = P.checkSTLEdges(stl)
    obad, ocounts = O.check_stl_edges(tri)
    print(f"  {name}: {tri.shape[0]} triangles, open-edge triangles {bad} (oracle {obad})", flush=True)
    assert bad == obad and np.array_equal(counts, ocounts)
    return stl, tri


def test_naca0012_fixture(psl, oracle_mod):
    P, O = psl, oracle_mod
    print("[stl fixture] NACA0012.STL", flush=True)
    stl, tri = _read_and_check(P, O, "NACA0012.STL")
    assert tri.shape[0] == 12996
    # rotateSTLAlongZ by minus the angle of attack (NACA0012Main.cu:165-166)
    radians = -10.0 * (math.pi / 180.0)
    P.rotateSTLAlongZ(stl, radians)
    rot = stl.triangles()
    c, s = math.cos(radians), math.sin(radians)
    t = tri.reshape(-1, 3).astype(np.float64)
    want = np.stack([t[:, 0] * c - t[:, 1] * s, t[:, 0] * s + t[:, 1] * c, t[:, 2]], 1).reshape(-1, 9)
    assert np.max(np.abs(rot - want)) < 2e-4  # FP32 of coordinates up to 500 mm
    # the driver's grid: 7143 x 7143 x 1 cells of 1.4 mm from (-3000, -5000, 0)
    n = int(math.ceil(10000.0 / 1.4))
    assert n == 7143
    info = P.InfoStruct(cellCountX=n, cellCountY=n, cellCountZ=1, res=1.4, ox=-3000.0, oy=-5000.0, oz=0.0)
    got = _voxelize(P, stl, info)
    ref = O.markers_inside_stl(rot, (0, 0, 0), (-3000.0, -5000.0, 0.0), 1.4, n, n, 1)
    cells = int(got.sum())
    area = cap_area(tri) / 1.4 ** 2
    print(f"  reference grid: {cells} solid cells (section area {area:.0f} cells)", flush=True)
    assert np.array_equal(got, ref)
    assert abs(cells / area - 1) < 1e-3
    # three-dimensional grid of 4 mm cells with offset grid and mesh origins:
    # the rays now end inside the span, on the caps at z = +-500
    stl.set_origin(3.3, -1.7, 0.9)
    info3 = P.InfoStruct(cellCountX=264, cellCountY=80, cellCountZ=262, res=4.0, ox=-527.1, oy=-161.3, oz=-522.7)
    got3 = _voxelize(P, stl, info3, inside=False)
    ref3 = O.markers_inside_stl(rot, (3.3, -1.7, 0.9), (-527.1, -161.3, -522.7), 4.0, 264, 80, 262, inside_value=0)
    print(f"  3-D grid: {int((got3 == 0).sum())} solid cells", flush=True)
    assert np.array_equal(got3, ref3)
    stl.close()


def test_cascade_vane_fixture(psl, oracle_mod):
    P, O = psl, oracle_mod
    print("[stl fixture] vane.STL", flush=True)
    stl, tri = _read_and_check(P, O, "vane.STL")
    assert tri.shape[0] == 3584
    # cascadeMain.cu:154-178 with its defaults h = 10, s = 10
    h, s, res = 10.0, 10.0, 0.02
    nx, ny = int((5 * s) / res), int((4.0 * h) / res)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=1, res=res)
    total = P.BoolArray(nx * ny)
    ref = np.zeros(nx * ny, np.uint8)
    for shift in range(-5, 6):
        ox = 0.5 * s + shift * s
        stl.set_origin(ox, h, 0.0)
        vane = P.BoolArray(nx * ny)
        P.applyMarkersInsideSTL(vane, stl, True, info)
        P.sumBoolArrays(vane, total, total)
        ref |= O.markers_inside_stl(tri, (ox, h, 0.0), (0, 0, 0), res, nx, ny, 1)
    got = total.get()
    print(f"  cascade grid {nx} x {ny}: {int(got.sum())} solid cells", flush=True)
    assert np.array_equal(got, ref)
    assert got.sum() > 0
    # a three-dimensional grid across the span at 0.25 mm, offset origins
    stl.set_origin(0.37, -0.21, 0.13)
    info3 = P.InfoStruct(cellCountX=72, cellCountY=48, cellCountZ=84, res=0.25, ox=-0.61, oy=-0.93, oz=-10.4)
    got3 = _voxelize(P, stl, info3)
    ref3 = O.markers_inside_stl(tri, (0.37, -0.21, 0.13), (-0.61, -0.93, -10.4), 0.25, 72, 48, 84)
    print(f"  3-D grid: {int(got3.sum())} solid cells", flush=True)
    assert np.array_equal(got3, ref3)
    stl.close()


@pytest.mark.parametrize("sx,sy", [(1.0, 0.5), (0.25, 0.75), (2.0, 1.0), (1 / 3, 2 / 3)])
def test_tilted_faces_on_cell_multiples(psl, oracle_mod, sx, sy):
    """Boxes whose tilted top face puts hit heights on exact multiples of a
    0.1 cell (include/STLFunctions.h:247-275, 297-320: z = a - (n . d) / n_z,
    then k = ceil(z / res)).  Whether z lands just below or above a multiple
    depends on each rounding of the hit height, so the masks match only if
    the library rounds every product and sum as the oracle does (no fused
    multiply-adds): an oracle built with contraction (-ffp-contract=fast
    -mfma) differs from the plain one in 88 to 143 cells of these masks."""
    P, O = psl, oracle_mod
    tri = tilted_box(0.05, 6.05, 0.05, 6.05, 0.05, 1.0, sx, sy)
    bad, _ = O.check_stl_edges(tri)
    assert bad == 0
    n = 80
    info = P.InfoStruct(cellCountX=n, cellCountY=n, cellCountZ=n, res=0.1)
    stl = P.STLStruct.from_triangles(tri)
    print(f"[stl ties] slopes ({sx:.3f}, {sy:.3f})", flush=True)
    got = _voxelize(P, stl, info)
    ref = O.markers_inside_stl(tri, (0, 0, 0), (0, 0, 0), 0.1, n, n, n)
    assert np.array_equal(got, ref)
    # the same with offset mesh and grid origins
    stl.set_origin(0.013, -0.027, 0.031)
    info2 = P.InfoStruct(cellCountX=n, cellCountY=n, cellCountZ=n, res=0.1, ox=-0.07, oy=0.03, oz=-0.11)
    got2 = _voxelize(P, stl, info2)
    ref2 = O.markers_inside_stl(tri, (0.013, -0.027, 0.031), (-0.07, 0.03, -0.11), 0.1, n, n, n)
    assert np.array_equal(got2, ref2)
    stl.close()
