"""The reference's own meshes through the oracle's voxelizer, pinned by what
the geometry fixes: both meshes are closed (every edge shared by exactly two
triangles), and on the NACA0012 driver's grid (a single cell layer inside the
span) the solid cell count times the cell area equals the area of the
section the mesh's top cap triangulates, to 0.1 %.  The GPU voxelizer is
compared with the oracle on the same grids in test_gpu_stl_fixtures."""
import math
import os

import numpy as np

import oracle as O
from tests.stl_shapes import GOLDEN_STL, cap_area, read_binary_stl


def test_fixture_meshes_are_closed():
    for name, count in (("NACA0012.STL", 12996), ("vane.STL", 3584)):
        tri = read_binary_stl(os.path.join(GOLDEN_STL, name))
        assert tri.shape == (count, 9)
        bad, counts = O.check_stl_edges(tri)
        assert bad == 0 and np.all(counts == 2), name


def test_naca0012_section_area():
    tri = read_binary_stl(os.path.join(GOLDEN_STL, "NACA0012.STL"))
    r = np.float32(-10.0 * math.pi / 180.0)
    v = tri.reshape(-1, 3).copy()
    x = v[:, 0] * np.cos(r) - v[:, 1] * np.sin(r)
    y = v[:, 0] * np.sin(r) + v[:, 1] * np.cos(r)
    v[:, 0], v[:, 1] = x, y
    m = O.markers_inside_stl(v.reshape(-1, 9), (0, 0, 0), (-3000.0, -5000.0, 0.0), 1.4, 7143, 7143, 1)
    cells = int(m.sum())
    area = cap_area(tri) / 1.4 ** 2
    assert abs(cells / area - 1) < 1e-3, (cells, area)


def test_vane_section_area():
    tri = read_binary_stl(os.path.join(GOLDEN_STL, "vane.STL"))
    m = O.markers_inside_stl(tri, (5.0, 10.0, 0.0), (0, 0, 0), 0.02, 2500, 2000, 1)
    area = cap_area(tri) / 0.02 ** 2
    assert abs(int(m.sum()) / area - 1) < 1e-3
