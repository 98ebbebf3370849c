"""Test meshes: a reader for binary STL fixtures and closed boxes with a tilted
top face (hit heights on exact multiples of the cell size)."""
import os

import numpy as np

GOLDEN_STL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stl")


def read_binary_stl(path: str) -> np.ndarray:
    """Triangles (n, 9) float32 of a binary STL: 80-byte header, uint32 count,
    then per triangle 12 float32 (normal, three vertices) and a uint16."""
    data = open(path, "rb").read()
    n = int(np.frombuffer(data[80:84], np.uint32)[0])
    assert len(data) == 84 + 50 * n, (path, len(data), n)
    rec = np.dtype([("normal", "<f4", 3), ("v", "<f4", 9), ("attr", "<u2")])
    return np.frombuffer(data[84:], rec, count=n)["v"].astype(np.float32)


def tilted_box(x0, x1, y0, y1, zb, a, sx, sy) -> np.ndarray:
    """Closed 12-triangle box over [x0, x1] x [y0, y1]: flat bottom at z = zb,
    top face on the plane z = a + sx x + sy y."""
    def zt(x, y):
        return a + sx * x + sy * y

    b = [(x0, y0, zb), (x1, y0, zb), (x1, y1, zb), (x0, y1, zb)]
    t = [(x, y, zt(x, y)) for (x, y, _) in b]
    quads = [(b[0], b[3], b[2], b[1]), (t[0], t[1], t[2], t[3]), (b[0], b[1], t[1], t[0]),
             (b[1], b[2], t[2], t[1]), (b[2], b[3], t[3], t[2]), (b[3], b[0], t[0], t[3])]
    tris = []
    for q in quads:
        tris.append(q[0] + q[1] + q[2])
        tris.append(q[0] + q[2] + q[3])
    return np.array(tris, np.float32)


def cap_area(tri: np.ndarray) -> float:
    """Area (FP64) of the triangles of an extruded mesh that lie in its top
    plane: the cross-section a z ray through the span sees."""
    t = np.asarray(tri, np.float64).reshape(-1, 3, 3)
    zmax = t[:, :, 2].max()
    top = t[np.all(np.abs(t[:, :, 2] - zmax) <= 1e-6 * max(1.0, abs(zmax)), axis=1)]
    cr = np.cross(top[:, 1] - top[:, 0], top[:, 2] - top[:, 0])
    return float(np.abs(cr[:, 2]).sum() / 2)
