"""Seeded random closed meshes through the STL path (include/STLFunctions.h:
checkSTLEdges, applyMarkersInsideSTL) against the oracle, bit-exact.

Each case is the convex hull of 8-40 random points (scipy's Qhull, outward
winding), some of them snapped to the grid's cell centres so that rays pass
through vertices and along edges, voxelized on a grid of random size,
resolution and origin with a random STL origin, for both inside values.  The
edge check must count no open-edge triangles, as the oracle does.
"""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

pytestmark = pytest.mark.gpu
CASES = 16


def _mesh(rng, origin, res, dims):
    npts = int(rng.integers(8, 41))
    lo = np.array(origin) + res * (np.array(dims) * 0.15)
    hi = np.array(origin) + res * (np.array(dims) * 0.85)
    pts = rng.uniform(lo, hi, (npts, 3))
    snap = rng.random(npts) < 0.3  # on cell centres: rays through vertices and edges
    pts[snap] = np.array(origin) + res * np.round((pts[snap] - np.array(origin)) / res)
    hull = ConvexHull(pts)
    tri = []
    centre = pts[hull.vertices].mean(axis=0)
    for s in hull.simplices:
        a, b, c = pts[s]
        if np.dot(np.cross(b - a, c - a), a - centre) < 0:
            b, c = c, b
        tri.append(np.concatenate([a, b, c]))
    return np.asarray(tri, np.float32)


@pytest.mark.parametrize("seed", range(CASES))
def test_random_convex_mesh_matches_oracle(psl, oracle_mod, seed):
    P, O = psl, oracle_mod
    rng = np.random.default_rng(13000 + seed)
    dims = tuple(int(x) for x in rng.integers(4, 40, 3))
    res = float(rng.choice([1.0, 0.5, 0.25, 0.1, 1.4]))
    origin = tuple(float(x) for x in rng.uniform(-5, 5, 3))
    stl_origin = tuple(float(x) for x in rng.uniform(-0.5, 0.5, 3) * res) if rng.integers(0, 2) else (0.0, 0.0, 0.0)
    tri = _mesh(rng, tuple(o - s for o, s in zip(origin, stl_origin)), res, dims)
    stl = P.STLStruct.from_triangles(tri)
    stl.set_origin(*stl_origin)
    bad, counts = P.checkSTLEdges(stl)
    obad, ocounts = O.check_stl_edges(tri)
    assert bad == obad == 0 and np.array_equal(counts, ocounts)
    info = P.InfoStruct(cellCountX=dims[0], cellCountY=dims[1], cellCountZ=dims[2], res=res,
                        ox=origin[0], oy=origin[1], oz=origin[2])
    for inside in (True, False):
        m = P.BoolArray(int(np.prod(dims)))
        P.applyMarkersInsideSTL(m, stl, inside, info)
        got = m.get()
        ref = O.markers_inside_stl(tri, stl_origin, origin, res, *dims, inside_value=int(inside))
        solid = int((got == int(inside)).sum())
        print(f"[random stl] seed {seed}: {tri.shape[0]} triangles, grid {dims} res {res}, inside={inside}: "
              f"{solid} cells marked", flush=True)
        assert np.array_equal(got, ref)
    stl.close()
