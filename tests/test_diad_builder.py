"""DIAD grid builder (buildDIADGrids, include/DIADFunctions.h:640-1046): CPU pins.

The oracle's list-based restatement (sorted cells, binary-search neighbours,
compaction, child generation, interface maps) is compared element for element
with an independent dense-box restatement (tests/diad_dense.py) on tiny
geometries, and checked against properties the construction must have:
children of every interface cell exist on the finer level, no finer cell hangs
under a coarse cell that was kept as plain fluid, the finest level keeps
exactly the fluid cells and their neighbours, and every esotwist neighbour is
the first existing cell along its periodic ray.  The marker rules stand in for
the drivers' getMarkers callbacks (sphere walls, a refinement cylinder and
sphere, a rule that clears refinement, an outlet plane).
"""
import numpy as np
import pytest

import oracle as O
from tests import diad_dense, random_diad
from tests.stl_shapes import tilted_box

import periodicshiftlbm_amd as P
from periodicshiftlbm_amd import cases


def _rules_sphere():
    return P.rules(
        P.marker_rule(P.RULE_BOUNCEBACK, P.SHAPE_SPHERE_LE, p=(0.0, 0.0, 0.0, 2.2)),
        P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_XCYLINDER, p=(0.0, 0.0, 0.0, 3.0, 6.0), levels=[0]),
        P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_SPHERE_LT, p=(0.0, 0.0, 0.0, 3.5), levels=[0]),
        P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_BOX, p=(4.0, -100, -100, 100, 100, 100), value=0, levels=[1]),
        P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(-1, -1, 0, -1, 0, -1), levels=[0]),
    )


def _case(name):
    """(base, levels, span, stls, rules)"""
    if name == "sphere_rules_3":
        base = dict(res=1.0, ox=-7.5, oy=-7.5, oz=-7.5, nu=0.01, dt=1.0, nx=24, ny=16, nz=16)
        return base, 3, 1, [], _rules_sphere()
    if name == "stl_2":
        base = dict(res=1.0, ox=-8.0, oy=-6.5, oz=-6.5, nu=0.02, dt=0.5, nx=22, ny=14, nz=14)
        tri = cases.icosphere(2, 2.6, (-0.3, 0.2, -0.1))
        r = P.rules(P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(-1, -1, 0, -1, 0, -1), levels=[0]))
        return base, 2, 2, [(tri, (0.3, -0.2, 0.1))], r
    if name == "stl_walls_1":
        base = dict(res=0.5, ox=-5.0, oy=-4.0, oz=-4.0, nu=0.02, dt=1.0, nx=20, ny=16, nz=16)
        tri = cases.icosphere(2, 2.0, (0.0, 0.0, 0.0))
        r = P.rules(P.marker_rule(P.RULE_BOUNCEBACK, P.SHAPE_INDEX_BOX, idx=(0, -1, 0, -1, 0, 0)),
                    P.marker_rule(P.RULE_BOUNCEBACK, P.SHAPE_INDEX_BOX, idx=(0, -1, 0, -1, -1, -1)))
        return base, 1, 0, [(tri, (0.0, 0.0, 0.0))], r
    if name == "ties_3":
        # cells of 0.1 from a non-dyadic origin: rule faces and the tilted top
        # face of an STL box fall on cell positions, where a fused
        # `i * res + ox` or hit height rounds the other way (an oracle built
        # with contraction gets other level sizes here)
        base = dict(res=0.1, ox=-1.2, oy=-1.08, oz=-1.32, nu=0.01, dt=1.0, nx=24, ny=20, nz=20)
        r = P.rules(
            P.marker_rule(P.RULE_BOUNCEBACK, P.SHAPE_SPHERE_LE, p=(0.0, 0.0, 0.0, 0.3)),
            P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_BOX, p=(-0.6, -0.5, -0.5, 0.7, 0.5, 0.6), levels=[0]),
            P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_SPHERE_LT, p=(0.05, 0.0, 0.0, 0.5), levels=[0, 1]),
            P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(-1, -1, 0, -1, 0, -1), levels=[0]),
        )
        tri = tilted_box(-0.85, 0.35, 0.15, 0.75, -0.35, 0.05, 1.0, 0.5)
        return base, 3, 1, [(tri, (0.0, 0.0, 0.0))], r
    raise KeyError(name)


CASES = ["sphere_rules_3", "stl_2", "stl_walls_1", "ties_3"]


@pytest.fixture(scope="module")
def built():
    out = {}
    for name in CASES:
        base, levels, span, stls, r = _case(name)
        out[name] = (O.diad_build(base, levels, span, 1e-3, stls, r), (base, levels, span, stls, r))
    return out


def test_rule_layout_matches():
    assert P.RULE_DTYPE == O.RULE_DTYPE
    assert P.RULE_DTYPE.itemsize == 72
    assert P.DIAD_ARRAYS == O.DIAD_ARRAYS


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_dense_restatement(built, name):
    lv, (base, levels, span, stls, r) = built[name]
    dense = diad_dense.build(O, base, levels, span, 1e-3, stls, r)
    assert len(dense) == levels
    for lev, (a, b) in enumerate(zip(lv, dense)):
        print(f"  {name} level {lev}: n={a['n']} bounceback={int(a['bounceback'].sum())} "
              f"c2f={a['c2f_read'].size} f2c={a['f2c_read'].size} outlet={a['outlet'].size}")
        assert a["n"] == b["n"] > 0
        for key in ("nx", "ny", "nz"):
            assert a[key] == b[key]
        for key in ("res", "ox", "oy", "oz"):
            assert np.float32(a[key]) == np.float32(b[key]), key
        for key in O.DIAD_ARRAYS:
            assert np.array_equal(a[key], b[key]), (lev, key)


@pytest.mark.parametrize("name", CASES)
def test_builder_properties(built, name):
    lv, (base, levels, span, stls, r) = built[name]
    _check_properties(lv, levels)


@pytest.mark.parametrize("seed", range(12))
def test_random_cases_oracle_matches_dense_and_properties(seed):
    """Seeded random sphere walls, refinement and outlet rules over two or
    three levels (tests/random_diad.py): the oracle equals the dense
    restatement on every array and has the construction's properties."""
    base, levels, span, stls, r = random_diad.case(seed)
    lv = O.diad_build(base, levels, span, 1e-3, stls, r)
    dense = diad_dense.build(O, base, levels, span, 1e-3, stls, r)
    print(f"  random {seed}: levels {[d['n'] for d in lv]}")
    for lev, (a, b) in enumerate(zip(lv, dense)):
        assert a["n"] == b["n"] > 0
        for key in O.DIAD_ARRAYS:
            assert np.array_equal(a[key], b[key]), (lev, key)
    _check_properties(lv, levels)


def _check_properties(lv, levels):
    for lev, d in enumerate(lv):
        i, j, k = d["ijk"].astype(np.int64)
        key = (k * d["ny"] + j) * d["nx"] + i
        assert np.all(np.diff(key) > 0), "cells sorted by (k, j, i), unique"
        n = d["n"]
        # esotwist neighbours lie along their rays
        for t, (dx, dy, dz) in enumerate(diad_dense.ESO_DIRS):
            m = d["nbr"][t]
            assert np.all((m >= 0) & (m < n))
            # no displacement along the axes the direction does not move in
            for comp, c in ((i[m] - i, dx), (j[m] - j, dy), (k[m] - k, dz)):
                if c == 0:
                    assert np.all(comp == 0)
        if lev == levels - 1:
            # finest: kept = fluid + fluid neighbours, bounceback = not fluid
            fluid = d["bounceback"] == 0
            fl_set = set(key[fluid].tolist())
            for c in np.flatnonzero(~fluid):
                nb = [((k[c] + dz) * d["ny"] + j[c] + dy) * d["nx"] + i[c] + dx
                      for dx, dy, dz in diad_dense.DIRS26]
                assert any(x in fl_set for x in nb), "solid cell kept without a fluid neighbour"
            continue
        f = lv[lev + 1]
        fi, fj, fk = f["ijk"].astype(np.int64)
        fkey = (fk * f["ny"] + fj) * f["nx"] + fi
        fset = set(fkey.tolist())
        for arr in ("c2f_read", "f2c_write"):
            for c in d[arr]:
                for dx in (0, 1):
                    for dy in (0, 1):
                        for dz in (0, 1):
                            ck = ((2 * k[c] + dz) * f["ny"] + 2 * j[c] + dy) * f["nx"] + 2 * i[c] + dx
                            assert ck in fset, f"child of interface cell {c} missing"
        assert np.array_equal(fi[d["c2f_write"]], 2 * i[d["c2f_read"]])
        assert np.array_equal(fk[d["f2c_read"]], 2 * k[d["f2c_write"]])
        plain = set(key[d["interface"] == 0].tolist())
        parent = ((fk // 2) * d["ny"] + fj // 2) * d["nx"] + fi // 2
        assert not any(p in plain for p in parent.tolist()), "fine cell under a plain coarse cell"
        assert f["res"] == np.float32(d["res"]) * np.float32(0.5)
