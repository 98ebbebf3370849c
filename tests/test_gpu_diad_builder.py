"""DIAD grid builder on the device (psl_diad_build) against the oracle.

Every level's cell table, bounceback and interface markers, esotwist
neighbours, interface maps and outlet list must equal the oracle's
restatement of buildDIADGrids (include/DIADFunctions.h:640-1046) exactly
(integer work: bit-exact).  The oracle is pinned on the CPU against a dense
restatement (tests/test_diad_builder.py).  Cases: sphere walls from marker
rules over three and four levels with refinement rules (one clears
refinement), an STL sphere with an STL origin offset over two levels, and a
single level with STL plus index-box walls.
"""
import numpy as np
import pytest

from tests import random_diad
from tests.test_diad_builder import CASES, _case, _rules_sphere

pytestmark = pytest.mark.gpu


def _gpu_build(P, base, levels, span, stls, rules, nu_phys=1e-3):
    info = P.InfoStruct(res=base["res"], ox=base["ox"], oy=base["oy"], oz=base["oz"], nu=base["nu"],
                        dtPhys=base["dt"], cellCountX=base["nx"], cellCountY=base["ny"], cellCountZ=base["nz"])
    meshes = []
    for tri, origin in stls:
        s = P.STLStruct.from_triangles(tri)
        s.set_origin(*origin)
        meshes.append(s)
    return P.buildDIADGrids(info, levels, span, nu_phys, meshes, rules)


def _compare(P, O, name, base, levels, span, stls, rules):
    ref = O.diad_build(base, levels, span, 1e-3, stls, rules)
    lv = _gpu_build(P, base, levels, span, stls, rules)
    print(f"[diad-build] {name}: {levels} levels, build {lv.build_ms:.1f} ms", flush=True)
    for lev in range(levels):
        d = lv.level(lev)
        r = ref[lev]
        info = d["info"]
        print(f"  level {lev}: n={d['n']} (oracle {r['n']}) bounceback={int(d['bounceback'].sum())} "
              f"c2f={d['c2f_read'].size} f2c={d['f2c_read'].size} outlet={d['outlet'].size}", flush=True)
        assert d["n"] == r["n"]
        assert (info.cellCountX, info.cellCountY, info.cellCountZ) == (r["nx"], r["ny"], r["nz"])
        for key in ("res", "ox", "oy", "oz", "nu"):
            got = np.float32(getattr(info, key))
            assert got == np.float32(r[key]), (lev, key, got, r[key])
        for key in P.DIAD_ARRAYS:
            assert np.array_equal(d[key], r[key]), (lev, key)
    lv.close()


@pytest.mark.parametrize("name", CASES)
def test_builder_matches_oracle(psl, oracle_mod, name):
    base, levels, span, stls, rules = _case(name)
    _compare(psl, oracle_mod, name, base, levels, span, stls, rules)


@pytest.mark.parametrize("seed", range(12))
def test_builder_random_cases_match_oracle(psl, oracle_mod, seed):
    """Seeded random sphere walls, refinement and outlet rules over two or
    three levels (tests/random_diad.py), pinned on the CPU against the dense
    restatement."""
    base, levels, span, stls, rules = random_diad.case(seed)
    _compare(psl, oracle_mod, f"random {seed}", base, levels, span, stls, rules)


def test_builder_four_levels(psl, oracle_mod):
    """Sphere walls over four levels (level 0 spans 48 x 32 x 32)."""
    P = psl
    base = dict(res=1.0, ox=-15.5, oy=-15.5, oz=-15.5, nu=0.01, dt=1.0, nx=48, ny=32, nz=32)
    rules = P.rules(
        P.marker_rule(P.RULE_BOUNCEBACK, P.SHAPE_SPHERE_LE, p=(0.0, 0.0, 0.0, 2.6)),
        P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_XCYLINDER, p=(0.0, 0.0, 0.0, 4.0, 10.0), levels=[0]),
        P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_SPHERE_LT, p=(0.0, 0.0, 0.0, 5.0), levels=[0, 1]),
        P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(-1, -1, 0, -1, 0, -1), levels=[0]),
    )
    _compare(P, oracle_mod, "sphere_rules_4", base, 4, 2, [], rules)


def test_builder_rejects_bad_input(psl):
    P = psl
    info = P.InfoStruct(res=1.0, cellCountX=8, cellCountY=8, cellCountZ=8)
    with pytest.raises(P.PslError):
        P.buildDIADGrids(info, 0, 1, 1e-3)
    big = P.InfoStruct(res=1.0, cellCountX=1 << 20, cellCountY=1, cellCountZ=1)
    with pytest.raises(P.PslError):
        P.buildDIADGrids(big, 3, 1, 1e-3)
    bad = _rules_sphere().copy()
    bad["shape"][0] = 17
    with pytest.raises(P.PslError):
        P.buildDIADGrids(info, 2, 1, 1e-3, (), bad)
