"""Seeded random wall-refined DIAD geometries through the whole level
recursion (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:200-209):
level updates, the interface transfers and the records the updates write for
them, against the oracle's recursion over the same tables (tests/eso_levels.py).

The levels come from the device builder on the random builder cases of
tests/random_diad.py (bit-exact against the oracle's builder,
tests/test_gpu_diad_builder.py), with a physical viscosity and time step that
give level 0 a lattice viscosity of 0.01-0.03, and the DIAD sphere driver's
flags: given velocity on the box faces of level 0 and given density on its
x-max face, ghost box faces on the finer levels, the builder's bounceback
cells.  Started from uniform flow at 0.03 with a random Smagorinsky constant.

* FP32: two blocks of two coarse steps, every level within 1e-5 relative of the
  oracle after each block (the tolerance of the other DIAD level tests).
* FP16: after four coarse steps every interface record equals what its cell
  holds, and every level is within 2e-3 absolute of the FP32 oracle.
"""
import numpy as np
import pytest

from tests import eso_levels as E
from tests import random_diad

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5
SEEDS = range(8)


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b), 1e-3)))


def _levels(P, seed, storage):
    base, levels, span, stls, rules = random_diad.case(seed)
    rng = np.random.default_rng(12000 + seed)
    nu0 = float(rng.uniform(0.01, 0.03))
    cs = float(rng.choice([0.0, 0.1]))
    dt = 1e-3
    nu_phys = nu0 * (base["res"] / 1000.0) ** 2 / dt
    info = P.InfoStruct(res=base["res"], ox=base["ox"], oy=base["oy"], oz=base["oz"], nu=nu0, dtPhys=dt,
                        cellCountX=base["nx"], cellCountY=base["ny"], cellCountZ=base["nz"])
    lv = P.buildDIADGrids(info, levels, span, nu_phys, (), rules)
    grids, oracle_levels = [], []
    for l in range(levels):
        g = P.DIADGridStruct.from_levels(lv, l, storage)
        li, _ = lv.info(l)
        shape = (li.cellCountX, li.cellCountY, li.cellCountZ)
        d = lv.level(l)
        flags = cases.diad_sphere_flags(d["ijk"], d["bounceback"], shape, l)
        g.set_flags(flags)
        g.set_given_values(cases.channel_given_values(0.03))
        g.set_smagorinsky(cs)
        g.fill_equilibrium(ux=0.03)
        grids.append(g)
        dd = dict(d, nx=shape[0], ny=shape[1], nz=shape[2])
        o = E.OracleLevel(dd, flags, cases.channel_given_values(0.03), li.nu, cs)
        o.set(g.get_f())
        oracle_levels.append(o)
    print(f"[random diad levels] seed {seed} {storage}: cells per level {[g.n for g in grids]}, cs {cs}", flush=True)
    return lv, grids, oracle_levels


@pytest.mark.parametrize("seed", SEEDS)
def test_random_diad_recursion_matches_oracle(psl, seed):
    P = psl
    lv, grids, ol = _levels(P, seed, "f32")
    for block in range(2):
        P.updateGrid(grids, 2)
        for _ in range(2):
            E.update_grid(ol)
        for l, (g, o) in enumerate(zip(grids, ol)):
            assert g.esotwistFlipper == o.flip
            err = _rel_err(g.get_f(), o.get())
            print(f"  after {2 * (block + 1)} coarse steps, level {l}: max rel err {err:.3e}", flush=True)
            assert err < RTOL
    for g in grids:
        g.close()
    lv.close()


@pytest.mark.parametrize("seed", SEEDS)
def test_random_diad_recursion_fp16(psl, seed):
    P = psl
    lv, grids, ol = _levels(P, seed, "f16")
    P.updateGrid(grids, 4)
    for _ in range(4):
        E.update_grid(ol)
    for l, (g, o) in enumerate(zip(grids, ol)):
        assert g.verify_interface_records() == 0
        err = float(np.max(np.abs(g.get_f().astype(np.float64) - o.get())))
        print(f"  level {l}: max abs err {err:.3e}", flush=True)
        assert err < 2e-3
    for g in grids:
        g.close()
    lv.close()
